# ChainRulesCore reverse rule of lossvalue (src/fdtd/func_grad.jl:1-9) over the libgpi
# binding (GPI.jl): the primal is the :forward_save run with its misfit, the pullback the
# adjoint half of gradient! (GPI.backprop!), scaled by the cotangent.  So Zygote or any
# ChainRules-aware AD differentiates the misfit without repeating the forward modelling.
#
# Include after GPI.jl inside module GeoPhyInv, where ChainRulesCore is available:
#     include("GPIChainRules.jl")
# The gradient it returns equals gradient!(g, m, loss, dobs, pa, mparams) scaled by the
# cotangent (the Python mirror, geophyinv/chainrules.py, is tested to do so bit for bit).
using ChainRulesCore

function ChainRulesCore.rrule(::typeof(lossvalue), m, loss, dobs::Records, pa::PFdtd,
                              mparams = Medium(pa.c.attrib_mod))
    J = lossvalue(m, loss, dobs, pa, mparams)   # :forward_save through GPI.mod_x_proc!
    function lossvalue_pullback(ΔJ)
        g = zero(m)
        GPI.backprop!(g, m, loss, dobs, pa, mparams)
        return NoTangent(), unthunk(ΔJ) .* g, NoTangent(), NoTangent(), NoTangent(), NoTangent()
    end
    return J, lossvalue_pullback
end
