# GeoPhyInv.jl binding of libgpi (include/gpi.h): the body of mod_x_proc!
# (src/fdtd/propagate.jl:138-261) as one ccall per supersource.
#
# Include this file inside module GeoPhyInv after src/fdtd/fdtd.jl, e.g. at the end of
# src/fdtd/fdtd.jl:
#     include("GPI.jl")
# and let mod_x_proc! (propagate.jl:138) delegate:
#     mod_x_proc!(pac::P_common, pap::Vector{P_x_worker_x_pw{N,B}}, activepw, src_flags) where {N,B} =
#         GPI.mod_x_proc!(pac, pap, activepw, src_flags)
# The reference's host preparation is unchanged: pac.mod / pac.dmod (medium.jl:128-186),
# pac.pml (cpml.jl:104-155), ssprayw / rinterpolatew (fdtd/ageom.jl:29-54) and the
# wavelets (source.jl:3-58) are what the library receives.  Arrays are the Threads backend's
# Data.Array (column-major, z fastest): their pointers are passed without copies.
module GPI

using SparseArrays

const GP = parentmodule(@__MODULE__)   # GeoPhyInv: FdtdAcoustic, _fd_order, _fd_npml

const lib = get(ENV, "GPI_LIB", joinpath(@__DIR__, "libgpi.so"))

# ---- C structs, field for field as in include/gpi.h ------------------------------------
struct Grid
    ndim::Int32
    order::Int32
    physics::Int32
    npml::Int32
    n::NTuple{3,Int64}
    pml_mask::UInt32
    rigid_mask::UInt32
    free_mask::UInt32
    reserved::UInt32
    dt::Float32
    dsI::NTuple{3,Float32}
end

struct CArray
    id::Int32
    reserved::Int32
    count::Int64
    data::Ptr{Float32}
end

struct CPml
    field::Int32
    len::Int32
    a::Ptr{Float32}
    b::Ptr{Float32}
    kI::Ptr{Float32}
end

struct Points
    field::Int32
    npts::Int32
    nnz::Int32
    reserved::Int32
    idx::Ptr{Int64}
    w::Ptr{Float32}
end

struct Shot
    nt::Int64
    nsrc::Int32
    nrec::Int32
    src::Ptr{Points}
    wavelets::Ptr{Ptr{Float32}}
    rec::Ptr{Points}
    records::Ptr{Ptr{Float32}}
    nfinal::Int32
    nsnap::Int32
    final_fields::Ptr{Int32}
    final_out::Ptr{Ptr{Float32}}
    snap_field::Int32
    reserved::Int32
    snap_steps::Ptr{Int64}
    snap_out::Ptr{Ptr{Float32}}
    illum::Ptr{Float64}
end

struct AdjShot
    nt::Int64
    nsrc1::Int32
    nsrc2::Int32
    src1::Ptr{Points}
    wav1::Ptr{Ptr{Float32}}
    src2::Ptr{Points}
    wav2::Ptr{Ptr{Float32}}
end

# enum gpi_field (names of src/fields.jl:9-42)
const FIELD = Dict(:p => 0, :vx => 1, :vy => 2, :vz => 3, :tauxx => 4, :tauyy => 5, :tauzz => 6,
    :tauxy => 7, :tauxz => 8, :tauyz => 9, :dpdx => 10, :dpdy => 11, :dpdz => 12, :dvxdx => 13,
    :dvydy => 14, :dvzdz => 15, :dvxdy => 16, :dvxdz => 17, :dvydx => 18, :dvydz => 19,
    :dvzdx => 20, :dvzdy => 21, :dtauxxdx => 22, :dtauyydy => 23, :dtauzzdz => 24,
    :dtauxydx => 25, :dtauxydy => 26, :dtauxzdx => 27, :dtauxzdz => 28, :dtauyzdy => 29,
    :dtauyzdz => 30)

# enum gpi_coef: the dependent (dmod) and independent (mod) names of MediumParameters
# (src/fdtd/medium.jl:81-95)
const COEF = Dict(:dtinvavxirho => 0, :dtinvavyirho => 1, :dtinvavzirho => 2, :dtK => 3,
    :dtM => 4, :dtlambda => 5, :dtavmu => 6, :dtavxzimu => 7, :dtavxyimu => 8,
    :dtavyzimu => 9, :rho => 10, :invK => 11, :invlambda => 12, :invmu => 13)

# GPI_FACE(axis, side) = 1 << (2 axis + side), axis 0 = z, 1 = y, 2 = x
const FACE_BIT = Dict(:zmin => 0, :zmax => 1, :ymin => 2, :ymax => 3, :xmin => 4, :xmax => 5)

facemask(faces) = reduce(|, (UInt32(1) << FACE_BIT[f] for f in faces if haskey(FACE_BIT, f)); init = UInt32(0))

# ---- context ---------------------------------------------------------------------------
last_error(ctx) = unsafe_string(ccall((:gpi_last_error, lib), Cstring, (Ptr{Cvoid},), ctx))

function check(ctx, rc)
    rc == 0 || error("libgpi error $(rc): $(last_error(ctx))")
    return nothing
end

function context(device::Integer = 0; hbm_budget::Integer = 0)
    ctx = Ref{Ptr{Cvoid}}(C_NULL)
    rc = ccall((:gpi_create, lib), Cint, (Ptr{Ptr{Cvoid}}, Cint, UInt64), ctx, device, hbm_budget)
    rc == 0 || error("gpi_create failed with code $(rc)")
    return ctx[]
end

destroy(ctx) = ccall((:gpi_destroy, lib), Cvoid, (Ptr{Cvoid},), ctx)

# ---- model: pac.mod, pac.dmod and pac.pml resident on the device ----------------------
function grid(pac)
    n = length.(pac.exmedium.grid)
    nd = length(n)
    n3 = nd == 2 ? (Int64(n[1]), Int64(1), Int64(n[2])) : (Int64(n[1]), Int64(n[2]), Int64(n[3]))
    physics = pac.attrib_mod isa GP.FdtdAcoustic ? Int32(0) : Int32(1)
    dsI = nd == 2 ? (Float32(pac.fc[:dzI]), 0.0f0, Float32(pac.fc[:dxI])) :
          (Float32(pac.fc[:dzI]), Float32(pac.fc[:dyI]), Float32(pac.fc[:dxI]))
    # pac.rigid_faces already holds pml_faces ∪ rigid_faces (src/fdtd/fdtd.jl:215)
    return Grid(Int32(nd), Int32(GP._fd_order), physics, Int32(GP._fd_npml), n3,
                facemask(pac.pml_faces), facemask(pac.rigid_faces), facemask(pac.stressfree_faces),
                UInt32(0), Float32(pac.fc[:dt]), dsI)
end

# every dmod and mod array: acoustic dtinvav{x,(y),z}irho, dtK, invK, rho; elastic
# dtinvav{x,(y),z}irho, dtavmu or dtav{xz,xy,yz}imu, dtM, dtlambda, invlambda, invmu, rho
function coef_arrays(pac)
    arrs = CArray[]
    for k in names(pac.dmod)[1]
        a = pac.dmod[k]
        push!(arrs, CArray(Int32(COEF[k]), Int32(0), Int64(length(a)), pointer(a)))
    end
    for k in names(pac.mod)[1]
        a = pac.mod[k]
        push!(arrs, CArray(Int32(COEF[k]), Int32(0), Int64(length(a)), pointer(a)))
    end
    return arrs
end

function pml_arrays(pac)
    pmls = CPml[]
    for df in names(pac.pml)[1]
        P = pac.pml[df]
        push!(pmls, CPml(Int32(FIELD[df]), Int32(length(P[:a])), pointer(P[:a]), pointer(P[:b]), pointer(P[:kI])))
    end
    return pmls
end

function model(ctx, pac)
    g = grid(pac)
    arrs = coef_arrays(pac)
    pmls = pml_arrays(pac)
    h = Ref{UInt64}(0)
    GC.@preserve pac arrs pmls begin
        rc = ccall((:gpi_model_create, lib), Cint,
                   (Ptr{Cvoid}, Ref{Grid}, Ptr{CArray}, Int32, Ptr{CPml}, Int32, Ref{UInt64}),
                   ctx, g, arrs, Int32(length(arrs)), pmls, Int32(length(pmls)), h)
    end
    check(ctx, rc)
    return h[]
end

# update!(pa, medium) (src/fdtd/medium.jl:128-140): coefficient arrays overwritten in place
function model_update!(ctx, mdl, pac)
    arrs = coef_arrays(pac)
    GC.@preserve pac arrs begin
        rc = ccall((:gpi_model_update, lib), Cint, (Ptr{Cvoid}, UInt64, Ptr{CArray}, Int32),
                   ctx, mdl, arrs, Int32(length(arrs)))
    end
    check(ctx, rc)
end

model_free(ctx, mdl) = check(ctx, ccall((:gpi_model_free, lib), Cint, (Ptr{Cvoid}, UInt64), ctx, mdl))

# ---- supersource inputs ----------------------------------------------------------------
# One CSC matrix (n_field x npts) as gpi_points arrays: nnz zero-based rows per column in
# ascending order (the column order of src/proj_mat.jl:229-247).  A column with fewer
# stored entries is padded with weight-0 entries on its first row, which add exact zeros.
function point_arrays(W::SparseMatrixCSC, nnz::Integer)
    np = size(W, 2)
    idx = zeros(Int64, nnz, np)
    w = zeros(Float32, nnz, np)
    rv = rowvals(W)
    nz = nonzeros(W)
    for j in 1:np
        r = nzrange(W, j)
        length(r) <= nnz || error("point $(j) has more than $(nnz) weights")
        for (k, q) in enumerate(r)
            idx[k, j] = rv[q] - 1
            w[k, j] = Float32(nz[q])
        end
        for k in (length(r) + 1):nnz
            idx[k, j] = isempty(r) ? 0 : rv[first(r)] - 1
        end
    end
    return idx, w
end

# wavelets[it][field][is] (fdtd.jl:431-436) as an nt x ns matrix, time fastest
wavelet_matrix(pass, f, nt) = Float32[pass.wavelets[it][f][is] for it in 1:nt, is in 1:length(pass.wavelets[1][f])]

nnz_of(pac) = length(pac.exmedium.grid) == 2 ? 4 : 8

# Points for the fields `fs` of a NamedStack of CSC matrices; `keep` holds the arrays
# the pointers refer to until the call returns.
function points(mats, fs, nnz, keep)
    out = Points[]
    for f in fs
        idx, w = point_arrays(mats[f], nnz)
        push!(keep, idx, w)
        push!(out, Points(Int32(FIELD[f]), Int32(size(mats[f], 2)), Int32(nnz), Int32(0), pointer(idx), pointer(w)))
    end
    return out
end

# The gpi_shot of supersource `pass` (one P_x_worker_x_pw_x_ss): sources, receivers,
# snapshots and illumination; returns the struct and the record arrays it fills.
function shot(pac, pass, keep)
    nt = pac.ic[:nt]
    nnz = nnz_of(pac)
    sf = names(pass.ssprayw)[1]
    rf = names(pass.rinterpolatew)[1]
    src = points(pass.ssprayw, sf, nnz, keep)
    rcv = points(pass.rinterpolatew, rf, nnz, keep)
    wav = [wavelet_matrix(pass, f, nt) for f in sf]
    rec = [zeros(Float32, nt, size(pass.rinterpolatew[f], 2)) for f in rf]
    wp = Ptr{Float32}[pointer(a) for a in wav]
    rp = Ptr{Float32}[pointer(a) for a in rec]
    push!(keep, src, rcv, wav, rec, wp, rp)
    nsnap = 0
    snap_field = Int32(0)
    snap_steps = Ptr{Int64}(C_NULL)
    snap_out = Ptr{Ptr{Float32}}(C_NULL)
    if haskey(FIELD, pac.snaps_field)   # :Nothing when no snapshots are asked for
        steps = Int64[it - 1 for it in pac.itsnaps]   # zero-based step after which to copy
        outs = Ptr{Float32}[pointer(pass.snaps[k]) for k in 1:length(pac.itsnaps)]
        push!(keep, steps, outs)
        nsnap = length(steps)
        snap_field = Int32(FIELD[pac.snaps_field])
        snap_steps = pointer(steps)
        snap_out = pointer(outs)
    end
    illum = Ptr{Float64}(C_NULL)
    if pac.illum_flag   # compute_illum! (fdtd.jl:570-581) on the stress grid, FP64
        ill = zeros(Float64, length.(pac.exmedium.grid)...)
        push!(keep, ill)
        illum = pointer(ill)
    end
    sh = Shot(Int64(nt), Int32(length(src)), Int32(length(rcv)), pointer(src), pointer(wp), pointer(rcv),
              pointer(rp), Int32(0), Int32(nsnap), Ptr{Int32}(C_NULL), Ptr{Ptr{Float32}}(C_NULL), snap_field,
              Int32(0), snap_steps, snap_out, illum)
    return sh, rf, rec
end

# records as update_datamat! reads them: pass.records[field][it] is the vector over receivers
function copy_records!(pass, rf, rec)
    for (i, f) in enumerate(rf), it in 1:size(rec[i], 1)
        copyto!(pass.records[f][it], view(rec[i], it, :))
    end
end

# ---- modes -----------------------------------------------------------------------------
# :forward (propagate.jl:138-261 with npw = 1)
function forward!(ctx, mdl, pac, pass)
    keep = Any[]
    sh, rf, rec = shot(pac, pass, keep)
    GC.@preserve keep begin
        rc = ccall((:gpi_forward, lib), Cint, (Ptr{Cvoid}, UInt64, Ref{Shot}), ctx, mdl, sh)
    end
    check(ctx, rc)
    copy_records!(pass, rf, rec)
end

# :forward_save: the same run; the returned store keeps the boundary strips and final
# snapshots of this supersource on the device (boundary.jl:217-306)
function forward_save!(ctx, mdl, pac, pass)
    keep = Any[]
    sh, rf, rec = shot(pac, pass, keep)
    store = Ref{UInt64}(0)
    GC.@preserve keep begin
        rc = ccall((:gpi_forward_save, lib), Cint, (Ptr{Cvoid}, UInt64, Ref{Shot}, Ref{UInt64}),
                   ctx, mdl, sh, store)
    end
    check(ctx, rc)
    copy_records!(pass, rf, rec)
    return store[]
end

# checkpointed :forward_save (no reference counterpart): whole states every `every` steps
function forward_checkpoint!(ctx, mdl, pac, pass, every::Integer)
    keep = Any[]
    sh, rf, rec = shot(pac, pass, keep)
    store = Ref{UInt64}(0)
    GC.@preserve keep begin
        rc = ccall((:gpi_forward_checkpoint, lib), Cint, (Ptr{Cvoid}, UInt64, Ref{Shot}, Int32, Ref{UInt64}),
                   ctx, mdl, sh, Int32(every), store)
    end
    check(ctx, rc)
    copy_records!(pass, rf, rec)
    return store[]
end

store_free(ctx, store) = check(ctx, ccall((:gpi_store_free, lib), Cint, (Ptr{Cvoid}, UInt64), ctx, store))

# gpi_adjoint_shot of the :adjoint run with npw = 2: pw1 = the forward sources with their
# type -1 wavelets (source.jl:8-19), pw2 = the adjoint sources sprayed with pw1's receiver
# weights (source.jl:142-156)
function adjoint_shot(pac, pass1, pass2, keep; first_wavefield = true)
    nt = pac.ic[:nt]
    nnz = nnz_of(pac)
    sf = first_wavefield ? names(pass1.ssprayw)[1] : Symbol[]
    af = [f for f in names(pass2.wavelets[1])[1] if f in names(pass1.rinterpolatew)[1]]
    s1 = points(pass1.ssprayw, sf, nnz, keep)
    s2 = points(pass1.rinterpolatew, af, nnz, keep)
    w1 = [wavelet_matrix(pass1, f, nt) for f in sf]
    w2 = [wavelet_matrix(pass2, f, nt) for f in af]
    p1 = Ptr{Float32}[pointer(a) for a in w1]
    p2 = Ptr{Float32}[pointer(a) for a in w2]
    push!(keep, s1, s2, w1, w2, p1, p2)
    return AdjShot(Int64(nt), Int32(length(s1)), Int32(length(s2)),
                   isempty(s1) ? Ptr{Points}(C_NULL) : pointer(s1), isempty(p1) ? Ptr{Ptr{Float32}}(C_NULL) : pointer(p1),
                   pointer(s2), pointer(p2))
end

# :adjoint, 2-D acoustic (gradient.jl:17-64): fills pass1.gradients[:invK] and [:rho]
function adjoint!(ctx, mdl, store, pac, pass1, pass2)
    keep = Any[]
    sh = adjoint_shot(pac, pass1, pass2, keep)
    gK = pass1.gradients[:invK]
    gR = pass1.gradients[:rho]
    GC.@preserve keep gK gR begin
        rc = ccall((:gpi_adjoint, lib), Cint,
                   (Ptr{Cvoid}, UInt64, UInt64, Ref{AdjShot}, Float32, Ptr{Float32}, Ptr{Float32}),
                   ctx, mdl, store, sh, Float32(pac.fc[:dtI]), gK, gR)
    end
    check(ctx, rc)
end

# checkpointed acoustic gradient (2-D and 3-D): FP64 gradients w.r.t. invK and rho
function adjoint_checkpoint!(ctx, mdl, store, pac, pass1, pass2)
    keep = Any[]
    sh = adjoint_shot(pac, pass1, pass2, keep; first_wavefield = false)
    gK = zeros(Float64, length.(pac.exmedium.grid)...)
    gR = zeros(Float64, length.(pac.exmedium.grid)...)
    GC.@preserve keep gK gR begin
        rc = ccall((:gpi_adjoint_checkpoint, lib), Cint,
                   (Ptr{Cvoid}, UInt64, UInt64, Ref{AdjShot}, Float32, Ptr{Float64}, Ptr{Float64}),
                   ctx, mdl, store, sh, Float32(pac.fc[:dtI]), gK, gR)
    end
    check(ctx, rc)
    pass1.gradients[:invK] .= gK
    pass1.gradients[:rho] .= gR
end

# checkpointed elastic gradient (2-D P-SV and 3-D): FP64 gradients w.r.t. invlambda,
# invmu, rho; the receivers of the stored run record velocities
function adjoint_checkpoint_elastic!(ctx, mdl, store, pac, pass1, pass2)
    keep = Any[]
    sh = adjoint_shot(pac, pass1, pass2, keep; first_wavefield = false)
    gl = zeros(Float64, length.(pac.exmedium.grid)...)
    gm = zero(gl)
    gr = zero(gl)
    GC.@preserve keep gl gm gr begin
        rc = ccall((:gpi_adjoint_checkpoint_elastic, lib), Cint,
                   (Ptr{Cvoid}, UInt64, UInt64, Ref{AdjShot}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}),
                   ctx, mdl, store, sh, gl, gm, gr)
    end
    check(ctx, rc)
    pass1.gradients[:invlambda] .= gl
    pass1.gradients[:invmu] .= gm
    pass1.gradients[:rho] .= gr
end

# FdtdAcoustic{Born}: background and scattered wavefields, records of the scattered one.
# dcoefs: the linearised perturbations d(dtK) and d(dtinvav*irho) as (name => array) pairs
function born!(ctx, mdl, pac, pass, dcoefs)
    keep = Any[]
    sh, rf, rec = shot(pac, pass, keep)
    d = CArray[CArray(Int32(COEF[k]), Int32(0), Int64(length(a)), pointer(a)) for (k, a) in dcoefs]
    GC.@preserve keep dcoefs d begin
        rc = ccall((:gpi_born, lib), Cint, (Ptr{Cvoid}, UInt64, Ref{Shot}, Ptr{CArray}, Int32),
                   ctx, mdl, sh, d, Int32(length(d)))
    end
    check(ctx, rc)
    copy_records!(pass, rf, rec)
end

# ---- the replacement of mod_x_proc! ----------------------------------------------------
# One context per worker process and one resident model per P_common; stores of the
# :forward_save run are kept per supersource until the :adjoint run consumes them.
const CTX = Ref{Ptr{Cvoid}}(C_NULL)
const MODELS = IdDict{Any,UInt64}()
const STORES = IdDict{Any,Dict{Int,UInt64}}()

function ctx!()
    CTX[] == C_NULL && (CTX[] = context(0))
    return CTX[]
end

# the device copy of pac's medium, refreshed from pac.mod / pac.dmod on every call (the
# reference re-derives dmod in update!(pa, medium), medium.jl:128-140)
function model!(pac)
    ctx = ctx!()
    if haskey(MODELS, pac)
        model_update!(ctx, MODELS[pac], pac)
    else
        MODELS[pac] = model(ctx, pac)
    end
    return MODELS[pac]
end

function mod_x_proc!(pac, pap, activepw, src_flags)
    ctx = ctx!()
    mdl = model!(pac)
    mode = pac.attrib_mod.mode
    stores = get!(STORES, pac, Dict{Int,UInt64}())
    for issp in 1:length(pap[1].ss)
        pass1 = pap[1].ss[issp]
        if mode == :forward
            forward!(ctx, mdl, pac, pass1)
        elseif mode == :forward_save
            haskey(stores, issp) && store_free(ctx, stores[issp])
            stores[issp] = forward_save!(ctx, mdl, pac, pass1)
        elseif mode == :adjoint
            haskey(stores, issp) || error("run the :forward_save mode before the :adjoint mode")
            adjoint!(ctx, mdl, stores[issp], pac, pass1, pap[2].ss[issp])
            store_free(ctx, pop!(stores, issp))
        else
            error("mode $(mode) is not handled by libgpi")
        end
    end
    return nothing
end

# ---- pullback of lossvalue (no reference counterpart) ----------------------------------
# The adjoint half of gradient! (src/fdtd/func_grad.jl:21-45) for a reverse rule whose
# primal is lossvalue (func_grad.jl:1-9): that :forward_save run left one store per
# supersource in STORES, so only the :adjoint run is made here (mod_x_proc! consumes the
# stores), followed by the same chain rule to the log-parameter vector g.
function backprop!(g, m, loss, dobs, pa, mparams)
    pac = pa.c
    haskey(STORES, pac) && !isempty(STORES[pac]) ||
        error("run lossvalue (the :forward_save mode) before its pullback")
    mode_save = pac.attrib_mod.mode
    GP.gradient!(pac.srcwav[2], loss, dobs, pac.data[1])   # adjoint sources
    reverse!(pac.srcwav[2])
    GP.update!(pa, pac.srcwav, [-1, 1], verbose = false)
    pac.attrib_mod.mode = :adjoint
    GP.update!(pa)
    mchunks = GP.chunk(m, length(mparams))
    gchunks = GP.chunk(g, length(mparams))
    for (i, mname) in enumerate(mparams)
        r = pac.ref_mod[mname]
        gm = GP.view_inner(pac.gradients[mname], GP._fd_npextend, pac.pml_faces)
        map!((gm1, x1) -> gm1 * exp(x1) * r, gm, gm, mchunks[i])
        copyto!(gchunks[i], gm)
    end
    pac.attrib_mod.mode = mode_save
    return g
end

# ---- the FWI outer loop (src/fwi/fwi.jl) -----------------------------------------------

# The filter estimate of update!(paw, pac) (fwi.jl:42-57) in place of Conv.operator +
# IterativeSolvers.lsmr!: g and d are nt x ntr (the supersource's fields side by side), s
# has nt samples with Conv.Pconv's symmetric lags, gs = g * s.  Returns (iterations, istop).
function conv_lsmr!(s::Vector{Float64}, gs::Matrix{Float64}, g::Matrix{Float64}, d::Matrix{Float64};
                    maxiter::Integer = 100, atol::Real = 1e-6, btol::Real = 1e-6, conlim::Real = 1e8)
    ctx = ctx!()
    nt, ntr = size(g)
    (size(d) == size(g) == size(gs) && length(s) == nt) || throw(DimensionMismatch("conv_lsmr!"))
    iters, istop = Ref{Int32}(0), Ref{Int32}(0)
    check(ctx, ccall((:gpi_conv_lsmr, lib), Cint,
                     (Ptr{Cvoid}, Int64, Int64, Ptr{Float64}, Ptr{Float64}, Int32, Float64, Float64, Float64,
                      Ptr{Float64}, Ptr{Float64}, Ref{Int32}, Ref{Int32}),
                     ctx, nt, ntr, g, d, maxiter, atol, btol, conlim, s, gs, iters, istop))
    return iters[], istop[]
end

# apply_proj_matrix! (proj_mat.jl:208-226) on the device: y = x contracted with the dense
# Float32 weight matrix of each axis (nin[a] x nout[a], as get_proj_matrix builds them).
function interp_apply!(y::Array{Float64}, x::Array{Float64}, mats::Vector{Matrix{Float32}})
    ctx = ctx!()
    nd = length(mats)
    nin = Int64[size(W, 1) for W in mats]
    nout = Int64[size(W, 2) for W in mats]
    (size(x) == Tuple(nin) && size(y) == Tuple(nout)) || throw(DimensionMismatch("interp_apply!"))
    append!(nin, ones(Int64, 3 - nd))
    append!(nout, ones(Int64, 3 - nd))
    ptrs = Ptr{Float32}[pointer(W) for W in mats]
    append!(ptrs, fill(Ptr{Float32}(C_NULL), 3 - nd))
    GC.@preserve mats check(ctx, ccall((:gpi_interp_apply, lib), Cint,
                                       (Ptr{Cvoid}, Int32, Ptr{Int64}, Ptr{Int64}, Ptr{Ptr{Float32}}, Ptr{Float64},
                                        Ptr{Float64}), ctx, nd, nin, nout, ptrs, x, y))
    return y
end

# L-BFGS directions with the pair history in HBM, for an optimizer loop in place of
# Optim.LBFGS (fwi.jl:66-75): direction!(lb, d, m, g) returns the pairs used, d = -H g.
mutable struct LBFGS
    h::Ptr{Cvoid}
end

function LBFGS(n::Integer, memory::Integer = 10)
    ctx = ctx!()
    h = Ref{Ptr{Cvoid}}(C_NULL)
    check(ctx, ccall((:gpi_lbfgs_create, lib), Cint, (Ptr{Cvoid}, Int64, Int32, Ref{Ptr{Cvoid}}), ctx, n, memory, h))
    lb = LBFGS(h[])
    finalizer(x -> ccall((:gpi_lbfgs_destroy, lib), Cvoid, (Ptr{Cvoid},), x.h), lb)
    return lb
end

function direction!(lb::LBFGS, d::Vector{Float64}, m::Vector{Float64}, g::Vector{Float64})
    pairs = Ref{Int32}(0)
    check(ctx!(), ccall((:gpi_lbfgs_direction, lib), Cint,
                        (Ptr{Cvoid}, Ptr{Float64}, Ptr{Float64}, Ptr{Float64}, Ref{Int32}), lb.h, m, g, d, pairs))
    return pairs[]
end

reset!(lb::LBFGS) = check(ctx!(), ccall((:gpi_lbfgs_reset, lib), Cint, (Ptr{Cvoid},), lb.h))

end # module GPI
