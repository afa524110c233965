"""Host: the elastic chain rule (invlambda, invmu, rho) -> (vp, vs, rho) of
geophyinv.func_grad.chainrule_elastic against central differences of a smooth function of
the independent parameters, evaluated through the derived-parameter formulas of the medium
(src/media/media.jl:111-130)."""
import numpy as np

from geophyinv import func_grad
from geophyinv.media import ElasticMedium, derived
from geophyinv.grid import Grid


def test_chain_rule_matches_central_differences():
    rng = np.random.Generator(np.random.PCG64(21))
    n = (6, 5)
    vp = 2000.0 + 800.0 * rng.random(n)
    vs = vp / (1.6 + 0.4 * rng.random(n))
    rho = 1800.0 + 400.0 * rng.random(n)
    # F(invlambda, invmu, rho) = sum(a*invlambda + b*invmu^2 + c*log(rho)) with random weights
    a, b, c = (rng.standard_normal(n) for _ in range(3))

    def indep(vp, vs, rho):
        il = 1.0 / ((vp * vp - 2.0 * vs * vs) * rho)
        im = 1.0 / (vs * vs * rho)
        return il, im, rho

    def F(vp, vs, rho):
        il, im, r = indep(vp, vs, rho)
        return float(np.sum(a * il * 1e10 + b * (im * 1e10) ** 2 + c * np.log(r)))

    il, im, r = indep(vp, vs, rho)
    gl, gm, gr = a * 1e10, 2.0 * b * 1e20 * im, c / r
    gvp, gvs, grho = func_grad.chainrule_elastic(gl, gm, gr, vp, vs, rho)
    for name, g in (("vp", gvp), ("vs", gvs), ("rho", grho)):
        d = rng.standard_normal(n)
        base = {"vp": vp, "vs": vs, "rho": rho}
        h = 1e-4 * float(base[name].mean())

        def at(t):
            p = dict(base)
            p[name] = base[name] + t * d
            return F(p["vp"], p["vs"], p["rho"])

        fd = (at(h) - at(-h)) / (2 * h)
        ad = float(np.sum(g * d))
        assert abs(ad / fd - 1) < 1e-6, (name, ad, fd)


def test_independent_parameters_follow_the_medium():
    """The formulas inside the chain rule are the medium's derived parameters (FP32)."""
    rng = np.random.Generator(np.random.PCG64(3))
    n = (4, 7)
    vp = (2500.0 + 500.0 * rng.random(n)).astype(np.float32)
    vs = (vp / 1.8).astype(np.float32)
    rho = (2000.0 + 100.0 * rng.random(n)).astype(np.float32)
    med = ElasticMedium([Grid(0.0, 5.0, n[0]), Grid(0.0, 5.0, n[1])], vp, vs, rho)
    il = 1.0 / ((vp.astype(np.float64) ** 2 - 2.0 * vs.astype(np.float64) ** 2) * rho)
    im = 1.0 / (vs.astype(np.float64) ** 2 * rho)
    assert np.allclose(derived(med, "invlambda"), il, rtol=1e-6)
    assert np.allclose(derived(med, "invmu"), im, rtol=1e-6)
