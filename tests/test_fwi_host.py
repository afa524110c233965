"""Host side of the inversion-grid projection (src/proj_mat.jl:180-226, src/fwi/func_grad.jl).

Pinned by what interpolation must satisfy: partition of unity, exact reproduction of
linear functions inside the grid, clamping outside it, and the transposed map being the
adjoint of the forward map (the identity the inversion-grid gradient relies on).
"""
import numpy as np
import pytest

import geophyinv as G
from geophyinv.fwi import apply_proj, interp_matrix
from geophyinv.grid import Grid
from helpers import fwi_experiment, layered_acoustic


def test_interp_matrix_partition_and_linear_reproduction():
    grid = Grid(10.0, 20.0, 9)         # inversion grid 10..170
    pts = Grid(0.0, 5.0, 41)           # modelling grid 0..200
    M = interp_matrix(pts, grid)       # (9, 41)
    assert M.shape == (9, 41)
    assert np.allclose(M.sum(axis=0), 1.0, atol=1e-7)
    x = pts.values()
    xi = grid.values()
    f = lambda v: 3.0 - 0.25 * v
    inside = (x >= xi[0]) & (x <= xi[-1])
    got = (f(xi)[:, None] * M).sum(axis=0)
    assert np.allclose(got[inside], f(x[inside]), rtol=1e-6)
    # outside the grid the weights clamp to the edge node
    assert np.allclose(got[x < xi[0]], f(xi[0]), rtol=1e-6)
    assert np.allclose(got[x > xi[-1]], f(xi[-1]), rtol=1e-6)
    # a point on a node has a single weight 1
    node = interp_matrix([xi[3]], grid)[:, 0]
    assert node[3] == 1.0 and np.count_nonzero(node) == 1


def test_transposed_projection_is_the_adjoint():
    rng = np.random.Generator(np.random.PCG64(3))
    mm = [Grid(0.0, 5.0, 30), Grid(0.0, 5.0, 37)]
    mi = [Grid.range(20.0, 125.0, length=6), Grid.range(20.0, 160.0, length=8)]
    P = [interp_matrix(a, b) for a, b in zip(mm, mi)]
    PT = [np.ascontiguousarray(p.T) for p in P]
    m = rng.standard_normal((6, 8))
    y = rng.standard_normal((30, 37))
    lhs = np.sum(apply_proj(m, P).astype(np.float64) * y)
    rhs = np.sum(m * apply_proj(y, PT).astype(np.float64))
    assert lhs == pytest.approx(rhs, rel=1e-5)


def test_seisinvexpt_model_vector_round_trip():
    med = layered_acoustic(nz=40, nx=52, seed=2)
    pa = fwi_experiment(med, nss=1, nt=50)
    paw = G.SeisInvExpt(pa, [G.Recs(a.nr, pa.tgrid, pa.rfields) for a in pa.ageom], [8, 10])
    assert [g.n for g in paw.migrid] == [8, 10]
    assert paw.migrid[0].first == pytest.approx(pa.medium.grid[0].first + 4 * pa.medium.grid[0].step)
    m = paw.get_modelvector()
    assert m.shape == (paw.nm,)
    # log-parameters of a layered model are O(0.1) around the reference value
    assert np.all(np.isfinite(m)) and np.abs(m).max() < 2.0
    paw._to_full(m)
    assert paw.mfull.shape == pa.model_vector().shape
