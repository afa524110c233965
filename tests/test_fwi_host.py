"""Host side of the inversion-grid projection (src/proj_mat.jl:180-226, src/fwi/func_grad.jl).

Pinned by what interpolation must satisfy: partition of unity, exact reproduction of
linear functions inside the grid, clamping outside it, and the transposed map being the
adjoint of the forward map (the identity the inversion-grid gradient relies on).
"""
import numpy as np
import pytest

import geophyinv as G
from geophyinv.fwi import interp_matrix, wolfe_search
from geophyinv.grid import Grid
from helpers import apply_proj_host, lbfgs_two_loop


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
    lhs = np.sum(apply_proj_host(m, P).astype(np.float64) * y)
    rhs = np.sum(m * apply_proj_host(y, PT).astype(np.float64))
    assert lhs == pytest.approx(rhs, rel=1e-5)


def test_two_loop_equals_the_dense_bfgs_inverse():
    """The two-loop recursion the device L-BFGS is checked against equals -H g with H the
    BFGS inverse Hessian built densely from H0 = gamma I through the same pairs
    (H <- (I - rho s y') H (I - rho y s') + rho s s'), and with no pairs it is -g."""
    rng = np.random.Generator(np.random.PCG64(7))
    n = 12
    A = rng.standard_normal((n, n))
    A = A @ A.T + n * np.eye(n)                      # SPD: every pair has s'y > 0
    pairs = []
    for _ in range(5):
        s = rng.standard_normal(n)
        pairs.append((s, A @ s))
    g = rng.standard_normal(n)
    s_n, y_n = pairs[-1]
    H = (s_n @ y_n) / (y_n @ y_n) * np.eye(n)
    for s, y in pairs:
        rho = 1.0 / (s @ y)
        V = np.eye(n) - rho * np.outer(y, s)
        H = V.T @ H @ V + rho * np.outer(s, s)
    assert np.allclose(lbfgs_two_loop(pairs, g), -H @ g, rtol=1e-11, atol=1e-12)
    assert np.array_equal(lbfgs_two_loop([], g), -g)
    # a mistake in either loop's order shows: reversing the pairs changes the direction
    assert not np.allclose(lbfgs_two_loop(pairs[::-1], g), -H @ g, rtol=1e-6)


def test_wolfe_search_satisfies_the_strong_wolfe_conditions():
    """The host line search of the device L-BFGS on the Rosenbrock function from several
    points and initial steps: the step it returns meets sufficient decrease and the strong
    curvature condition."""
    def fg(x):
        f = 100.0 * (x[1] - x[0] ** 2) ** 2 + (1 - x[0]) ** 2
        g = np.array([-400.0 * x[0] * (x[1] - x[0] ** 2) - 2 * (1 - x[0]), 200.0 * (x[1] - x[0] ** 2)])
        return f, g
    for x0, a0 in (((-1.2, 1.0), 1.0), ((0.5, -0.3), 1e-3), ((2.0, 2.0), 10.0)):
        x = np.array(x0)
        f0, g0 = fg(x)
        d = -g0
        a, f, g, nfev = wolfe_search(fg, x, f0, g0, d, a0)
        assert a is not None and nfev <= 20
        assert f <= f0 + 1e-4 * a * (g0 @ d)
        assert abs(g @ d) <= 0.9 * abs(g0 @ d)
        assert f == pytest.approx(fg(x + a * d)[0])
