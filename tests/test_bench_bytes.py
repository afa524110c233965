"""The roofline's algorithmic bytes in bench.py: per-cell figures of the two-pass staggered
leapfrog with the reference's coefficient layout (every array read once per pass that uses
it, outputs written once, CPML memory variables read and written in the strips), and the
passes' own compulsory bytes beside them."""
import importlib.util
import os

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
bench = importlib.util.module_from_spec(spec)
spec.loader.exec_module(bench)


def test_2d_acoustic_step_bytes_match_the_worked_example():
    # 201^2 + 20-cell CPML = 241^2 cells: 58 081 x 48 B + 2 x 9 640 x 16 B per step
    vel, stress = bench.reference_bytes((241, 241), 20, 0)
    assert vel + stress == 58081 * 48 + 2 * 9640 * 16


def test_per_cell_figures_by_counting_arrays():
    # a grid without strips (npml = 0) leaves the per-cell figures: floats read + written
    n2, n3 = (100, 100), (20, 20, 20)
    N2, N3 = 100 * 100, 20 * 20 * 20
    assert bench.reference_bytes(n2, 0, 0) == (4 * (5 + 2) * N2, 4 * (4 + 1) * N2)
    assert bench.reference_bytes(n2, 0, 1) == (4 * (7 + 2) * N2, 4 * (8 + 3) * N2)
    assert bench.reference_bytes(n3, 0, 0) == (4 * (7 + 3) * N3, 4 * (5 + 1) * N3)
    assert bench.reference_bytes(n3, 0, 1) == (4 * (12 + 3) * N3, 4 * (14 + 6) * N3)


def test_strips_add_one_memory_variable_per_derivative_and_axis():
    # one strip cell per axis costs 8 B per memory variable (read + write); 2-D elastic has
    # two derivatives per axis in each pass, 3-D elastic three
    for shape, npml, phys, per_pass in (((300, 400), 10, 0, 8), ((300, 400), 10, 1, 16),
                                        ((40, 50, 60), 5, 0, 8), ((40, 50, 60), 5, 1, 24)):
        base = bench.reference_bytes(shape, 0, phys)
        full = bench.reference_bytes(shape, npml, phys)
        N = 1
        for n in shape:
            N *= n
        strips = sum(N * 2 * npml / n for n in shape)
        assert abs(full[0] - base[0] - per_pass * strips) < 1e-6 * full[0]
        assert abs(full[1] - base[1] - per_pass * strips) < 1e-6 * full[1]


def test_own_bytes_never_exceed_the_reference_layout():
    # the passes form the dt-scaled coefficients from rho, invlambda and invmu, so they read
    # at most as many arrays as the reference's layout asks for
    k2, r2 = bench.algorithmic_bytes(2081, 2081, 40, 1), bench.reference_bytes((2081, 2081), 40, 1)
    k3, r3 = bench.algorithmic_bytes_3d_elastic((360, 360, 360), 20), bench.reference_bytes((360, 360, 360), 20, 1)
    ka, ra = bench.algorithmic_bytes_3d((440, 440, 440), 20), bench.reference_bytes((440, 440, 440), 20, 0)
    for k, r in ((k2, r2), (k3, r3)):
        assert k[0] < r[0] and k[1] < r[1]
    assert ka == ra
