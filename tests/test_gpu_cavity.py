"""The lid-driven cavity of the headline benchmark (D3Q19 TRT, magic 3/16,
tau = 0.6, bounce-back walls on every face, the lid a solid layer moving at
ux = 0.05), built by bench.make_cavity.

At a reduced size the FP64 GPU run is compared with the CPU oracle, in the
dense and the sparse layout. At the full 256^3 x 2000 steps the check is a
property that does not depend on an oracle run: the geometry and the lid
motion are symmetric under the mirror j -> n-1-j, so the flow is too (den,
ux, uz even, uy odd). FP64 rounding does not respect the mirror (the
population sums run in a fixed order), hence a tolerance of 1e-10 relative
to the largest speed.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle_cavity(n):
    import jyulb_oracle as jo
    from jyulb.defs import TRT, WALL
    import bench
    s = jo.OracleSolver(np.array([n, n, n], dtype=np.int32), TRT, (0.6 - 0.5) / 3.0, np.zeros(3),
                        np.full(6, WALL, dtype=np.int32))
    phase, den, vel = bench.cavity_inputs(n)
    s.get_phase()[...] = phase
    s.get_den()[...] = den
    s.get_vel()[...] = vel
    s.finalize_init()
    return s


@pytest.mark.parametrize("layout", ["rows", "sparse"])
def test_cavity_matches_oracle(layout):
    import bench
    n, calls = 48, (37, 24)
    print(f"cavity {n}^3 {layout} {calls}", flush=True)
    ref = _oracle_cavity(n)
    got, nfl = bench.make_cavity(n, "fp64", layout)
    assert nfl == n * n * (n - 1)
    fl = np.ones((n, n, n), dtype=bool)
    fl[:, :, n - 1] = False
    for steps in calls:
        ref.evolve(steps)
        got.evolve(steps)
        for a, b in ((ref.get_den(), got.get_den()), (ref.get_vel(), got.get_vel())):
            a, b = np.asarray(a)[fl], np.asarray(b)[fl]
            assert np.max(np.abs(a - b)) <= 1e-12 * np.max(np.abs(a))


def test_cavity_full_size_mirror_symmetry():
    import bench
    n = 256
    s, nfl = bench.make_cavity(n, "fp64")
    s.evolve(2000)
    den = np.asarray(s.get_den())[:, :, : n - 1]
    vel = np.asarray(s.get_vel())[:, :, : n - 1]
    umax = float(np.max(np.abs(vel)))
    assert 0.01 < umax <= 0.05
    mir = (slice(None), slice(None, None, -1), slice(None))
    assert np.max(np.abs(den - den[mir])) <= 1e-10
    for a, sign in ((0, 1.0), (1, -1.0), (2, 1.0)):
        err = np.max(np.abs(vel[..., a] - sign * vel[..., a][mir]))
        assert err <= 1e-10 * umax, (a, err)
    # the flow is not trivially symmetric: the lid drives a nonzero uy
    assert float(np.max(np.abs(vel[..., 1]))) > 1e-4 * umax
