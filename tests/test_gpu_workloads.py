"""The throughput workload (porous medium) and the precision workload (Zou-He
duct) at their own parameters.

Porous medium: seeded random spheres of radius 6..12 up to 20 % solid, D3Q19
TRT (magic 3/16), tau = 0.6, gravity gx = 1e-6, periodic on every face
(bench.porous_medium; full size 768 x 384 x 384, 113M cells, which the
reference's 32-bit indices cannot address, isothermal.pxd:136).
Duct: regularized collision, tau = 0.55, parabolic FIXED_VEL inlet of peak
0.04 at face 1X, FIXED_DEN outlet (rho = 1) at X1, WALL on both y and both z
faces (bench.duct; full size 512 x 128 x 128). Its Zou-He layers meet the
walls along the duct's edges, where a site has bounce-back links and a Zou-He
face (isothermal.pyx:1280-1291: Zou-He runs after the streaming and
bounce-back).

At reduced sizes the FP64 GPU runs are compared with the CPU oracle, in the
dense and the sparse layout (<= 1e-12 relative, bit-exact in practice). At
full size the checks are properties that need no oracle run: mass
conservation and layout invariance for the medium, the mirror and transpose
symmetries of the duct.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _oracle(lat, rlx, kvisc, g, bc, phase, den=None, vel=None):
    import jyulb_oracle as jo
    s = jo.OracleSolver(np.array(lat, dtype=np.int32), rlx, kvisc, np.asarray(g, dtype=np.float64),
                        np.asarray(bc, dtype=np.int32))
    s.get_phase()[...] = phase
    if den is not None:
        s.get_den()[...] = den
    if vel is not None:
        s.get_vel()[...] = vel
    s.finalize_init()
    return s


def _compare(ref, got, what):
    exact = True
    for name in ("get_den", "get_vel"):
        a = np.asarray(getattr(ref, name)())
        b = np.asarray(getattr(got, name)())
        scale = float(np.max(np.abs(a)))
        err = float(np.max(np.abs(a - b)))
        assert err <= 1e-12 * scale, f"{what} {name}: {err:.3e} (scale {scale:.3e})"
        exact &= err == 0.0
    return exact


@pytest.mark.parametrize("layout", ["rows", "sparse"])
def test_porous_medium_parameters_match_oracle(layout):
    """The medium's exact parameters at 96 x 48 x 48 (spheres of radius
    6..12, 20 % solid, seed 5), two evolve calls of odd and even length."""
    import bench
    from jyulb.defs import TRT, PERIODIC
    from jyulb.geometry import porous_spheres
    lat, calls = (96, 48, 48), (61, 40)
    solid = porous_spheres(lat, 6, 12, 0.20, 5)
    print(f"porous medium {lat} {layout} calls {calls}, solid {solid.mean():.3f}", flush=True)
    got, nfl = bench.porous_medium("fp64", lat=lat, layout=layout, solid=solid)
    ref = _oracle(lat, TRT, (0.6 - 0.5) / 3.0, [1e-6, 0, 0], [PERIODIC] * 6, np.where(solid, 0, 1))
    assert nfl == int((~solid).sum())
    exact = True
    for st in calls:
        ref.evolve(st)
        got.evolve(st)
        exact &= _compare(ref, got, f"porous {layout} after {st}")
    print(f"bit-exact: {exact}", flush=True)


def test_porous_medium_full_size_mass_and_layouts():
    """768 x 384 x 384, 500 FP64 steps: the total mass (FP64 reduction on the
    device) is conserved to 1e-12 relative, the dense and the sparse layout
    give the same fields bit for bit, and gravity drives a flow along +x."""
    import bench
    from jyulb.geometry import porous_spheres
    lat = (768, 384, 384)
    solid = porous_spheres(lat, 6, 12, 0.20, 5)
    fields = {}
    for layout in ("rows", "sparse"):
        print(f"porous medium {lat} fp64 {layout}: 500 steps", flush=True)
        s, nfl = bench.porous_medium("fp64", layout=layout, solid=solid)
        m0 = s.flow_info()["tot_den"]
        assert m0 == float(nfl)  # rho = 1 everywhere at the start
        s.evolve(500)
        info = s.flow_info()
        print(f"{layout}: mass {info['tot_den']!r} (start {m0!r}), ave ux {info['ave_ux']:.3e}, "
              f"max |u| {info['max_u2']:.3e}", flush=True)
        assert abs(info["tot_den"] - m0) <= 1e-12 * m0
        assert info["ave_ux"] > 0 and np.isfinite(info["max_u2"])
        fields[layout] = (np.array(s.get_den()), np.array(s.get_vel()))
        del s
    for a, b in zip(fields["rows"], fields["sparse"]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("layout", ["rows", "sparse"])
def test_duct_parameters_match_oracle(layout):
    """The duct's boundaries at 64 x 16 x 16: FIXED_VEL parabola / FIXED_DEN
    on x, WALL on y and z, regularized, tau = 0.55. The 4 x 2 Zou-He edge
    rows (a Zou-He face and bounce-back links on the same site) are checked
    on their own as well."""
    import bench
    from jyulb.defs import REGULARIZATION, WALL, FIXED_DEN, FIXED_VEL
    lat, calls = (64, 16, 16), (45, 30)
    print(f"duct {lat} {layout} calls {calls}", flush=True)
    got, nfl = bench.duct("fp64", lat=lat, layout=layout)
    ph = np.asarray(got.get_phase()).copy()
    ref = _oracle(lat, REGULARIZATION, (0.55 - 0.5) / 3.0, [0, 0, 0],
                  [FIXED_VEL, FIXED_DEN, WALL, WALL, WALL, WALL], ph, np.asarray(got.get_den()).copy(),
                  np.asarray(got.get_vel()).copy())
    exact = True
    for st in calls:
        ref.evolve(st)
        got.evolve(st)
        exact &= _compare(ref, got, f"duct {layout} after {st}")
        for i in (0, lat[0] - 1):
            for j in (0, lat[1] - 1):
                for k in (0, lat[2] - 1):
                    a = np.asarray(ref.get_vel())[i, j, k]
                    b = np.asarray(got.get_vel())[i, j, k]
                    assert np.array_equal(a, b), (i, j, k, a, b)
    # the inlet layer carries the prescribed profile (Zou-He), the walls slow the edges
    v = np.asarray(got.get_vel())
    assert v[0, lat[1] // 2, lat[2] // 2, 0] > v[0, 1, 1, 0] > 0
    print(f"bit-exact: {exact}", flush=True)


def test_duct_full_size_symmetries():
    """512 x 128 x 128, 5000 FP64 steps. Geometry, inlet profile and walls
    are symmetric under y -> -y, z -> -z and y <-> z (ny = nz), so is the
    flow: den and ux even, the mirrored component odd, uy and uz exchanged
    by the transpose. Rounding does not respect the mirrors (the population
    sums run in a fixed order): tolerance 1e-10 relative to max |u|."""
    import bench
    s, nfl = bench.duct("fp64")
    print("duct 512x128x128 fp64: 5000 steps", flush=True)
    s.evolve(5000)
    den = np.asarray(s.get_den())
    vel = np.asarray(s.get_vel())
    assert np.isfinite(vel).all()
    umax = float(np.max(np.abs(vel)))
    assert 0.03 < umax < 0.2
    tol = 1e-10 * umax
    my = (slice(None), slice(None, None, -1), slice(None))
    mz = (slice(None), slice(None), slice(None, None, -1))
    for mir, odd in ((my, 1), (mz, 2)):
        assert np.max(np.abs(den - den[mir])) <= 1e-10
        for a in range(3):
            sign = -1.0 if a == odd else 1.0
            err = float(np.max(np.abs(vel[..., a] - sign * vel[..., a][mir])))
            assert err <= tol, (mir, a, err)
    dt = np.transpose(den, (0, 2, 1))
    vt = np.transpose(vel, (0, 2, 1, 3))
    assert np.max(np.abs(den - dt)) <= 1e-10
    assert np.max(np.abs(vel[..., 0] - vt[..., 0])) <= tol
    assert np.max(np.abs(vel[..., 1] - vt[..., 2])) <= tol
    assert np.max(np.abs(vel[..., 2] - vt[..., 1])) <= tol
    # a developing flow, not a trivially symmetric one
    assert float(np.max(np.abs(vel[..., 1]))) > 1e-5 * umax
