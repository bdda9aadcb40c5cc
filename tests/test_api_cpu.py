"""Host-side behaviour of the jyulb drop-in API that needs no GPU.

Argument checks of the constructors (isothermal.pyx:196-206, 378-426) raise
the reference's exception types and messages before any device call; the
enum values match defs.pxd:17-103; the analytic profiles of flow_field
(flow_field.pyx:20-123) have their closed forms; and without a HIP device the
evolve path refuses to run (no CPU fallback).
"""
import numpy as np
import pytest

import jyulb
from jyulb import defs, flow_field
from jyulb.D3Q19 import D3Q19_H10
from jyulb.isothermal import Solver


def test_enum_values_match_reference():
    assert (defs.FACE_1X, defs.FACE_X1, defs.FACE_1Y, defs.FACE_Y1, defs.FACE_1Z, defs.FACE_Z1) == tuple(range(6))
    assert (defs.WALL, defs.PERIODIC, defs.FIXED_DEN, defs.FIXED_VEL, defs.SYMMETRIC) == (0, 1, 2, 3, 4)
    assert (defs.BGK, defs.TRT, defs.REGULARIZATION) == (0, 1, 2)
    assert (defs.SOLID, defs.FLUID) == (0, 1)
    assert (defs.X, defs.Y, defs.Z) == (0, 1, 2)


def test_lattice_vectors_and_weights():
    d = D3Q19_H10()
    c = np.array([d.LI, d.LJ, d.LK]).T
    assert c.shape == (19, 3)
    # opposite vectors: opp(l) = 18 - l, rest at 9
    np.testing.assert_array_equal(c + c[::-1], 0)
    assert tuple(c[9]) == (0, 0, 0)
    assert len({tuple(v) for v in c}) == 19
    w = np.array([d.W0 if l == 9 else d.W1 if abs(c[l]).sum() == 1 else d.W2 for l in range(19)])
    assert abs(w.sum() - 1.0) < 1e-15
    np.testing.assert_allclose(w @ (c[:, 0] ** 2), 1.0 / d.AS2, rtol=0, atol=1e-15)


@pytest.mark.parametrize("args, msg", [
    (([0, 4, 4], 0, 0.1), "Invalid lattice size input."),
    (([4, 4, 4], 5, 0.1), "Invalid relaxation scheme input."),
    (([4, 4, 4], 0, 0.0), "Invalid kinematic viscosity input."),
    (([4, 4, 4], 0, 0.1, None, [7, 0, 0, 0, 0, 0]), "Invalid boundary condition input."),
    (([4, 4, 4], 0, 0.1, None, [1, 0, 0, 0, 0, 0]), "Invalid periodic bc setup (in x-dir.)."),
    (([4, 4, 4], 0, 0.1, None, [0, 0, 0, 1, 0, 0]), "Invalid periodic bc setup (in y-dir.)."),
    (([4, 4, 4], 0, 0.1, None, [0, 0, 0, 0, 1, 4]), "Invalid periodic bc setup (in z-dir.)."),
    (([4, 4, 4], 0, 0.1, None, [3, 2, 2, 0, 0, 0]), "Invalid fixed variable bc setup."),
    (([4, 4, 4], 0, 0.1, None, [0, 0, 0, 3, 2, 0]), "Invalid fixed variable bc setup."),
    (([4, 4, 4], 0, 0.1, None, [0, 2, 0, 0, 0, 3]), "Invalid fixed variable bc setup."),
])
def test_constructor_argument_checks(args, msg):
    with pytest.raises(ValueError) as e:
        Solver(np.array(args[0], dtype=np.int32), *args[1:])
    assert str(e.value) == msg


def test_with_iflow_shape_check():
    phase = np.ones((4, 3, 2), dtype=np.uint8)
    with pytest.raises(ValueError, match="y-dimension"):
        Solver.with_iflow(phase, np.ones((4, 2, 2)), np.zeros((4, 3, 2, 3)))
    with pytest.raises(TypeError):
        Solver.with_iflow(phase, None, np.zeros((4, 3, 2, 3)))


def test_no_device_fails_loudly():
    """Valid arguments on a host without a GPU: the solver refuses to run
    rather than computing anything on the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match="no HIP device"):
        Solver(np.array([4, 4, 4], dtype=np.int32), defs.TRT, 0.1)


def test_variant_names_checked_before_the_device():
    """Precision and layout names are validated on the host: every listed name
    passes to the device step, an unknown one raises ValueError."""
    import torch
    from jyulb import _native
    lat = np.array([4, 4, 4], dtype=np.int32)
    for kw in ({"precision": "fp24"}, {"layout": "dense"}, {"layout": 7}):
        with pytest.raises(ValueError, match="unknown"):
            Solver(lat, defs.TRT, 0.1, **kw)
    assert {k for k in _native.LAYOUTS if isinstance(k, str)} == {"rows", "soa", "soa_padded", "sparse", "auto"}
    assert [_native.LAYOUTS[k] for k in ("rows", "soa", "soa_padded", "sparse", "auto")] == [0, 1, 2, 3, 4]
    if not torch.cuda.is_available():
        for kw in ({"layout": "sparse"}, {"precision": "fp32_storage", "layout": 3}):
            with pytest.raises(RuntimeError, match="no HIP device"):
                Solver(lat, defs.TRT, 0.1, **kw)


def test_flow_field_closed_forms():
    H, nu, pg = 10, 0.1, -2e-5
    dist, ana = np.zeros(H), np.zeros(H)
    flow_field.poise_plate_steady_state_pgrad(H, nu, pg, dist, ana)
    y = np.arange(H) + 0.5
    np.testing.assert_allclose(dist, y - H / 2, rtol=0, atol=1e-15)
    np.testing.assert_allclose(ana, -pg * y * (H - y) / (2 * nu), rtol=1e-14, atol=0)
    um = np.zeros(H)
    flow_field.poise_plate_steady_state_umax(H, 0.03, dist, um)
    np.testing.assert_allclose(um, 0.03 * (1 - ((y - H / 2) / (H / 2)) ** 2), rtol=1e-14, atol=0)
    cv = np.zeros((H, 2))
    flow_field.couette_plate_steady_state(H, 0.01, -0.02, 0.05, 0.03, dist, cv)
    np.testing.assert_allclose(cv[:, 0], 0.01 + y * 0.04 / H, rtol=1e-14, atol=0)
    np.testing.assert_allclose(cv[:, 1], -0.02 + y * 0.05 / H, rtol=1e-14, atol=0)


def test_package_exports():
    assert jyulb.__version__ == "0.2.0"
    from jyulb import adjoint  # noqa: F401
    assert callable(adjoint.flux_objective)
    assert callable(adjoint.flux_gradient)


def test_checkpoint_interval_choice(monkeypatch):
    """interval='auto' of the gradient entry points: the fewest checkpoint
    segments whose memory fits the budget. Lattice 10 x 4 x 70 has
    10 * 4 * 128 = 5120 padded sites: a population array is 778,240 bytes in
    FP64, one step of (rho, u*) record 163,840, and the reverse pass holds a
    cotangent array and 3 acceleration cotangents per site (901,120 bytes),
    and a second cotangent array where the adjoint steps take their pull
    form (from 1 GiB of cotangents, or as LB_ADJ_PULL says)."""
    from types import SimpleNamespace
    from jyulb.adjoint import checkpoint_interval
    from jyulb._native import LB_FP64, LB_FP32_SHIFTED
    monkeypatch.delenv("LB_ADJ_PULL", raising=False)
    s64 = SimpleNamespace(engine=SimpleNamespace(lat=(10, 4, 70), precision=LB_FP64))
    s32 = SimpleNamespace(engine=SimpleNamespace(lat=(10, 4, 70), precision=LB_FP32_SHIFTED))
    fb, rb, extra = 778_240, 163_840, 901_120
    one = 2 * fb + 100 * rb + extra  # one segment: 18,841,600 bytes
    assert checkpoint_interval(s64, 100, one) == 100
    # a byte less: two segments of 50 (3 arrays + 50 steps of record), not the
    # largest interval that fits (which would recompute more)
    assert checkpoint_interval(s64, 100, one - 1) == 50
    assert checkpoint_interval(s64, 100, 3 * fb + 50 * rb + extra) == 50
    assert checkpoint_interval(s64, 100, 3 * fb + 50 * rb + extra - 1) == 34
    # FP32 halves every figure
    assert checkpoint_interval(s32, 100, one // 2) == 100
    assert checkpoint_interval(s32, 100, one // 2 - 1) == 50
    # the dissipation objective also holds 23 seed values per site
    assert checkpoint_interval(s64, 100, one, objective="dissipation") == 50
    assert checkpoint_interval(s64, 100, one + 23 * 5120 * 8, objective="dissipation") == 100
    # the least footprint is 5 segments of 20 (6 arrays + 20 steps): 8,847,360 bytes
    assert checkpoint_interval(s64, 100, 6 * fb + 20 * rb + extra) == 20
    with pytest.raises(MemoryError):
        checkpoint_interval(s64, 100, 6 * fb + 20 * rb + extra - 1)
    with pytest.raises(ValueError):
        checkpoint_interval(s64, 0, 1 << 40)
    # pull-form adjoint steps hold a second cotangent array: forced here ...
    monkeypatch.setenv("LB_ADJ_PULL", "1")
    assert checkpoint_interval(s64, 100, one + fb) == 100
    assert checkpoint_interval(s64, 100, one + fb - 1) == 50
    monkeypatch.delenv("LB_ADJ_PULL")
    # ... and by default from 1 GiB of cotangents: 512^3 in FP64, 134,217,728
    # sites, 20,401,094,656 bytes per array
    big = SimpleNamespace(engine=SimpleNamespace(lat=(512, 512, 512), precision=LB_FP64))
    bfb, brb = (19 * 8) << 27, (4 * 8) << 27
    bextra = 2 * bfb + ((3 * 8) << 27)
    assert checkpoint_interval(big, 1, 2 * bfb + brb + bextra) == 1
    with pytest.raises(MemoryError):
        checkpoint_interval(big, 1, 2 * bfb + brb + bextra - 1)
    monkeypatch.setenv("LB_ADJ_PULL", "0")
    assert checkpoint_interval(big, 1, 2 * bfb + brb + bextra - bfb) == 1


@pytest.mark.parametrize("lat", [(1024, 1024, 2048), (1024, 1024, 2049), (4096, 1024, 1024)])
def test_lattice_limit_is_checked_before_the_device(lat):
    """Sites are indexed with 32-bit integers and counted with ints: lb_create
    refuses nx * ny * round_up(nz, 64) >= 2^31 with one message, before it
    touches a device (so this runs on the CPU). 1024 x 1024 x 2048 is exactly
    2^31 padded cells; 2049 pads to 2112."""
    import ctypes
    from jyulb import _native
    L = _native.lib()
    cfg = _native.LbConfig()
    for a in range(3):
        cfg.lat[a] = lat[a]
    cfg.rlx_scheme, cfg.kvisc, cfg.precision, cfg.layout = 0, 0.1, 2, 0
    for f in range(6):
        cfg.bc[f] = 1
    h = ctypes.c_void_p()
    rc = L.lb_create(ctypes.byref(cfg), ctypes.byref(h))
    msg = L.lb_last_error().decode()
    padded = lat[0] * lat[1] * ((lat[2] + 63) // 64 * 64)
    assert padded >= 2 ** 31
    assert rc == _native.LB_EINVAL
    assert msg == (f"lattice of {padded} cells (z padded to a multiple of 64) exceeds the engine's "
                   "limit of 2^31 padded cells")


def test_lattice_below_the_limit_passes_the_check():
    """One row fewer than 2^31 padded cells passes the size check: without a
    device lb_create then fails for want of a GPU, not for the size."""
    import ctypes
    import torch
    from jyulb import _native
    if torch.cuda.is_available():
        pytest.skip("a GPU is present: lb_create would allocate 2^31 sites")
    L = _native.lib()
    cfg = _native.LbConfig()
    for a, n in enumerate((1023, 1024, 2048)):
        cfg.lat[a] = n
    cfg.rlx_scheme, cfg.kvisc, cfg.precision, cfg.layout = 0, 0.1, 2, 0
    for f in range(6):
        cfg.bc[f] = 1
    h = ctypes.c_void_p()
    assert L.lb_create(ctypes.byref(cfg), ctypes.byref(h)) == _native.LB_EHIP
    assert "no HIP device" in L.lb_last_error().decode()


def test_porous_spheres_refuses_an_impossible_solid_fraction():
    """Open boundary layers that leave too few cells for the requested solid
    fraction raise instead of looping forever."""
    from jyulb.geometry import porous_spheres
    with pytest.raises(ValueError, match="open boundary layers"):
        porous_spheres((12, 10, 9), 4, 8, 0.10, 4, open_faces=(0, 1), open_depth=8)
    s = porous_spheres((20, 12, 10), 2, 4, 0.20, 5)
    assert 0.20 <= s.mean() < 0.3
