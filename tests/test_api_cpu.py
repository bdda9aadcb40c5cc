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
    assert {k for k in _native.LAYOUTS if isinstance(k, str)} == {"rows", "soa", "soa_padded", "sparse"}
    assert [_native.LAYOUTS[k] for k in ("rows", "soa", "soa_padded", "sparse")] == [0, 1, 2, 3]
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
    from jyulb import adjoint, bcs  # noqa: F401
    assert callable(adjoint.flux_objective)
    assert callable(adjoint.flux_gradient)


def test_checkpoint_interval_choice():
    """interval='auto' of the gradient entry points: the largest interval whose
    checkpoints and (rho, u*) record fit the budget. Lattice 10 x 4 x 70 has
    10 * 4 * 128 = 5120 padded sites: a population array is 778,240 bytes in
    FP64 and one step of record 163,840."""
    from types import SimpleNamespace
    from jyulb.adjoint import checkpoint_interval
    from jyulb._native import LB_FP64, LB_FP32_SHIFTED
    s64 = SimpleNamespace(engine=SimpleNamespace(lat=(10, 4, 70), precision=LB_FP64))
    s32 = SimpleNamespace(engine=SimpleNamespace(lat=(10, 4, 70), precision=LB_FP32_SHIFTED))
    # one segment: 2 arrays + 100 steps of record = 17,940,480 bytes
    assert checkpoint_interval(s64, 100, 17_940_480) == 100
    # a byte less: 2 segments need 3 arrays (2,334,720) + i steps, i <= 95
    assert checkpoint_interval(s64, 100, 17_940_479) == 95
    # FP32 halves every figure
    assert checkpoint_interval(s32, 100, 17_940_480 // 2) == 100
    assert checkpoint_interval(s32, 100, 17_940_480 // 2 - 1) == 95
    # too small for any interval (the least footprint is 20 segments: 6 arrays + 20 steps)
    with pytest.raises(MemoryError):
        checkpoint_interval(s64, 100, 6 * 778_240 + 17 * 163_840)
    with pytest.raises(ValueError):
        checkpoint_interval(s64, 0, 1 << 40)
