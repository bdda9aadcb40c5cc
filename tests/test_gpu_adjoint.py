"""Adjoint AA step vs central finite differences of the forward model.

Objective: mean outlet x-flux J = <rho u*_x> over the fluid sites of face X1
after `steps` steps. Gradients with respect to the per-site acceleration field
(sampled sites, all components) and to the Zou-He inlet velocity data
(through the amplitude A of the inlet profile, dJ/dA = sum dJ/dv * profile).
Finite differences: FP64 forward on the GPU from the same initial state,
central, h = 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from jyulb.defs import BGK, TRT, REGULARIZATION, WALL, PERIODIC, FIXED_DEN, FIXED_VEL, SYMMETRIC
from jyulb.defs import FACE_1X, FACE_X1, X


def make_channel(rlx, precision, lat=(14, 9, 7), steps_bc=(WALL, SYMMETRIC), seed=3, solid=0.08, layout="rows"):
    from jyulb.isothermal import Solver
    nx, ny, nz = lat
    bc = np.array([FIXED_VEL, FIXED_DEN, WALL, WALL, steps_bc[0], steps_bc[1]], dtype=np.int32)
    if steps_bc[0] == PERIODIC:
        bc[4] = bc[5] = PERIODIC
    s = Solver(np.array(lat, dtype=np.int32), rlx, 0.2 / 3.0, np.array([1e-5, 0.0, 0.0]), bc,
               precision=precision, layout=layout)
    rng = np.random.default_rng(seed)
    ph = s.get_phase()
    ph[...] = 1
    core = rng.random(lat) < solid
    core[0, :, :] = False
    core[-1, :, :] = False
    ph[core] = 0
    prof = np.zeros((ny + 2, nz + 2))
    eta_y = (np.arange(ny) + 0.5) / ny - 0.5
    eta_z = (np.arange(nz) + 0.5) / nz - 0.5
    prof[1:-1, 1:-1] = (1 - (2 * eta_y[:, None]) ** 2) * (1 - (2 * eta_z[None, :]) ** 2)
    amp = 0.02
    s.get_vel()[0, :, :, 0] = amp * prof[1:-1, 1:-1]
    s.get_acc()[...] = 2e-6 * (rng.random(lat + (3,)) - 0.5)
    s.finalize_init()
    return s, prof, amp


def _inputs(s):
    s._set_bc_data()
    return (np.ascontiguousarray(s.get_acc()).copy(), s.bc_data_den_face_1X.copy(),
            s.bc_data_vel_face_1X.copy(), np.array(s.gravity))


OBJ_DISS = 6  # LB_OBJ_DISSIPATION


def _J(eng, acc, fden, fvel, g, steps, obj=FACE_X1):
    eng.restore()
    eng.set_gravity(g)
    eng.upload_acc(acc)
    eng.set_face_data(FACE_1X, fden, fvel)
    if obj == OBJ_DISS:  # evaluated on the final step's collision input
        return eng.adj_forward(steps, steps, OBJ_DISS, 0)
    eng.evolve(steps)
    return eng.objective(FACE_X1, X)


def _adjoint(s, acc, fden, fvel, g, steps, interval, obj=FACE_X1):
    eng = s.engine
    eng.restore()
    eng.set_gravity(g)
    eng.upload_acc(acc)
    eng.set_face_data(FACE_1X, fden, fvel)
    J = eng.adj_forward(steps, interval, obj, X)
    ga, gf = eng.adj_backward(FACE_1X)
    return J, ga, gf


def _fd_checks(s_adj, s_fd, prof, amp, steps, interval, ncells, seed, obj=FACE_X1):
    acc, fden, fvel, g = _inputs(s_fd)
    s_fd.engine.snapshot()
    s_adj.engine.snapshot()
    J, ga, gf = _adjoint(s_adj, acc, fden, fvel, g, steps, interval, obj)
    J0 = _J(s_fd.engine, acc, fden, fvel, g, steps, obj)
    h = 1e-6
    rng = np.random.default_rng(seed)
    fl = np.argwhere(s_fd.get_phase() != 0)
    pick = fl[rng.choice(len(fl), size=ncells, replace=False)]
    adj, fd = [], []
    for (i, j, k) in pick:
        a = int(rng.integers(0, 3))
        ap = acc.copy()
        ap[i, j, k, a] += h
        am = acc.copy()
        am[i, j, k, a] -= h
        fd.append((_J(s_fd.engine, ap, fden, fvel, g, steps, obj) - _J(s_fd.engine, am, fden, fvel, g, steps, obj))
                  / (2 * h))
        adj.append(ga[i, j, k, a])
    # inlet amplitude: v = A * profile (x component)
    vp = fvel.copy()
    vp[:, :, 0] = (amp + h) * prof
    vm = fvel.copy()
    vm[:, :, 0] = (amp - h) * prof
    fdA = (_J(s_fd.engine, acc, fden, vp, g, steps, obj) - _J(s_fd.engine, acc, fden, vm, g, steps, obj)) / (2 * h)
    adjA = float(np.sum(gf[:, :, 0] * prof))
    return J, J0, np.array(adj), np.array(fd), adjA, fdA


@pytest.mark.parametrize("rlx", [BGK, TRT, REGULARIZATION], ids=["bgk", "trt", "reg"])
def test_adjoint_fp64_matches_finite_differences(rlx):
    print(f"adjoint fp64 rlx={rlx}", flush=True)
    s, prof, amp = make_channel(rlx, "fp64")
    J, J0, adj, fd, adjA, fdA = _fd_checks(s, s, prof, amp, steps=60, interval=16, ncells=16, seed=7)
    assert abs(J - J0) <= 1e-13 * abs(J0)
    scale = np.max(np.abs(fd))
    assert scale > 0
    err = np.max(np.abs(adj - fd)) / scale
    assert err < 1e-6, (err, adj, fd)
    assert abs(adjA - fdA) <= 1e-6 * abs(fdA), (adjA, fdA)


@pytest.mark.parametrize("rlx", [BGK, TRT, REGULARIZATION], ids=["bgk", "trt", "reg"])
def test_adjoint_fp32_within_1e4_of_fp64_fd(rlx):
    """FP32 forward/adjoint against FP64 central differences: the north-star
    bar of 1e-4 relative (relative to the largest sampled gradient)."""
    print(f"adjoint fp32 rlx={rlx}", flush=True)
    s32, prof, amp = make_channel(rlx, "fp32")
    s64, _, _ = make_channel(rlx, "fp64")
    J, J0, adj, fd, adjA, fdA = _fd_checks(s32, s64, prof, amp, steps=60, interval=16, ncells=16, seed=11)
    scale = np.max(np.abs(fd))
    err = np.max(np.abs(adj - fd)) / scale
    assert err < 1e-4, (err, adj, fd)
    assert abs(adjA - fdA) <= 1e-4 * abs(fdA), (adjA, fdA)


def test_adjoint_periodic_and_checkpoint_invariance():
    """Gradients do not depend on the checkpoint interval; periodic faces
    (halo-crossing sites in the special-site kernel) are covered."""
    s, prof, amp = make_channel(TRT, "fp64", steps_bc=(PERIODIC, PERIODIC), seed=5)
    acc, fden, fvel, g = _inputs(s)
    s.engine.snapshot()
    J1, ga1, gf1 = _adjoint(s, acc, fden, fvel, g, 37, 5)
    J2, ga2, gf2 = _adjoint(s, acc, fden, fvel, g, 37, 37)
    J3, ga3, gf3 = _adjoint(s, acc, fden, fvel, g, 37, 1)
    assert J1 == J2 == J3
    np.testing.assert_array_equal(ga1, ga2)
    np.testing.assert_array_equal(ga1, ga3)
    np.testing.assert_array_equal(gf1, gf2)
    # and against finite differences at a few sites
    J, J0, adj, fd, adjA, fdA = _fd_checks(s, s, prof, amp, steps=37, interval=5, ncells=8, seed=2)
    assert np.max(np.abs(adj - fd)) <= 1e-6 * np.max(np.abs(fd))
    assert abs(adjA - fdA) <= 1e-6 * abs(fdA)


@pytest.mark.parametrize("layout,obj", [("rows", FACE_X1), ("sparse", OBJ_DISS)])
def test_adjoint_forward_record_reused_and_invalidated(layout, obj):
    """The forward pass records (rho, u*) over its last segment and the reverse
    pass skips that segment's recompute. A second reverse pass (interval <
    steps: the record was overwritten by an earlier segment) recomputes
    instead; both agree bit for bit, and with interval >= steps nothing is
    recomputed. An input set between the passes refuses the reverse pass
    (its recompute would run on other inputs than the forward pass's)."""
    s, prof, amp = make_channel(TRT, "fp64", seed=6, layout=layout)
    acc, fden, fvel, g = _inputs(s)
    eng = s.engine
    eng.snapshot()
    ref = None
    for interval in (29, 6, 40):
        print(f"adjoint record {layout} obj {obj} interval {interval}", flush=True)
        J, ga, gf = _adjoint(s, acc, fden, fvel, g, 29, interval, obj)
        again = eng.adj_backward(FACE_1X)
        np.testing.assert_array_equal(again[0], ga)
        np.testing.assert_array_equal(again[1], gf)
        eng.set_face_data(FACE_1X, fden, fvel)
        with pytest.raises(RuntimeError, match="needs the lb_adj_forward pass"):
            eng.adj_backward(FACE_1X)
        if ref is None:
            ref = (J, ga, gf)
        assert J == ref[0]
        np.testing.assert_array_equal(ga, ref[1])
        np.testing.assert_array_equal(gf, ref[2])
    assert np.max(np.abs(ref[1])) > 0 and np.max(np.abs(ref[2])) > 0


def test_flux_gradient_auto_interval():
    """interval='auto' takes one segment on a small lattice and gives the
    gradients of a checkpointed run bit for bit."""
    from jyulb.adjoint import flux_gradient, checkpoint_interval
    a, _, _ = make_channel(TRT, "fp32", seed=8)
    b, _, _ = make_channel(TRT, "fp32", seed=8)
    assert checkpoint_interval(a, 33) == 33
    Ja, gaa, gfa = flux_gradient(a, 33, "auto", grad_face=FACE_1X)
    Jb, gab, gfb = flux_gradient(b, 33, 4, grad_face=FACE_1X)
    assert Ja == Jb
    np.testing.assert_array_equal(gaa, gab)
    np.testing.assert_array_equal(gfa, gfb)
    np.testing.assert_array_equal(np.asarray(a.get_vel()), np.asarray(b.get_vel()))


def test_adjoint_failed_forward_leaves_no_reverse_pass():
    """A forward pass that fails after its steps (objective face without a
    fluid site) has overwritten the checkpoints: the reverse pass of the
    earlier, successful forward pass is refused afterwards."""
    from jyulb.isothermal import Solver
    from jyulb.defs import PERIODIC
    lat = (8, 6, 5)
    s = Solver(np.array(lat, dtype=np.int32), BGK, 0.1, np.array([1e-5, 0.0, 0.0]),
               np.array([PERIODIC] * 6, dtype=np.int32))
    ph = s.get_phase()
    ph[...] = 1
    ph[-1, :, :] = 0  # the x = nx-1 layer is solid
    s.finalize_init()
    eng = s.engine
    eng.adj_forward(6, 4, FACE_1X, X)
    ga, _ = eng.adj_backward()
    assert np.max(np.abs(ga)) > 0
    with pytest.raises(ValueError, match="no fluid site"):
        eng.adj_forward(6, 4, FACE_X1, X)
    with pytest.raises(RuntimeError, match="needs the lb_adj_forward pass"):
        eng.adj_backward()


def test_adjoint_reverse_pass_tied_to_its_forward_pass():
    """An evolve call after the forward pass moves the state on: the reverse
    pass is refused. A reverse pass right after its forward pass leaves the
    populations, the step count and the AA parity as the forward pass left
    them, so evolving on gives the same bits as evolving throughout."""
    s, _, _ = make_channel(TRT, "fp64", seed=12)
    t, _, _ = make_channel(TRT, "fp64", seed=12)
    eng = s.engine
    s._set_bc_data()
    s._load_inputs()
    t._set_bc_data()
    t._load_inputs()
    eng.adj_forward(7, 3, FACE_X1, X)
    eng.evolve(3)
    with pytest.raises(RuntimeError, match="needs the lb_adj_forward pass"):
        eng.adj_backward(FACE_1X)
    eng.adj_forward(7, 3, FACE_X1, X)  # from step 10
    ga, gf = eng.adj_backward(FACE_1X)
    assert np.max(np.abs(ga)) > 0
    assert eng.info()["steps"] == 17 and eng.info()["next_odd"] == 0
    eng.evolve(5)
    t.engine.evolve(22)
    s._pull_fields()
    t._pull_fields()
    np.testing.assert_array_equal(np.asarray(s.get_den()), np.asarray(t.get_den()))
    np.testing.assert_array_equal(np.asarray(s.get_vel()), np.asarray(t.get_vel()))


def test_autograd_backward_refused_after_another_forward():
    """Two objectives evaluated on one solver: the first one's backward would
    run against the second forward pass's checkpoints, so it raises."""
    import torch
    from jyulb.adjoint import flux_objective
    s, prof, amp = make_channel(BGK, "fp64", seed=13)
    acc = torch.tensor(np.ascontiguousarray(s.get_acc()), device="cuda", requires_grad=True)
    J1 = flux_objective(acc, None, s, 12, 4, FACE_X1, X)
    J2 = flux_objective(acc, None, s, 12, 4, FACE_X1, X)
    with pytest.raises(RuntimeError, match="another adjoint forward pass"):
        (J1 + J2).backward()
    J3 = flux_objective(acc, None, s, 12, 4, FACE_X1, X)
    J3.backward()  # its own forward pass: fine
    assert acc.grad is not None and float(acc.grad.abs().max()) > 0


def test_adjoint_torch_autograd_function():
    """The torch.autograd.Function returns the engine's gradients."""
    import torch
    from jyulb.adjoint import flux_objective
    s, prof, amp = make_channel(BGK, "fp64", seed=9)
    acc, fden, fvel, g = _inputs(s)
    s.engine.snapshot()
    J, ga, gf = _adjoint(s, acc, fden, fvel, g, 40, 8)
    s.engine.restore()
    acc_t = torch.tensor(acc, device="cuda", requires_grad=True)
    fv_t = torch.tensor(fvel, device="cuda", requires_grad=True)
    Jt = flux_objective(acc_t, fv_t, s, 40, 8, FACE_X1, X, FACE_1X)
    Jt.backward()
    assert abs(float(Jt.detach()) - J) <= 1e-14 * abs(J)
    np.testing.assert_allclose(acc_t.grad.cpu().numpy(), ga, rtol=0, atol=1e-14 * np.max(np.abs(ga)))
    np.testing.assert_allclose(fv_t.grad.cpu().numpy(), gf, rtol=0, atol=1e-14 * np.max(np.abs(gf)))


@pytest.mark.parametrize("rlx", [BGK, TRT, REGULARIZATION], ids=["bgk", "trt", "reg"])
def test_adjoint_matches_torch_fp64_autograd(rlx):
    """Gradients of the engine's FP64 adjoint against torch.autograd through
    an independent FP64 PyTorch restatement of the forward model (tests/torch_lb.py)
    started from the same post-collision state."""
    import torch
    import torch_lb
    print(f"torch reference rlx={rlx}", flush=True)
    lat = (10, 6, 5)
    s, prof, amp = make_channel(rlx, "fp64", lat=lat, steps_bc=(WALL, SYMMETRIC), seed=21)
    acc, fden, fvel, g = _inputs(s)
    s.engine.snapshot()
    f0 = s.post_collision_f()[:, 1:-1, 1:-1, 1:-1].reshape(19, -1)
    steps = 30
    J, ga, gf = _adjoint(s, acc, fden, fvel, g, steps, 7)
    ref = torch_lb.TorchLB(np.asarray(s.get_phase()).copy(), s.bc, rlx, 0.2 / 3.0)
    acc_t = torch.tensor(acc.reshape(-1, 3), dtype=torch.float64, requires_grad=True)
    v_t = torch.tensor(fvel.reshape(-1, 3), dtype=torch.float64, requires_grad=True)
    rho, us = ref.run(torch.tensor(f0), acc_t, v_t, 1.0, torch.tensor(g), steps)
    nx = lat[0]
    out = torch.tensor(np.argwhere(np.asarray(s.get_phase()) != 0)[:, 0] == nx - 1)
    Jt = (rho[out] * us[0][out]).mean()
    Jt.backward()
    assert abs(float(Jt) - J) <= 1e-12 * abs(J), (float(Jt), J)
    gt = acc_t.grad.numpy().reshape(ga.shape)
    fluid = np.asarray(s.get_phase()) != 0
    scale = np.max(np.abs(gt[fluid]))
    assert np.max(np.abs(ga[fluid] - gt[fluid])) <= 1e-9 * scale
    gvt = v_t.grad.numpy().reshape(gf.shape)
    np.testing.assert_allclose(gf[1:-1, 1:-1], gvt[1:-1, 1:-1], rtol=0, atol=1e-9 * np.max(np.abs(gvt)))


def make_porous_channel(precision, lat=(128, 64, 64), seed=4, peak=0.02, layout="rows"):
    """Adjoint workload: porous channel of seeded spheres (radius 4..8, 10 %
    solid, none in 8-layer buffers behind the inlet and outlet faces), BGK
    tau = 0.7, gravity gx = 1e-5, parabolic inlet (FIXED_VEL at 1X, peak
    `peak`), outlet density 1 (FIXED_DEN at X1), walls in y and z."""
    from jyulb.isothermal import Solver
    from jyulb.geometry import porous_spheres
    nx, ny, nz = lat
    s = Solver(np.array(lat, dtype=np.int32), BGK, 0.2 / 3.0, np.array([1e-5, 0.0, 0.0]),
               np.array([FIXED_VEL, FIXED_DEN, WALL, WALL, WALL, WALL], dtype=np.int32), precision=precision,
               layout=layout)
    solid = porous_spheres(lat, 4, 8, 0.10, seed, open_faces=(FACE_1X, FACE_X1), open_depth=8)
    s.get_phase()[...] = np.where(solid, 0, 1)
    prof = np.zeros((ny + 2, nz + 2))
    ey = (np.arange(ny) + 0.5) / ny - 0.5
    ez = (np.arange(nz) + 0.5) / nz - 0.5
    prof[1:-1, 1:-1] = (1 - (2 * ey[:, None]) ** 2) * (1 - (2 * ez[None, :]) ** 2)
    s.get_vel()[0, :, :, 0] = peak * prof[1:-1, 1:-1]
    s.finalize_init()
    return s, prof, peak


def test_adjoint_porous_channel_full_size_fp32_vs_fp64_fd():
    """128 x 64 x 64 porous channel, 1000 steps, checkpoint every 50: FP32
    adjoint gradients of the mean outlet x-flux with respect to the body
    force at 64 seeded random fluid cells and to the inlet amplitude, against
    FP64 central differences (h = 1e-6). Bar: 1e-4 relative to the largest
    sampled gradient (cells) and to the amplitude gradient."""
    print("porous channel 128x64x64, 1000 steps, fp32 adjoint vs fp64 FD", flush=True)
    s32, prof, amp = make_porous_channel("fp32")
    s64, _, _ = make_porous_channel("fp64")
    J, J0, adj, fd, adjA, fdA = _fd_checks(s32, s64, prof, amp, steps=1000, interval=50, ncells=64, seed=64)
    scale = np.max(np.abs(fd))
    err = np.max(np.abs(adj - fd)) / scale
    print(f"J fp32 {J:.9e} fp64 {J0:.9e}; cell gradients max rel err {err:.2e}; "
          f"amplitude {adjA:.9e} vs FD {fdA:.9e}", flush=True)
    assert abs(J - J0) <= 1e-5 * abs(J0)
    assert err < 1e-4, (err, adj, fd)
    assert abs(adjA - fdA) <= 1e-4 * abs(fdA), (adjA, fdA)


@pytest.mark.parametrize("precision", ["fp16", "fp32_storage"])
def test_adjoint_rejects_other_precisions(precision):
    s, prof, amp = make_channel(BGK, precision)
    acc, fden, fvel, g = _inputs(s)
    with pytest.raises(ValueError, match="adjoint runs in FP64 or FP32"):
        s.engine.adj_forward(10, 5, FACE_X1, X)


def test_adjoint_identical_with_xcd_tile_order(monkeypatch):
    """The XCD tile order (used on large lattices) does not change a bit of
    the forward objective or the gradients."""
    out = []
    for env in (None, "0"):
        if env is not None:
            monkeypatch.setenv("LB_XCD_REMAP_CELLS", env)
        s, prof, amp = make_channel(TRT, "fp32", steps_bc=(PERIODIC, PERIODIC), seed=8)
        acc, fden, fvel, g = _inputs(s)
        s.engine.snapshot()
        out.append(_adjoint(s, acc, fden, fvel, g, 30, 7))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[0][1], out[1][1])
    np.testing.assert_array_equal(out[0][2], out[1][2])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("rlx,steps_bc", [(BGK, (WALL, SYMMETRIC)), (TRT, (PERIODIC, PERIODIC)),
                                          (REGULARIZATION, (WALL, SYMMETRIC))], ids=["bgk", "trt", "reg"])
def test_adjoint_identical_in_sparse_layout(precision, rlx, steps_bc):
    """The sparse layout's adjoint kernel reads and writes the cotangents at
    the forward step's odd-step locations; objective and both gradients are
    bit-identical to the dense layout's (Zou-He inlet/outlet, walls,
    symmetric or periodic faces, seeded solids)."""
    out = []
    for layout in ("rows", "sparse"):
        s, prof, amp = make_channel(rlx, precision, steps_bc=steps_bc, seed=8, layout=layout)
        acc, fden, fvel, g = _inputs(s)
        s.engine.snapshot()
        out.append(_adjoint(s, acc, fden, fvel, g, 30, 7))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][2], out[0][2])


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
@pytest.mark.parametrize("pre_steps", [0, 1], ids=["from_odd", "from_even"])
@pytest.mark.parametrize("layout", ["sparse", "rows"])
@pytest.mark.parametrize("steps_bc", [(PERIODIC, PERIODIC), (WALL, SYMMETRIC)], ids=["periodic", "symmetric"])
def test_adjoint_pull_forms_identical(monkeypatch, precision, pre_steps, layout, steps_bc):
    """The forward and recompute passes in place or as pull steps
    (LB_SP_PULL), and the adjoint steps in place or in pull form
    (LB_ADJ_PULL, a second cotangent array), from either parity, with
    periodic or wall and symmetric faces beside the Zou-He inlet and outlet.
    Objective and both gradients are bit-identical in every combination, for
    the flux objective with segments of unequal length (30 steps, checkpoint
    every 7) and for the dissipation objective."""
    from jyulb import adjoint
    out, diss = {}, {}
    for sp in ("0", "1"):
        for ap in ("0", "1"):
            monkeypatch.setenv("LB_SP_PULL", sp)
            monkeypatch.setenv("LB_ADJ_PULL", ap)
            print(f"pull forms {layout} {steps_bc} {precision} pre {pre_steps} LB_SP_PULL={sp} LB_ADJ_PULL={ap}",
                  flush=True)
            s, prof, amp = make_channel(TRT, precision, steps_bc=steps_bc, seed=8, layout=layout)
            assert s.engine.pull_steps() == (sp == "1")
            if pre_steps:
                s.engine.evolve(pre_steps)
            acc, fden, fvel, g = _inputs(s)
            s.engine.snapshot()
            out[sp, ap] = _adjoint(s, acc, fden, fvel, g, 30, 7)
            s.engine.restore()
            diss[sp, ap] = adjoint.dissipation_gradient(s, 20, 8, grad_face=FACE_1X)
    ref = out["0", "0"]
    assert np.any(ref[1] != 0) and np.any(ref[2] != 0)
    for key, r in out.items():
        assert r[0] == ref[0], key
        np.testing.assert_array_equal(r[1], ref[1], err_msg=str(key))
        np.testing.assert_array_equal(r[2], ref[2], err_msg=str(key))
    dref = diss["0", "0"]
    for key, r in diss.items():
        assert r[0] == dref[0], key
        np.testing.assert_array_equal(r[1], dref[1], err_msg=str(key))
        np.testing.assert_array_equal(r[2], dref[2], err_msg=str(key))


def test_adjoint_porous_channel_sparse_equals_rows():
    """Full-size porous channel (128x64x64, FP32, 60 steps, checkpoint every
    25): the sparse layout's gradients equal the dense layout's bit for bit."""
    out = []
    for layout in ("rows", "sparse"):
        s, prof, amp = make_porous_channel("fp32", layout=layout)
        acc, fden, fvel, g = _inputs(s)
        s.engine.snapshot()
        out.append(_adjoint(s, acc, fden, fvel, g, 60, 25))
    assert out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][2], out[0][2])


@pytest.mark.parametrize("rlx", [BGK, TRT, REGULARIZATION], ids=["bgk", "trt", "reg"])
def test_dissipation_adjoint_fp64_matches_finite_differences(rlx):
    """The dissipation objective (mean Pi_neq:Pi_neq of the final step's
    collision input, seeded by k_diss) against FP64 central differences, for
    the acceleration field and the Zou-He inlet amplitude."""
    print(f"dissipation adjoint fp64 rlx={rlx}", flush=True)
    s, prof, amp = make_channel(rlx, "fp64")
    J, J0, adj, fd, adjA, fdA = _fd_checks(s, s, prof, amp, steps=60, interval=16, ncells=16, seed=9, obj=OBJ_DISS)
    assert J > 0 and abs(J - J0) <= 1e-13 * abs(J0)
    scale = np.max(np.abs(fd))
    assert scale > 0
    err = np.max(np.abs(adj - fd)) / scale
    assert err < 1e-6, (err, adj, fd)
    assert abs(adjA - fdA) <= 1e-6 * abs(fdA), (adjA, fdA)


def test_dissipation_adjoint_fp32_within_1e4_of_fp64_fd():
    s32, prof, amp = make_channel(TRT, "fp32")
    s64, _, _ = make_channel(TRT, "fp64")
    J, J0, adj, fd, adjA, fdA = _fd_checks(s32, s64, prof, amp, steps=60, interval=16, ncells=16, seed=13,
                                           obj=OBJ_DISS)
    scale = np.max(np.abs(fd))
    err = np.max(np.abs(adj - fd)) / scale
    assert err < 1e-4, (err, adj, fd)
    assert abs(adjA - fdA) <= 1e-4 * abs(fdA), (adjA, fdA)


@pytest.mark.parametrize("precision", ["fp64", "fp32"])
def test_dissipation_gradient_sparse_equals_rows(precision):
    """jyulb.adjoint.dissipation_gradient in both layouts: objective and
    gradients bit-identical."""
    from jyulb import adjoint
    out = []
    for layout in ("rows", "sparse"):
        s, prof, amp = make_channel(TRT, precision, seed=19, layout=layout)
        out.append(adjoint.dissipation_gradient(s, 20, 8, grad_face=FACE_1X))
    assert out[0][0] > 0 and out[0][0] == out[1][0]
    np.testing.assert_array_equal(out[1][1], out[0][1])
    np.testing.assert_array_equal(out[1][2], out[0][2])


@pytest.mark.parametrize("rlx", [BGK, TRT, REGULARIZATION], ids=["bgk", "trt", "reg"])
def test_dissipation_adjoint_matches_torch_fp64_autograd(rlx):
    """The dissipation objective's gradients against torch.autograd through
    the FP64 PyTorch restatement (tests/torch_lb.py)."""
    import torch
    import torch_lb
    lat = (10, 6, 5)
    s, prof, amp = make_channel(rlx, "fp64", lat=lat, steps_bc=(WALL, SYMMETRIC), seed=23)
    acc, fden, fvel, g = _inputs(s)
    s.engine.snapshot()
    f0 = s.post_collision_f()[:, 1:-1, 1:-1, 1:-1].reshape(19, -1)
    steps = 30
    J, ga, gf = _adjoint(s, acc, fden, fvel, g, steps, 7, obj=OBJ_DISS)
    ref = torch_lb.TorchLB(np.asarray(s.get_phase()).copy(), s.bc, rlx, 0.2 / 3.0)
    acc_t = torch.tensor(acc.reshape(-1, 3), dtype=torch.float64, requires_grad=True)
    v_t = torch.tensor(fvel.reshape(-1, 3), dtype=torch.float64, requires_grad=True)
    ref.run(torch.tensor(f0), acc_t, v_t, 1.0, torch.tensor(g), steps)
    Jt = ref.dissipation()
    Jt.backward()
    assert abs(float(Jt.detach()) - J) <= 1e-11 * abs(J), (float(Jt.detach()), J)
    gt = acc_t.grad.numpy().reshape(ga.shape)
    fluid = np.asarray(s.get_phase()) != 0
    scale = np.max(np.abs(gt[fluid]))
    assert np.max(np.abs(ga[fluid] - gt[fluid])) <= 1e-9 * scale
    gvt = v_t.grad.numpy().reshape(gf.shape)
    np.testing.assert_allclose(gf[1:-1, 1:-1], gvt[1:-1, 1:-1], rtol=0, atol=1e-9 * np.max(np.abs(gvt)))


def test_dissipation_objective_torch_function():
    """The autograd Function form of the dissipation objective returns the
    gradients of dissipation_gradient."""
    import torch
    from jyulb import adjoint
    s, prof, amp = make_channel(TRT, "fp64", seed=17)
    acc, fden, fvel, g = _inputs(s)
    s.engine.snapshot()
    J0, ga0, gf0 = _adjoint(s, acc, fden, fvel, g, 24, 8, obj=OBJ_DISS)
    s.engine.restore()
    a = torch.tensor(acc, device="cuda", requires_grad=True)
    v = torch.tensor(fvel, device="cuda", requires_grad=True)
    J = adjoint.dissipation_objective(a, v, s, 24, 8, inlet_face=FACE_1X)
    J.backward()
    assert abs(float(J.detach()) - J0) <= 1e-14 * abs(J0)
    np.testing.assert_allclose(a.grad.cpu().numpy(), ga0, rtol=0, atol=1e-14 * np.max(np.abs(ga0)))
    np.testing.assert_allclose(v.grad.cpu().numpy(), gf0, rtol=0, atol=1e-14 * np.max(np.abs(gf0)))
