"""Drop-in API behaviour on the GPU: calc_flow_info, the host fields that
Solver.evolve reads and writes in place, with_iflow, phase edits after
finalize_init, the automatic layout, and the FP16 storage guard.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

from jyulb.defs import BGK, TRT, REGULARIZATION, WALL, PERIODIC, SYMMETRIC, FIXED_DEN, FIXED_VEL


def _flow_info_numpy(den, vel, fluid, rho0=1.0):
    """calc_flow_info (include/solver/solver.cpp:64-129) restated in numpy."""
    d, v = den[fluid], vel[fluid]
    u2 = np.sqrt((v * v).sum(1))
    n = d.size
    return dict(tot_jx=(d * v[:, 0]).sum(), tot_jy=(d * v[:, 1]).sum(), tot_jz=(d * v[:, 2]).sum(),
                tot_j2=(d * u2).sum(), tot_den=d.sum(), ave_ux=v[:, 0].sum() / n, ave_uy=v[:, 1].sum() / n,
                ave_uz=v[:, 2].sum() / n, ave_u2=u2.sum() / n, max_u2=max(u2.max(), 0.0),
                max_den=max(d.max(), rho0), min_den=min(d.min(), rho0), fluid_sites=float(n))


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16", "fp32_storage"])
def test_flow_info_matches_numpy(precision):
    """Every field of calc_flow_info, reduced in FP64 on the device, against
    numpy on the downloaded den/vel (a porous channel with walls, gravity and
    a Zou-He inlet, so that every sum is nonzero). The density extrema start
    from the reference density, as in the reference: an all-compressed field
    reports min_den = rho0."""
    import bench
    s, nfl = bench.porous_channel(precision, lat=(48, 24, 20))
    s.evolve(150)
    den, vel = np.asarray(s.get_den()), np.asarray(s.get_vel())
    fl = np.asarray(s.get_phase()) != 0
    got = s.flow_info()
    ref = _flow_info_numpy(den, vel, fl)
    print(precision, {k: f"{got[k]:.6e}" for k in ref}, flush=True)
    assert got.keys() == ref.keys()
    for k in ref:
        scale = {"tot_jx": abs(ref["tot_j2"]), "tot_jy": abs(ref["tot_j2"]), "tot_jz": abs(ref["tot_j2"]),
                 "ave_ux": ref["ave_u2"], "ave_uy": ref["ave_u2"], "ave_uz": ref["ave_u2"]}.get(k, abs(ref[k]))
        assert abs(got[k] - ref[k]) <= 1e-12 * scale, (k, got[k], ref[k])
    assert got["fluid_sites"] == nfl
    assert got["max_den"] > 1.0 > got["min_den"] and got["ave_ux"] > 0


def test_phase_edit_after_finalize_raises():
    """The engine builds its links from the phase at finalize_init; the
    reference re-reads phase every step (isothermal.pyx:1122-1248). An edit
    that turns a site solid or fluid afterwards is refused, not ignored; an
    edit that keeps every site's state (another nonzero FLUID value) is not."""
    from jyulb.isothermal import Solver
    s = Solver(np.array([8, 6, 5], dtype=np.int32), TRT, 0.1, np.array([1e-5, 0, 0]), [PERIODIC] * 6)
    ph = s.get_phase()
    ph[...] = 1
    ph[3, 2, 2] = 0
    s.finalize_init()
    s.evolve(3)
    ph[1, 1, 1] = 7  # still fluid
    s.evolve(3)
    ph[4, 4, 4] = 0
    with pytest.raises(RuntimeError, match="phase field was changed after finalize_init"):
        s.evolve(1)
    ph[4, 4, 4] = 1
    assert s.evolve(2) == 8


def test_host_fields_keep_solid_values_and_halos():
    """Solver.evolve writes den/vel of the fluid sites only: values the user
    put at SOLID sites and in WALL halos stay; the halos of periodic and
    symmetric faces hold the halo pass of step T-1 (isothermal.pyx:1408-1415),
    like the CPU oracle's."""
    import jyulb_oracle as jo
    from jyulb.isothermal import Solver
    lat, bc = np.array([7, 6, 5], dtype=np.int32), np.array([PERIODIC, PERIODIC, SYMMETRIC, WALL, WALL, WALL])
    outs = []
    for mk in (jo.OracleSolver, Solver):
        s = mk(lat, BGK, 0.1, np.array([1e-5, 2e-6, 0.0]), bc)
        ph = s.get_phase()
        ph[...] = 1
        ph[2, 3, 1] = 0
        s.finalize_init()
        s.evolve(4)
        s.den[3, 4, 2] = 1.5  # (interior 2, 3, 1): a SOLID site, never written
        s.vel[0, 0, 0] = [1.0, 2.0, 3.0]  # a corner halo no pass touches
        s.den[3, 7, 3] = 0.7  # the WALL halo of face Y1
        s.evolve(5)
        outs.append((np.array(s.den), np.array(s.vel)))
    for a, b in zip(outs[0], outs[1]):
        np.testing.assert_array_equal(b, a)
    assert outs[1][0][3, 4, 2] == 1.5 and outs[1][0][3, 7, 3] == 0.7
    assert list(outs[1][1][0, 0, 0]) == [1.0, 2.0, 3.0]


def test_with_iflow_matches_oracle():
    """Solver.with_iflow (isothermal.pyx:301-360): phase and the initial
    den/vel prescribed at construction, here with gravity and walls, against
    the oracle built the same way through its views."""
    import jyulb_oracle as jo
    from jyulb.isothermal import Solver
    rng = np.random.default_rng(21)
    shape = (9, 7, 6)
    phase = (rng.random(shape) > 0.1).astype(np.uint8)
    iden = 1.0 + 0.01 * (rng.random(shape) - 0.5)
    ivel = 0.02 * (rng.random(shape + (3,)) - 0.5)
    ivel[phase == 0] = 0.0
    bc = np.array([PERIODIC, PERIODIC, WALL, WALL, SYMMETRIC, SYMMETRIC], dtype=np.int32)
    g = np.array([1e-5, 0.0, -1e-5])
    got = Solver.with_iflow(phase, iden, ivel, REGULARIZATION, 0.07, g, bc)
    ref = jo.OracleSolver(np.array(shape, dtype=np.int32), REGULARIZATION, 0.07, g, bc)
    ref.get_phase()[...] = phase
    ref.get_den()[...] = iden
    ref.get_vel()[...] = ivel
    assert got.finalize_init() == ref.finalize_init() == int(phase.sum())
    for st in (5, 8):
        got.evolve(st)
        ref.evolve(st)
        np.testing.assert_array_equal(np.asarray(got.den), np.asarray(ref.den))
        np.testing.assert_array_equal(np.asarray(got.vel), np.asarray(ref.vel))


@pytest.mark.parametrize("solid_fraction, expect", [(0.0, "rows"), (0.01, "rows"), (0.2, "sparse")])
def test_auto_layout_choice(solid_fraction, expect):
    """layout='auto' (the default) resolves at finalize_init: sparse from 5 %
    solid cells on. The result is the same in either layout."""
    from jyulb.isothermal import Solver
    lat = (32, 16, 24)
    rng = np.random.default_rng(3)
    phase = (rng.random(lat) >= solid_fraction).astype(np.uint8)
    res = []
    for layout in ("auto", "rows"):
        s = Solver(np.array(lat, dtype=np.int32), TRT, 0.1, np.array([1e-5, 0, 0]), [PERIODIC] * 6, layout=layout)
        assert s.engine.layout() == layout
        s.get_phase()[...] = phase
        s.finalize_init()
        assert s.engine.layout() == (expect if layout == "auto" else "rows")
        s.evolve(7)
        res.append(np.array(s.get_vel()))
    np.testing.assert_array_equal(res[0], res[1])


def _gravity_channel(precision, g, steps, calls=1):
    from jyulb.isothermal import Solver
    s = Solver(np.array([4, 34, 4], dtype=np.int32), BGK, 0.1, np.array([g, 0.0, 0.0]), [PERIODIC] * 6,
               precision=precision)
    ph = s.get_phase()
    ph[...] = 1
    ph[:, 0, :] = 0
    ph[:, -1, :] = 0
    s.finalize_init()
    for _ in range(calls):
        s.evolve(steps // calls)
    return s


def test_fp16_guard_gravity_channel_from_rest():
    """FP16 storage of a gravity-driven channel started at rest: the initial
    scale is chosen for a fluid at rest (2^23), and the flow accelerates to
    u ~ 0.05, where |f - w rho0| reaches 8e-3, i.e. 7e4 in storage units,
    beyond the f16 range. The guard rescales the stored populations by powers
    of two as they grow; the run stays finite and within FP16 accuracy of the
    FP64 run."""
    g = 3.9e-5  # u_max = g H^2 / (8 nu) = 0.05 for H = 32, nu = 0.1
    s16 = _gravity_channel("fp16", g, 4000, calls=4)
    s64 = _gravity_channel("fp64", g, 4000, calls=4)
    v16, v64 = np.asarray(s16.get_vel()), np.asarray(s64.get_vel())
    assert np.isfinite(v16).all()
    umax = float(np.max(np.abs(v64)))
    print(f"FP16 channel: u_max {umax:.4f}, scale {s16.engine.info()['fp16_scale']}", flush=True)
    assert 0.04 < umax < 0.06
    assert s16.engine.info()["fp16_scale"] <= 2.0 ** 19  # rescaled down from the rest-state 2^23
    fl = np.asarray(s64.get_phase()) != 0
    err = float(np.max(np.abs(v16[fl] - v64[fl]))) / umax
    print(f"FP16 vs FP64: max rel velocity error {err:.3e}", flush=True)
    assert err <= 1e-1


@pytest.mark.parametrize("layout", ["rows", "sparse"])
def test_fp16_guard_with_pull_steps(layout, monkeypatch):
    """The FP16 guard's rescales act on the current population array, which
    the pull steps swap every step: forced pull steps give the in-place
    steps' bits, rescales included."""
    from jyulb.isothermal import Solver
    out = []
    for pull in ("1", "0"):
        monkeypatch.setenv("LB_SP_PULL", pull)
        s = Solver(np.array([4, 34, 4], dtype=np.int32), BGK, 0.1, np.array([3.9e-5, 0.0, 0.0]), [PERIODIC] * 6,
                   precision="fp16", layout=layout)
        ph = s.get_phase()
        ph[...] = 1
        ph[:, 0, :] = 0
        ph[:, -1, :] = 0
        s.finalize_init()
        for _ in range(3):
            s.evolve(333)
        assert s.engine.pull_steps() == (pull == "1")
        out.append((np.array(s.get_vel()), s.engine.info()["fp16_scale"]))
    assert out[0][1] == out[1][1] < 2.0 ** 23
    np.testing.assert_array_equal(out[0][0], out[1][0])


def test_fp16_guard_zou_he_inlet_from_rest():
    """A Zou-He inlet of 0.05 switched on for a fluid at rest: the scale chosen
    at finalize_init for the fluid at rest (2^23) would overflow on the first
    step (|f - w rho0| ~ 3e-2 at the inlet). The guard sizes it again from the
    face data at the start of the call that sees them, by an exact power of
    two (2^-9), so the run is bit for bit the run with that scale (2^14) from
    the start, and finite. Against FP64 this impulsive start keeps the FP16
    storage error of its acoustic transient (rel. L2 1.7e-2 measured with
    either scale)."""
    from jyulb.isothermal import Solver
    from jyulb.defs import REGULARIZATION

    def run(prec, **kw):
        s = Solver(np.array([48, 12, 12], dtype=np.int32), REGULARIZATION, (0.55 - 0.5) / 3.0, np.zeros(3),
                   np.array([FIXED_VEL, FIXED_DEN, WALL, WALL, WALL, WALL], dtype=np.int32), precision=prec, **kw)
        s.get_phase()[...] = 1
        s.finalize_init()
        s.evolve(10)
        s.get_vel()[0, :, :, 0] = 0.05  # the inlet layer: _set_bc_data copies it to the face data
        s.evolve(300)
        return s

    auto, fixed, s64 = run("fp16"), run("fp16", fp16_scale=2.0 ** 14), run("fp64")
    assert auto.engine.info()["fp16_scale"] == 2.0 ** 14
    v16, vfix, v64 = (np.asarray(x.get_vel()) for x in (auto, fixed, s64))
    assert np.isfinite(v16).all()
    np.testing.assert_array_equal(v16, vfix)
    l2 = float(np.sqrt(np.sum((v16 - v64) ** 2) / np.sum(v64 ** 2)))
    print(f"FP16 inlet: rel. L2 velocity error against FP64 {l2:.3e}", flush=True)
    assert l2 <= 3e-2


@pytest.mark.parametrize("nbytes, offset", [(8, 0), (2 ** 20 + 8 * 37, 0), (3 * 2 ** 27, 0), (1000, 3), (7, 8)])
def test_stream_copy_matches_source(nbytes, offset):
    """lb_stream_copy (the engine's device copy for checkpoints and
    snapshots): whole 8-byte elements by its own kernel, any other size or
    alignment through hipMemcpyAsync; bytes outside the range stay as they
    were."""
    import torch
    from jyulb import _native
    n = nbytes + offset + 64
    src = torch.randint(0, 256, (n,), dtype=torch.uint8, device="cuda")
    dst = torch.full((n,), 0xAB, dtype=torch.uint8, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    _native.stream_copy(dst.data_ptr() + offset, src.data_ptr() + offset, nbytes, st)
    torch.cuda.synchronize()
    assert torch.equal(dst[offset:offset + nbytes], src[offset:offset + nbytes])
    assert bool((dst[:offset] == 0xAB).all()) and bool((dst[offset + nbytes:] == 0xAB).all())


@pytest.mark.parametrize("pull", ["0", "1"])
@pytest.mark.parametrize("layout", ["rows", "sparse"])
def test_f2_holds_the_step_before_the_last(layout, pull, monkeypatch):
    """The reference's f2 (isothermal.pxd:150-157) after evolve holds the
    post-collision populations of the step before the last one (its arrays
    swap every step, isothermal.pyx:1408-1415): with keep_f2=True, f2 after
    5 + 3 steps equals f1 after 7, bit for bit, with in-place and with pull
    steps; before the first evolve call f2 is zero; without keep_f2 it is
    refused."""
    from jyulb.isothermal import Solver
    monkeypatch.setenv("LB_SP_PULL", pull)
    lat = np.array([12, 10, 9], dtype=np.int32)
    bc = np.array([PERIODIC, PERIODIC, WALL, WALL, SYMMETRIC, SYMMETRIC], dtype=np.int32)

    def make(**kw):
        s = Solver(lat, TRT, 0.08, np.array([2e-5, 0.0, 1e-6]), bc, layout=layout, **kw)
        ph = s.get_phase()
        ph[...] = 1
        ph[4:7, 3:6, 2:5] = 0
        s.finalize_init()
        return s

    a, b = make(keep_f2=True), make()
    assert not np.any(a.f2)
    a.evolve(5)
    a.evolve(3)
    b.evolve(5)
    b.evolve(2)
    np.testing.assert_array_equal(a.f2, b.f1)
    b.evolve(1)
    np.testing.assert_array_equal(a.f1, b.f1)
    with pytest.raises(AttributeError, match="keep_f2=True"):
        b.f2
