"""Pins of the CPU FP64 oracle that do not come from the reference's output.

The golden vectors (test_oracle_golden.py) pin the oracle end to end against
the reference's own solver. The checks here pin its parts against what the
mathematics fixes, so that a dropped term, a wrong sign or a transposed index
in one operator fails on its own:

  * Hermite projection (D3Q19.pyx:460-599): moments of P(h) reproduce h;
    f_to_h (395-452) is the brute-force moment sum and inverts P on H10.
  * Collisions (D3Q19.pyx:655-940): exact mass conservation, the Guo
    momentum update j -> j + rho a, the stored half-force shifted velocity;
    TRT with equal rates is BGK; the regularized scheme is BGK on populations
    in the H10 span; BGK with unit rate is the projected equilibrium plus the
    Guo forcing moments; the TRT magic parameter (isothermal.pyx:284) is 3/16.
  * Zou-He faces (bcs.pyx:31-716): known populations untouched, prescribed
    density / momentum reproduced, for all 6 faces and both kinds.
  * Whole steps (isothermal.pyx:1112-1418): the uniform-force known answer
    u(t) = u0 + (t - 1/2) a in a periodic box, mass conservation with
    bounce-back walls, Couette flow between Ladd walls and TRT (magic 3/16)
    Poiseuille flow between bounce-back walls at their analytic steady states
    (flow_field.pyx:20-123), which these schemes reproduce to round-off.
"""
import numpy as np
import pytest

import jyulb_oracle as jo

LI = np.array([-1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1])
LJ = np.array([-1, 0, 0, 0, 1, -1, -1, -1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 0, 1])
LK = np.array([0, -1, 0, 1, 0, -1, 0, 1, -1, 0, 1, -1, 0, 1, 0, -1, 0, 1, 0])
C = np.stack([LI, LJ, LK], axis=1).astype(np.float64)  # (19, 3)
H2_PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]  # H10 order: XX XY XZ YY YZ ZZ
BGK, TRT, REG = jo.BGK, jo.TRT, jo.REGULARIZATION
FIXED_DEN, FIXED_VEL = jo.FIXED_DEN, jo.FIXED_VEL


def moments(f):
    """Brute-force H10 moments of f: rho, j, sum (c_a c_b - delta_ab/3) f."""
    h = np.zeros(10)
    h[0] = f.sum()
    h[1:4] = C.T @ f
    for q, (a, b) in enumerate(H2_PAIRS):
        h[4 + q] = np.sum((C[:, a] * C[:, b] - (a == b) / 3.0) * f)
    return h


def rand_f(rng, rho=1.0, amp=0.02):
    """Populations near rest equilibrium with a random perturbation."""
    w = np.where(np.abs(C).sum(1) == 0, 1 / 3, np.where(np.abs(C).sum(1) == 1, 1 / 18, 1 / 36))
    return rho * w * (1.0 + amp * (rng.random(19) - 0.5) * 20)


def prm_for(rlx, tau=0.8):
    we = 1.0 / tau
    wo = 8.0 * (2.0 - we) / (8.0 - we) if rlx == TRT else 0.0
    return np.array([we, wo])


# ---------------------------------------------------------------------------
# Hermite projection and moments
# ---------------------------------------------------------------------------

def test_weights_are_the_d3q19_set():
    f = jo.project(np.array([1.0, 0, 0, 0, 0, 0, 0, 0, 0, 0]))
    assert np.isclose(f.sum(), 1.0, rtol=0, atol=1e-15)
    assert np.allclose(f[np.abs(C).sum(1) == 0], 1 / 3, rtol=0, atol=1e-16)
    assert np.allclose(f[np.abs(C).sum(1) == 1], 1 / 18, rtol=0, atol=1e-16)
    assert np.allclose(f[np.abs(C).sum(1) == 2], 1 / 36, rtol=0, atol=1e-16)


@pytest.mark.parametrize("seed", range(5))
def test_projection_reproduces_moments(seed):
    rng = np.random.default_rng(seed)
    h = rng.random(10) - 0.5
    f = jo.project(h)
    np.testing.assert_allclose(moments(f), h, rtol=0, atol=1e-15)


@pytest.mark.parametrize("seed", range(5))
def test_f_to_h_is_the_moment_sum_and_inverts_projection(seed):
    rng = np.random.default_rng(100 + seed)
    f = rng.random(19)
    np.testing.assert_allclose(jo.f_to_h(f), moments(f), rtol=0, atol=1e-14)
    h = rng.random(10) - 0.5
    np.testing.assert_allclose(jo.f_to_h(jo.project(h)), h, rtol=0, atol=1e-15)


# ---------------------------------------------------------------------------
# collisions
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("rlx", [BGK, TRT, REG], ids=["bgk", "trt", "reg"])
@pytest.mark.parametrize("seed", range(4))
def test_collision_invariants(rlx, seed):
    rng = np.random.default_rng(seed)
    f = rand_f(rng, rho=1.0 + 0.1 * rng.random())
    acc = 1e-3 * (rng.random(3) - 0.5)
    rho, j = f.sum(), C.T @ f
    den, vel, fp = jo.collide(rlx, f, acc, prm_for(rlx, 0.6 + rng.random()))
    assert abs(den - rho) <= 1e-15 * rho
    assert abs(fp.sum() - rho) <= 1e-14
    np.testing.assert_allclose(C.T @ fp, j + rho * acc, rtol=0, atol=1e-15)
    np.testing.assert_allclose(vel, j / rho + 0.5 * acc, rtol=0, atol=1e-16)


@pytest.mark.parametrize("seed", range(4))
def test_trt_with_equal_rates_is_bgk(seed):
    rng = np.random.default_rng(seed)
    f = rand_f(rng)
    acc = 1e-3 * (rng.random(3) - 0.5)
    w = 0.5 + rng.random()
    _, v1, f1 = jo.collide(TRT, f, acc, np.array([w, w]))
    _, v2, f2 = jo.collide(BGK, f, acc, np.array([w, 0.0]))
    np.testing.assert_allclose(f1, f2, rtol=0, atol=1e-15)
    np.testing.assert_array_equal(v1, v2)


@pytest.mark.parametrize("seed", range(4))
def test_regularized_is_bgk_in_the_hermite_span(seed):
    rng = np.random.default_rng(seed)
    h = np.concatenate([[1.0 + 0.1 * rng.random()], 0.02 * (rng.random(9) - 0.5)])
    f = jo.project(h)
    acc = 1e-3 * (rng.random(3) - 0.5)
    prm = prm_for(BGK, 0.55 + rng.random())
    _, _, fr = jo.collide(REG, f, acc, prm)
    _, _, fb = jo.collide(BGK, f, acc, prm)
    np.testing.assert_allclose(fr, fb, rtol=0, atol=1e-15)
    # ... but not on populations with ghost (non-H10) content
    g = f + 1e-3 * (rng.random(19) - 0.5)
    _, _, gr = jo.collide(REG, g, acc, prm)
    _, _, gb = jo.collide(BGK, g, acc, prm)
    assert np.max(np.abs(gr - gb)) > 1e-6


@pytest.mark.parametrize("seed", range(4))
def test_bgk_unit_rate_is_equilibrium_plus_guo_forcing(seed):
    """omega = 1: f* = P(h) with h0 = rho, h1 = j* + rho a / 2,
    h2_ab = j*_a j*_b / rho + (a_a j*_b + a_b j*_a) / 2 (Guo forcing with
    the half-force shifted momentum j* = j + rho a / 2)."""
    rng = np.random.default_rng(seed)
    f = rand_f(rng, rho=1.0 + 0.2 * rng.random())
    a = 1e-2 * (rng.random(3) - 0.5)
    rho = f.sum()
    js = C.T @ f + 0.5 * rho * a
    h = np.zeros(10)
    h[0] = rho
    h[1:4] = js + 0.5 * rho * a
    for q, (p, r) in enumerate(H2_PAIRS):
        h[4 + q] = js[p] * js[r] / rho + 0.5 * (a[p] * js[r] + a[r] * js[p])
    _, _, fp = jo.collide(BGK, f, a, np.array([1.0, 0.0]))
    np.testing.assert_allclose(fp, jo.project(h), rtol=0, atol=1e-15)


def test_trt_magic_parameter():
    for tau in (0.51, 0.6, 0.8, 1.3, 3.0):
        s = jo.OracleSolver(np.array([2, 2, 2]), TRT, (tau - 0.5) / 3.0)
        we, wo = s.rlx_prm
        assert abs((1 / we - 0.5) * (1 / wo - 0.5) - 3 / 16) < 1e-14
        assert abs(1 / we - tau) < 1e-14


# ---------------------------------------------------------------------------
# Zou-He faces
# ---------------------------------------------------------------------------

@pytest.mark.parametrize("kind", [FIXED_DEN, FIXED_VEL], ids=["den", "vel"])
@pytest.mark.parametrize("face", range(6))
def test_zouhe_reproduces_prescribed_moments(face, kind):
    rng = np.random.default_rng(10 * face + kind)
    axis, sgn_in = face // 2, (1 if face % 2 == 0 else -1)
    unknown = C[:, axis] == sgn_in
    for _ in range(4):
        f = rand_f(rng)
        den = 1.0 + 0.05 * (rng.random() - 0.5)
        vel = 0.05 * (rng.random(3) - 0.5)
        g = jo.zouhe(face, kind, den, vel, f)
        np.testing.assert_array_equal(g[~unknown], f[~unknown])
        assert np.any(g[unknown] != f[unknown])
        rho, j = g.sum(), C.T @ g
        tang = [a for a in range(3) if a != axis]
        if kind == FIXED_DEN:
            assert abs(rho - den) <= 1e-15
            np.testing.assert_allclose(j[tang], den * vel[tang], rtol=0, atol=1e-16)
        else:
            np.testing.assert_allclose(j, rho * vel, rtol=0, atol=1e-16)
            # the density closes the normal momentum balance
            known = C[:, axis] == 0
            out = C[:, axis] == -sgn_in
            expect = (f[known].sum() + 2 * f[out].sum()) / (1.0 - sgn_in * vel[axis])
            assert abs(rho - expect) <= 1e-15


# ---------------------------------------------------------------------------
# whole steps
# ---------------------------------------------------------------------------

def _solver(lat, rlx, kvisc, g, bc):
    return jo.OracleSolver(np.array(lat), rlx, kvisc, np.array(g, dtype=np.float64), np.array(bc))


@pytest.mark.parametrize("rlx", [BGK, TRT, REG], ids=["bgk", "trt", "reg"])
def test_uniform_force_known_answer(rlx):
    """Periodic all-fluid box, uniform rho, u0 and acceleration: the initial
    state has no gradients, streaming changes nothing and each collision adds
    rho a to j, so the stored velocity after t steps is u0 + (t - 1/2) a."""
    a = np.array([2e-5, -1e-5, 3e-5])
    u0 = np.array([0.01, 0.02, -0.015])
    s = _solver((4, 3, 5), rlx, 0.1, a, [jo.PERIODIC] * 6)
    s.get_phase()[...] = 1
    s.get_den()[...] = 1.03
    s.get_vel()[...] = u0
    s.finalize_init()
    t = 0
    for n in (1, 1, 5):
        s.evolve(n)
        t += n
        np.testing.assert_allclose(s.get_vel(), np.broadcast_to(u0 + (t - 0.5) * a, s.get_vel().shape),
                                   rtol=0, atol=1e-15)
        np.testing.assert_allclose(s.get_den(), 1.03, rtol=0, atol=1e-14)


@pytest.mark.parametrize("rlx", [BGK, TRT, REG], ids=["bgk", "trt", "reg"])
def test_mass_conserved_with_bounce_back(rlx):
    rng = np.random.default_rng(4)
    lat = (6, 5, 7)
    s = _solver(lat, rlx, 0.05, [1e-4, 0, -2e-4], [jo.WALL, jo.WALL, jo.PERIODIC, jo.PERIODIC,
                                                    jo.SYMMETRIC, jo.WALL])
    ph = s.get_phase()
    ph[...] = 1
    ph[rng.random(lat) < 0.15] = 0
    s.get_den()[...] = 1.0 + 0.02 * (rng.random(lat) - 0.5)
    fl = ph != 0
    s.get_vel()[fl] = 0.01 * (rng.random((int(fl.sum()), 3)) - 0.5)  # SOLID cells at rest
    s.finalize_init()
    m0 = s.post_collision_f()[:, 1:-1, 1:-1, 1:-1][:, fl].sum()
    s.evolve(40)
    m1 = s.get_den()[fl].sum()
    assert abs(m1 - m0) <= 1e-13 * m0
    assert abs(s.post_collision_f()[:, 1:-1, 1:-1, 1:-1][:, fl].sum() - m0) <= 1e-13 * m0


@pytest.mark.parametrize("rlx", [BGK, TRT, REG], ids=["bgk", "trt", "reg"])
def test_couette_between_ladd_walls_is_exact(rlx):
    """Solid layers at z = 0 (at rest) and z = nz-1 (moving, Ladd's
    bounce-back), periodic x, y: the steady profile is the linear Couette
    profile with the walls half a spacing outside the fluid, for any tau."""
    from jyulb.flow_field import couette_plate_steady_state
    H, U, V = 6, 0.04, -0.01
    lat = (3, 2, H + 2)
    s = _solver(lat, rlx, 0.12, [0, 0, 0], [jo.PERIODIC] * 6)
    ph = s.get_phase()
    ph[...] = 1
    ph[:, :, 0] = 0
    ph[:, :, -1] = 0
    s.get_vel()[:, :, -1, 0] = U
    s.get_vel()[:, :, -1, 1] = V
    s.finalize_init()
    s.evolve(1500)
    dist = np.zeros(H)
    ana = np.zeros((H, 2))
    couette_plate_steady_state(H, 0.0, 0.0, U, V, dist, ana)
    got = s.get_vel()[1, 1, 1:-1, :2]
    np.testing.assert_allclose(got, ana, rtol=0, atol=1e-13 * U)
    np.testing.assert_allclose(s.get_vel()[:, :, 1:-1, 2], 0.0, rtol=0, atol=1e-13 * U)


def test_trt_poiseuille_between_bounce_back_walls_is_exact():
    """Gravity-driven flow between solid rows (periodic x, z), TRT with the
    magic parameter 3/16: the parabola with the walls half a spacing outside
    the fluid is reproduced to round-off. BGK is not exact (tau-dependent
    wall slip) and is checked against the golden vectors instead."""
    from jyulb.flow_field import poise_plate_steady_state_pgrad
    H, nu, g = 8, 0.1, 1e-5
    s = _solver((2, H + 2, 2), TRT, nu, [g, 0, 0], [jo.PERIODIC] * 6)
    ph = s.get_phase()
    ph[...] = 1
    ph[:, 0, :] = 0
    ph[:, -1, :] = 0
    s.finalize_init()
    s.evolve(4000)
    dist = np.zeros(H)
    ana = np.zeros(H)
    poise_plate_steady_state_pgrad(H, nu, -g, dist, ana)
    got = s.get_vel()[0, 1:-1, 0, 0]
    np.testing.assert_allclose(got, ana, rtol=0, atol=1e-12 * ana.max())
    y = np.arange(H) + 0.5  # distance from the lower wall, half a spacing outside the fluid
    np.testing.assert_allclose(ana, g * y * (H - y) / (2 * nu), rtol=1e-14, atol=0)
    # BGK at the same tau leaves a visible slip error
    s = _solver((2, H + 2, 2), BGK, nu, [g, 0, 0], [jo.PERIODIC] * 6)
    s.get_phase()[...] = ph
    s.finalize_init()
    s.evolve(4000)
    assert np.max(np.abs(s.get_vel()[0, 1:-1, 0, 0] - ana)) > 1e-4 * ana.max()
