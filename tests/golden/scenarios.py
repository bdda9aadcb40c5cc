"""Seeded scenario definitions shared by the golden-vector writer and the tests.

A scenario is a plain dict (JSON-serialisable) describing one run of the
jyulb isothermal solver:

  lat        interior lattice size (nx, ny, nz)
  rlx        collision id (BGK=0, TRT=1, REGULARIZATION=2)
  kvisc      kinematic viscosity
  g          gravity (3,)
  bc         boundary condition per face (6,), FACE_1X..FACE_Z1
  seed       numpy default_rng seed for the synthetic fields
  solid      fraction of interior cells made SOLID (random)
  moving     fraction of SOLID cells that get a nonzero wall velocity
  vel_amp    amplitude of the random initial velocity field
  den_amp    amplitude of the random initial density perturbation
  acc_amp    amplitude of the random per-cell acceleration field (0 = none)
  calls      list of evolve() step counts, executed in order
  mutate     if true, gravity / acc / solid velocities / densities are changed
             between evolve() calls (the host-array aliasing semantics of
             isothermal.pyx:1465-1527 and the per-call _set_bc_data refresh)
  faces      optional Zou-He face data written into den/vel before
             finalize_init (inlet profile / outlet density), see build_inputs
  den_profile  optional density per index along one axis (fluid sites)
  vel_profile  optional list: one velocity component of a face layer per
             index along one of the layer's axes
  solid_layers, layer_vel  boundary layers made SOLID, and layer velocities

The generator functions below produce the *interior* input arrays that the
reference Solver receives through its aliasing views (get_phase/get_den/
get_vel/get_acc, isothermal.pyx:1465-1512).
"""
import numpy as np

BGK, TRT, REG = 0, 1, 2
WALL, PERIODIC, FIXED_DEN, FIXED_VEL, SYMMETRIC = 0, 1, 2, 3, 4
SOLID, FLUID = 0, 1


def _face_slices(face, shape):
    nx, ny, nz = shape
    return {
        0: (0, slice(None), slice(None)),
        1: (nx - 1, slice(None), slice(None)),
        2: (slice(None), 0, slice(None)),
        3: (slice(None), ny - 1, slice(None)),
        4: (slice(None), slice(None), 0),
        5: (slice(None), slice(None), nz - 1),
    }[face]


def build_inputs(sc):
    """Return (phase, den, vel, acc) interior arrays for a scenario."""
    rng = np.random.default_rng(sc["seed"])
    shape = tuple(sc["lat"])
    phase = np.full(shape, FLUID, dtype=np.uint8)
    solid = sc.get("solid", 0.0)
    if solid > 0:
        phase[rng.random(shape) < solid] = SOLID
    for face, kind in sc.get("solid_layers", []):
        # whole boundary layers made SOLID (walls/plates/lids inside the domain)
        phase[_face_slices(face, shape)] = SOLID
    den = 1.0 + sc.get("den_amp", 0.0) * (rng.random(shape) - 0.5)
    vel = sc.get("vel_amp", 0.0) * (rng.random(shape + (3,)) - 0.5)
    acc = np.zeros(shape + (3,))
    if sc.get("acc_amp", 0.0) > 0:
        acc = sc["acc_amp"] * (rng.random(shape + (3,)) - 0.5)
    # solid cells: zero velocity except the "moving" ones
    sol = phase == SOLID
    vel[sol] = 0.0
    mov = sc.get("moving", 0.0)
    if mov > 0:
        pick = sol & (rng.random(shape) < mov)
        vel[pick] = sc.get("wall_amp", 0.02) * (rng.random((int(pick.sum()), 3)) - 0.5)
    for face, vx, vy, vz in sc.get("layer_vel", []):
        sl = _face_slices(face, shape)
        vel[sl + (0,)] = vx
        vel[sl + (1,)] = vy
        vel[sl + (2,)] = vz
    # Zou-He face data: inlet velocity profile / outlet density written into
    # the face layer of the fields (the data _set_bc_data copies, isothermal.pyx:476-502)
    for fd in sc.get("faces", []):
        sl = _face_slices(fd["face"], shape)
        if "den" in fd:
            den[sl] = fd["den"]
        if "vel" in fd:
            v = np.asarray(fd["vel"], dtype=np.float64)
            prof = 1.0 + 0.3 * (rng.random(den[sl].shape) - 0.5) if fd.get("profile") else 1.0
            for c in range(3):
                vel[sl + (c,)] = v[c] * prof
    # per-layer profiles written through the views, as the example drivers do:
    # a density per index along an axis (examples/poiseuille_pdiff.py:93-97)
    # and one velocity component of a face layer per index along another axis
    # (examples/poiseuille_invel.py:88-98)
    if "den_profile" in sc:
        dp = sc["den_profile"]
        sh = [1, 1, 1]
        sh[dp["axis"]] = -1
        prof = np.broadcast_to(np.asarray(dp["values"], dtype=np.float64).reshape(sh), shape)
        fl = phase != SOLID
        den[fl] = prof[fl]
    for vp in sc.get("vel_profile", []):
        sl = _face_slices(vp["face"], shape)
        layer = vel[sl + (vp["component"],)]
        ax = [a for a in range(3) if a != vp["face"] // 2].index(vp["axis"])
        sh = [1, 1]
        sh[ax] = -1
        layer[...] = np.asarray(vp["values"], dtype=np.float64).reshape(sh)
    return phase, den, vel, acc


def mutation(sc, call_index, phase):
    """Changes applied through the views before evolve call `call_index`>0.

    Returns dict with optional keys: g (3,), acc (interior array), solid_vel
    (interior array; applied at SOLID cells only), den_scale (interior array
    multiplied into den at FLUID cells)."""
    if not sc.get("mutate") or call_index == 0:
        return {}
    rng = np.random.default_rng(sc["seed"] * 1000 + call_index)
    shape = tuple(sc["lat"])
    out = {}
    out["g"] = list(np.asarray(sc["g"]) * (1.0 + 0.5 * call_index))
    out["acc"] = 1e-5 * (rng.random(shape + (3,)) - 0.5)
    sv = 0.01 * (rng.random(shape + (3,)) - 0.5)
    sv[phase != SOLID] = 0.0
    out["solid_vel"] = sv
    out["den_scale"] = 1.0 + 1e-3 * (rng.random(shape) - 0.5)
    return out


def _sc(name, **kw):
    d = dict(name=name, g=[0.0, 0.0, 0.0], kvisc=0.1, calls=[1, 2, 4], seed=1)
    d.update(kw)
    return d


def small_scenarios():
    """Small lattices covering every collision x boundary combination."""
    out = []
    P, W, S, FD, FV = PERIODIC, WALL, SYMMETRIC, FIXED_DEN, FIXED_VEL
    for rlx, rn in ((BGK, "bgk"), (TRT, "trt"), (REG, "reg")):
        s = 10 + rlx
        out.append(_sc(f"{rn}_periodic", lat=[6, 5, 4], rlx=rlx, bc=[P] * 6, seed=s,
                       g=[1e-5, -2e-5, 3e-5], vel_amp=0.02, den_amp=0.01, solid=0.12,
                       moving=0.5, acc_amp=2e-5, kvisc=0.08))
        out.append(_sc(f"{rn}_walls", lat=[7, 6, 5], rlx=rlx, bc=[P, P, W, W, W, W], seed=s + 100,
                       g=[2e-5, 0.0, -1e-5], vel_amp=0.01, den_amp=0.01, solid=0.1,
                       moving=0.6, kvisc=0.05, mutate=True))
        out.append(_sc(f"{rn}_symmetric", lat=[6, 7, 5], rlx=rlx, bc=[S, S, P, P, W, S], seed=s + 200,
                       g=[0.0, 1e-5, 2e-5], vel_amp=0.02, den_amp=0.02, solid=0.1,
                       moving=0.5, kvisc=0.12, mutate=True))
        out.append(_sc(f"{rn}_symwall", lat=[5, 6, 7], rlx=rlx, bc=[W, S, S, W, P, P], seed=s + 250,
                       g=[1e-5, 1e-5, 0.0], vel_amp=0.02, den_amp=0.02, solid=0.08,
                       moving=0.5, kvisc=0.1))
        out.append(_sc(f"{rn}_zouhe_x", lat=[8, 6, 5], rlx=rlx, bc=[FV, FD, W, W, P, P], seed=s + 300,
                       g=[1e-6, 0.0, 0.0], vel_amp=0.005, den_amp=0.005, solid=0.05,
                       kvisc=0.1, mutate=True,
                       faces=[dict(face=0, vel=[0.02, 0.003, -0.002], profile=True),
                              dict(face=1, den=1.001)]))
        out.append(_sc(f"{rn}_zouhe_y", lat=[5, 8, 6], rlx=rlx, bc=[S, W, FD, FD, P, P], seed=s + 400,
                       g=[0.0, 0.0, 1e-6], vel_amp=0.005, den_amp=0.005, solid=0.05,
                       kvisc=0.09,
                       faces=[dict(face=2, den=1.003, vel=[0.001, 0.0, 0.002]),
                              dict(face=3, den=0.998)]))
        # FIXED_DEN at 1Z with FIXED_VEL at Z1: exercises the face-Z1 refresh
        # condition of _set_bc_data (isothermal.pyx:500)
        out.append(_sc(f"{rn}_zouhe_z", lat=[6, 5, 8], rlx=rlx, bc=[P, P, W, S, FD, FV], seed=s + 500,
                       g=[0.0, 0.0, 0.0], vel_amp=0.005, den_amp=0.005, solid=0.05,
                       kvisc=0.07, mutate=True,
                       faces=[dict(face=4, den=1.002, vel=[0.002, 0.0, 0.0]),
                              dict(face=5, vel=[0.0, 0.001, -0.01], profile=True)]))
        out.append(_sc(f"{rn}_zouhe_neg", lat=[7, 5, 6], rlx=rlx, bc=[FD, FV, P, P, S, W], seed=s + 600,
                       g=[0.0, 1e-6, 0.0], vel_amp=0.005, den_amp=0.004, solid=0.05,
                       kvisc=0.11,
                       faces=[dict(face=0, den=0.999),
                              dict(face=1, vel=[-0.015, 0.001, 0.0], profile=True)]))
    # single-layer periodic direction (PERIOD_W = 1 as in examples/lid_driven_cavity.py:42)
    out.append(_sc("reg_thin_periodic", lat=[6, 1, 7], rlx=REG, bc=[W, W, P, P, W, W], seed=77,
                   g=[0.0, 0.0, 0.0], vel_amp=0.01, den_amp=0.01, solid=0.0, kvisc=0.05,
                   solid_layers=[(5, 1)], layer_vel=[(5, 0.05, 0.0, 0.0)]))
    # longer run
    out.append(_sc("trt_long", lat=[9, 8, 7], rlx=TRT, bc=[P, P, W, S, P, P], seed=91,
                   g=[1e-5, 0.0, 1e-5], vel_amp=0.01, den_amp=0.01, solid=0.1, moving=0.3,
                   kvisc=0.03, calls=[25, 25]))
    return out


def example_scenarios():
    """Re-expressions of the reference's example drivers (examples/*.py)."""
    X, Y, Z = 0, 1, 2
    P, W, S, FD, FV = PERIODIC, WALL, SYMMETRIC, FIXED_DEN, FIXED_VEL
    out = []
    # examples/poiseuille_grav.py:31-66: plates at faces 1X/X1 (WALL), flow in Y
    out.append(dict(name="ex_poiseuille_grav", lat=[11, 10, 5], rlx=BGK, kvisc=0.2 / 6.0,
                    g=[0.0, 1e-5, 0.0], bc=[W, W, P, P, P, P], seed=0, calls=[500] * 20))
    # examples/poiseuille_grav_sym.py: SYMMETRIC centre line at X1, TRT, flow in Z
    out.append(dict(name="ex_poiseuille_grav_sym", lat=[6, 4, 10], rlx=TRT, kvisc=0.2 / 6.0,
                    g=[0.0, 0.0, 1e-5], bc=[W, S, P, P, P, P], seed=0, calls=[500] * 20))
    # examples/couette.py:30-95: moving solid plates (Ladd), BGK
    out.append(dict(name="ex_couette", lat=[12, 5, 4], rlx=BGK, kvisc=0.5 / 6.0, g=[0.0, 0.0, 0.0],
                    bc=[P] * 6, seed=0, calls=[500] * 20,
                    solid_layers=[(0, 1), (1, 1)],
                    layer_vel=[(0, 0.0, 0.0, 0.0), (1, 0.0, 1e-5, 2e-5)]))
    # examples/poiseuille_pdiff.py:31-101 and poiseuille_pdiff_sym.py: flow in z
    # driven by FIXED_DEN faces at 1Z/Z1, plates at 1Y (WALL) and Y1 (WALL, or
    # SYMMETRIC for the half channel), x periodic; the initial density is the
    # linear pressure profile, computed in the example's operation order
    for name, H, rlx, kvisc, y1 in (("ex_poiseuille_pdiff", 11, BGK, 2.0 / 6.0, W),
                                    ("ex_poiseuille_pdiff_sym", 6, TRT, 0.2 / 6.0, S)):
        L, Wd = 10, 5
        pgrad = 1.0 * -1e-5
        pdiff = (L - 1.0) * pgrad
        pin = (1.0 / 3.0) * 1.0 + (-0.5 * pdiff)
        out.append(dict(name=name, lat=[Wd, H, L], rlx=rlx, kvisc=kvisc, g=[0.0, 0.0, 0.0],
                        bc=[P, P, W, y1, FD, FD], seed=0, calls=[500] * 20,
                        den_profile=dict(axis=Z, values=[3.0 * (pin + l * pgrad) for l in range(L)])))
    # examples/poiseuille_invel.py:31-98 and poiseuille_invel_sym.py: FIXED_VEL
    # inlet at 1Z carrying the parabola of peak 1e-3 (poise_plate_steady_state_umax
    # over H, or over 2H for the half channel), FIXED_DEN outlet at Z1
    for name, H, rlx, kvisc, y1, Hp in (("ex_poiseuille_invel", 11, BGK, 2.0 / 6.0, W, 11),
                                        ("ex_poiseuille_invel_sym", 6, TRT, 0.2 / 6.0, S, 12)):
        L, Wd, umax = 10, 5, 1e-3
        ana = []
        for x in range(Hp):
            rel = ((0.5 + x) - 0.5 * Hp) / (0.5 * Hp)
            ana.append(umax * (1.0 - rel * rel))
        out.append(dict(name=name, lat=[Wd, H, L], rlx=rlx, kvisc=kvisc, g=[0.0, 0.0, 0.0],
                        bc=[P, P, W, y1, FV, FD], seed=0, calls=[500] * 20,
                        vel_profile=[dict(face=4, component=Z, axis=Y, values=ana[:H])]))
    # examples/lid_driven_cavity.py:40-107: regularized, 102 x 1 x 102 with all
    # faces periodic (PERIOD_W = 1 in y), solid layers at z = 0 and z = H+1
    # (the lid, moving at ux = 0.05) and at x = 0 and x = 101; two of the
    # example's 25 calls of 5000 steps
    out.append(dict(name="ex_lid_driven_cavity", lat=[102, 1, 102], rlx=REG, kvisc=0.05 / 6.0,
                    g=[0.0, 0.0, 0.0], bc=[P] * 6, seed=0, calls=[5000, 5000],
                    solid_layers=[(4, 1), (5, 1), (0, 1), (1, 1)], layer_vel=[(5, 0.05, 0.0, 0.0)]))
    # Plain gravity-driven Poiseuille channel, D3Q19 BGK, FP64: 32x34x8, SOLID
    # rows j=0 and j=33, periodic everywhere, tau=0.8, gx=1e-6, 10000 steps
    out.append(dict(name="c1_poiseuille", lat=[32, 34, 8], rlx=BGK, kvisc=0.1, g=[1e-6, 0.0, 0.0],
                    bc=[P] * 6, seed=0, calls=[10000], solid_layers=[(2, 1), (3, 1)]))
    return out


def run_scenario(Solver, sc, record_f=True):
    """Drive a Solver class with the jyulb API through a scenario.

    Follows the usage pattern of the reference's example drivers
    (examples/couette.py:73-107): construct, write phase/den/vel/acc through
    the aliasing views, finalize_init, then a sequence of evolve() calls with
    optional mutations between them. Returns a dict of numpy arrays:
    nfluid, sites, and per call den/vel (full halo-padded arrays) and, if the
    solver exposes it, the post-collision populations `f_post` (19,NI,NJ,NK)."""
    lat = np.array(sc["lat"], dtype=np.int32)
    g = np.array(sc["g"], dtype=np.float64)
    bc = np.array(sc["bc"], dtype=np.int32)
    solver = Solver(lat, sc["rlx"], sc["kvisc"], g, bc)
    phase, den, vel, acc = build_inputs(sc)
    pv = np.asarray(solver.get_phase())
    dv = np.asarray(solver.get_den())
    vv = np.asarray(solver.get_vel())
    av = np.asarray(solver.get_acc())
    pv[...] = phase
    dv[...] = den
    vv[...] = vel
    av[...] = acc
    out = {}
    out["nfluid"] = np.array(solver.finalize_init())
    out["sites"] = np.array(solver.get_fluid_lat_sites())
    for c, steps in enumerate(sc["calls"]):
        m = mutation(sc, c, phase)
        if "g" in m:
            solver.gravity = np.array(m["g"])
        if "acc" in m:
            av[...] = m["acc"]
        if "solid_vel" in m:
            sol = phase == SOLID
            vv[sol] = m["solid_vel"][sol]
        if "den_scale" in m:
            flu = phase != SOLID
            dv[flu] = dv[flu] * m["den_scale"][flu]
        out[f"tsteps_{c}"] = np.array(solver.evolve(steps))
        out[f"den_{c}"] = np.array(solver.den, copy=True)
        out[f"vel_{c}"] = np.array(solver.vel, copy=True)
        if record_f and hasattr(solver, "post_collision_f"):
            out[f"f_{c}"] = solver.post_collision_f()
    return out
