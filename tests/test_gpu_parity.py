"""HIP path vs the reference (golden vectors) and vs the CPU oracle.

Runs the jyulb Solver, whose per-site work happens in the fused AA kernel of
libjyulb_hip.so, through the same scenarios the golden vectors were written
with. The FP64 bar is the north-star tolerance, |gpu - ref| <= 1e-12 * max|ref|
per field; the test also records whether the match is bit-exact.
"""
import glob
import json
import os

import numpy as np
import pytest

import scenarios

pytestmark = pytest.mark.gpu

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))
RTOL = 1e-12


def _solver_cls():
    from jyulb.isothermal import Solver
    return Solver


def _fluid(a, sites):
    s = sites.astype(np.int64) + 1
    return a[:, s[0], s[1], s[2]]


def _close(got, ref, what):
    scale = max(float(np.max(np.abs(ref))), 1e-300)
    err = float(np.max(np.abs(got - ref))) if ref.size else 0.0
    assert err <= RTOL * scale, f"{what}: max abs diff {err:.3e} (scale {scale:.3e})"
    return err == 0.0


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_gpu_matches_reference_golden(path):
    z = np.load(path)
    sc = json.loads(str(z["scenario"]))
    print(f"scenario {sc['name']} lat={sc['lat']} rlx={sc['rlx']} bc={sc['bc']} calls={sc['calls']}", flush=True)
    out = scenarios.run_scenario(_solver_cls(), sc, record_f=True)
    assert int(out["nfluid"]) == int(z["nfluid"])
    np.testing.assert_array_equal(out["sites"], z["sites"])
    exact = True
    for k in z.files:
        if k in ("scenario", "nfluid", "sites"):
            continue
        a, b = z[k], out[k]
        if k.startswith("tsteps"):
            assert int(a) == int(b)
            continue
        if k.startswith("f_"):
            a, b = _fluid(a, z["sites"]), _fluid(b, z["sites"])
        exact &= _close(b, a, f"{sc['name']} {k}")
    print(f"bit-exact: {exact}", flush=True)


def _random_case(seed, bc, rlx, lat, steps):
    return dict(name=f"rand{seed}", lat=lat, rlx=rlx, bc=bc, seed=seed, kvisc=0.07,
                g=[1e-5, -1e-5, 2e-5], vel_amp=0.02, den_amp=0.01, solid=0.15, moving=0.4,
                acc_amp=1e-5, calls=steps, mutate=True,
                faces=[])


CASES = [
    ([1, 1, 1, 1, 1, 1], 0, [33, 17, 70]),
    ([4, 4, 1, 1, 0, 4], 1, [20, 31, 66]),
    ([0, 4, 4, 0, 1, 1], 2, [18, 23, 129]),
    ([1, 1, 4, 4, 4, 4], 1, [16, 9, 5]),
]


@pytest.mark.parametrize("layout", ["rows", "sparse"])
@pytest.mark.parametrize("ci", range(len(CASES)))
def test_gpu_matches_oracle_larger(ci, layout):
    import jyulb_oracle as jo
    bc, rlx, lat = CASES[ci]
    sc = _random_case(100 + ci, bc, rlx, lat, [3, 8, 21])
    print(f"case {ci} lat={lat} bc={bc} rlx={rlx} layout={layout}", flush=True)
    ref = scenarios.run_scenario(jo.OracleSolver, sc)
    Solver = _solver_cls()
    got = scenarios.run_scenario(lambda *a: Solver(*a, layout=layout), sc)
    for k in ref:
        if k in ("nfluid", "sites") or k.startswith("tsteps"):
            continue
        a, b = ref[k], got[k]
        if k.startswith("f_"):
            a, b = _fluid(a, ref["sites"]), _fluid(b, ref["sites"])
        _close(b, a, f"case {ci} {k}")


@pytest.mark.parametrize("layout", [0, 1, 2, 3], ids=["rows", "soa", "soa_padded", "sparse"])
@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_gpu_fp64_bit_exact(path, layout):
    """The FP64 kernel keeps the reference's operation order and is built
    without FMA contraction, so it reproduces the reference bit for bit, in
    every population layout."""
    z = np.load(path)
    sc = json.loads(str(z["scenario"]))
    Solver = _solver_cls()

    def mk(*a):
        return Solver(*a, layout=layout)

    out = scenarios.run_scenario(mk, sc, record_f=True)
    for k in z.files:
        if k in ("scenario", "nfluid", "sites") or k.startswith("tsteps"):
            continue
        a, b = z[k], out[k]
        if k.startswith("f_"):
            a, b = _fluid(a, z["sites"]), _fluid(b, z["sites"])
        np.testing.assert_array_equal(b, a, err_msg=f"{sc['name']} {k} layout {layout}")


@pytest.mark.parametrize("layout", [0, 3], ids=["rows", "sparse"])
@pytest.mark.parametrize("name", ["trt_periodic", "bgk_zouhe_x", "reg_symwall"])
def test_gpu_fp64_bit_exact_with_xcd_tile_order(name, layout, monkeypatch):
    """Large lattices deal the step kernel's tiles to the XCDs in contiguous
    ranges; forcing that order on small cases must not change a bit."""
    monkeypatch.setenv("LB_XCD_REMAP_CELLS", "0")
    path = os.path.join(os.path.dirname(__file__), "golden", name + ".npz")
    test_gpu_fp64_bit_exact(path, layout)


@pytest.mark.parametrize("layout", [0, 1, 3], ids=["rows", "soa", "sparse"])
@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_gpu_fp64_bit_exact_with_pull_steps(path, layout, monkeypatch):
    """Pull steps (odd steps that gather like the in-place odd step and store
    into a second array like the even step), forced on for every fixture:
    still the reference's bits, moving walls, Zou-He faces, periodic and
    symmetric faces included."""
    monkeypatch.setenv("LB_SP_PULL", "1")
    test_gpu_fp64_bit_exact(path, layout)


@pytest.mark.parametrize("tile", ["128", "256", "512"])
@pytest.mark.parametrize("name", ["trt_walls", "reg_zouhe_y", "bgk_symmetric"])
def test_gpu_fp64_bit_exact_with_step_tile(name, tile, monkeypatch):
    """The bulk step's workgroup size is chosen per solver (512 threads for
    the dense FP64 kernel); any size gives the same bits, in the dense and
    the sparse layout."""
    monkeypatch.setenv("LB_STEP_TILE", tile)
    path = os.path.join(os.path.dirname(__file__), "golden", name + ".npz")
    test_gpu_fp64_bit_exact(path, 0)
    test_gpu_fp64_bit_exact(path, 3)


@pytest.mark.parametrize("name", ["trt_long", "ex_couette", "ex_poiseuille_grav_sym", "c1_poiseuille"])
def test_gpu_fp64_bit_exact_with_launch_tuning(name, monkeypatch):
    """The first long evolve call times its own steps under each candidate
    launch configuration and keeps the fastest; forced on for these small
    lattices, the FP64 results still equal the reference's bit for bit, in
    the rows layout and in the sparse layout (whose candidates are the tile
    orders)."""
    monkeypatch.setenv("LB_AUTOTUNE", "1")
    monkeypatch.setenv("LB_AUTOTUNE_MIN_CELLS", "0")
    path = os.path.join(os.path.dirname(__file__), "golden", name + ".npz")
    test_gpu_fp64_bit_exact(path, 0)
    test_gpu_fp64_bit_exact(path, 3)


@pytest.mark.parametrize("layout", ["auto", "sparse"])
@pytest.mark.parametrize("precision", ["fp32", "fp32_storage", "fp16"])
@pytest.mark.parametrize("name", ["trt_long", "ex_couette"])
def test_launch_tuning_changes_no_bit(name, precision, layout, monkeypatch):
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", name + ".npz"))
    sc = json.loads(str(z["scenario"]))
    Solver = _solver_cls()
    monkeypatch.setenv("LB_AUTOTUNE_MIN_CELLS", "0")
    outs = []
    for env in ("0", "1"):
        monkeypatch.setenv("LB_AUTOTUNE", env)
        outs.append(scenarios.run_scenario(lambda *a: Solver(*a, precision=precision, layout=layout), sc,
                                           record_f=True))
    for k in outs[0]:
        np.testing.assert_array_equal(outs[1][k], outs[0][k], err_msg=f"{name} {precision} {k}")


# Reduced-precision variants on every reference scenario: all boundary types
# (periodic, walls, symmetric, Zou-He on each face, moving walls, mutations
# between evolve calls) through the shifted-population forms. The density
# (about 1) is compared in absolute terms, the velocity relative to its
# largest magnitude in the reference field. Most scenarios run 0-60 steps
# from random states; the gravity-driven examples run 20 calls of 500 steps
# to a near-steady state, where the error settles at about one storage ulp
# times the slow-mode relaxation time. Measured worst cases over all
# scenarios and calls:
#   FP32 arithmetic: density 1.5e-7, velocity 4.1e-5 (relative);
#   FP32 storage / FP64 arithmetic: 6.5e-9, 5.8e-6;
#   FP16 storage (11-bit mantissa): 3.0e-5, 4.3e-2 (gravity examples).
_RED_TOL = {"fp32": (5e-6, 1e-4), "fp32_storage": (1e-7, 2e-5), "fp16": (5e-3, 1e-1)}


@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp32_storage"])
@pytest.mark.parametrize("path", [p for p in GOLD if "c1_poiseuille" not in p],
                         ids=[os.path.basename(p)[:-4] for p in GOLD if "c1_poiseuille" not in p])
def test_reduced_precision_on_reference_scenarios(path, precision):
    z = np.load(path)
    sc = json.loads(str(z["scenario"]))
    Solver = _solver_cls()

    def mk(*a):
        return Solver(*a, precision=precision)

    out = scenarios.run_scenario(mk, sc, record_f=False)
    s = z["sites"].astype(np.int64) + 1
    tol_den, tol_vel = _RED_TOL[precision]
    worst_den, worst_vel = 0.0, 0.0
    for k in z.files:
        if not (k.startswith("den_") or k.startswith("vel_")):
            continue
        a = z[k][s[0], s[1], s[2]]
        b = out[k][s[0], s[1], s[2]]
        err = float(np.max(np.abs(a - b)))
        if k.startswith("den_"):
            worst_den = max(worst_den, err)
        else:
            worst_vel = max(worst_vel, err / max(float(np.max(np.abs(a))), 1e-300))
    print(f"{sc['name']} {precision}: worst density abs err {worst_den:.2e}, velocity rel err {worst_vel:.2e}",
          flush=True)
    assert worst_den <= tol_den and worst_vel <= tol_vel, (sc["name"], precision, worst_den, worst_vel)
