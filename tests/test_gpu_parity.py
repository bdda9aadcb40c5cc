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


@pytest.mark.parametrize("ci", range(len(CASES)))
def test_gpu_matches_oracle_larger(ci):
    import jyulb_oracle as jo
    bc, rlx, lat = CASES[ci]
    sc = _random_case(100 + ci, bc, rlx, lat, [3, 8, 21])
    print(f"case {ci} lat={lat} bc={bc} rlx={rlx}", flush=True)
    ref = scenarios.run_scenario(jo.OracleSolver, sc)
    got = scenarios.run_scenario(_solver_cls(), sc)
    for k in ref:
        if k in ("nfluid", "sites") or k.startswith("tsteps"):
            continue
        a, b = ref[k], got[k]
        if k.startswith("f_"):
            a, b = _fluid(a, ref["sites"]), _fluid(b, ref["sites"])
        _close(b, a, f"case {ci} {k}")


@pytest.mark.parametrize("layout", [0, 1, 2], ids=["rows", "soa", "soa_padded"])
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


@pytest.mark.parametrize("name", ["trt_periodic", "bgk_zouhe_x", "reg_symwall"])
def test_gpu_fp64_bit_exact_with_xcd_tile_order(name, monkeypatch):
    """Large lattices deal the step kernel's tiles to the XCDs in contiguous
    ranges; forcing that order on small cases must not change a bit."""
    monkeypatch.setenv("LB_XCD_REMAP_CELLS", "0")
    path = os.path.join(os.path.dirname(__file__), "golden", name + ".npz")
    test_gpu_fp64_bit_exact(path, 0)
