"""The CPU oracle against the golden vectors written by the reference solver.

The fixtures (tests/golden/*.npz, written by tests/golden/make_golden.py from
the reference jyulb package) hold den/vel (halo-padded) and the post-collision
populations after each evolve() call. The oracle restates the reference's
arithmetic in the same order, so the bar is bit-exact equality.
"""
import glob
import json
import os

import numpy as np
import pytest

import jyulb_oracle as jo
import scenarios

GOLD = sorted(glob.glob(os.path.join(os.path.dirname(__file__), "golden", "*.npz")))


def _fluid(a, sites):
    s = sites.astype(np.int64) + 1
    return a[:, s[0], s[1], s[2]]


@pytest.mark.parametrize("path", GOLD, ids=[os.path.basename(p)[:-4] for p in GOLD])
def test_oracle_matches_reference_bit_exact(path):
    z = np.load(path)
    sc = json.loads(str(z["scenario"]))
    out = scenarios.run_scenario(jo.OracleSolver, sc)
    assert int(out["nfluid"]) == int(z["nfluid"])
    np.testing.assert_array_equal(out["sites"], z["sites"])
    for k in z.files:
        if k in ("scenario", "nfluid", "sites"):
            continue
        a, b = z[k], out[k]
        if k.startswith("f_"):
            a, b = _fluid(a, z["sites"]), _fluid(b, z["sites"])
        np.testing.assert_array_equal(b, a, err_msg=f"{sc['name']} {k}")
