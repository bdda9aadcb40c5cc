"""Write the golden vectors under tests/golden/ from the reference jyulb solver.

This script runs only in the build container (where /root/reference exists)
and is test infrastructure. It copies the reference's jyulb Cython sources
into a temporary build directory OUTSIDE this repository, compiles them there
with the reference's own compiler flags (setup.py:23-69: C++, -fopenmp -O3),
imports the result and runs the seeded scenarios of scenarios.py. Only the
inputs (the scenario dicts) and the outputs (den, vel, post-collision
populations) are committed, as .npz fixtures; no reference source or build
product enters the repository.

Usage:  OMP_NUM_THREADS=1 python tests/golden/make_golden.py [--only NAME ...]

One OpenMP thread: Cython 3 makes the reference's nogil loops re-acquire the
GIL for exception checks, so more threads only add contention; results do
not depend on the thread count (every site is updated independently).
"""
import argparse
import json
import os
import shutil
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import scenarios  # noqa: E402

REF = "/root/reference/jyulb"
MODULES = ["defs", "D3Q19", "bcs", "flow_field", "isothermal"]


def build_reference(workdir):
    for fn in os.listdir(REF):
        if fn.endswith((".pyx", ".pxd")):
            shutil.copy(os.path.join(REF, fn), workdir)
    from setuptools import Extension
    from setuptools.dist import Distribution
    from Cython.Build import cythonize
    exts = [Extension(m, [os.path.join(workdir, m + ".pyx")],
                      include_dirs=[workdir, np.get_include()],
                      extra_compile_args=["-fopenmp", "-O3"],
                      extra_link_args=["-fopenmp", "-O3"], language="c++")
            for m in MODULES]
    exts = cythonize(exts, include_path=[workdir], quiet=True,
                     compiler_directives={"language_level": 2})
    dist = Distribution({"ext_modules": exts})
    cmd = dist.get_command_obj("build_ext")
    cmd.build_lib = workdir
    cmd.build_temp = os.path.join(workdir, "tmp")
    cmd.ensure_finalized()
    cmd.run()
    sys.path.insert(0, workdir)
    import isothermal  # noqa: F401
    return isothermal


class RefSolver:
    """Thin adapter: the reference Solver plus an unpacked f1 getter."""

    def __init__(self, mod):
        self.mod = mod

    def __call__(self, *a):
        s = self.mod.Solver(*a)
        return _Wrap(s)


class _Wrap:
    def __init__(self, s):
        self._s = s

    def __getattr__(self, k):
        return getattr(self._s, k)

    @property
    def gravity(self):
        return self._s.gravity

    @gravity.setter
    def gravity(self, g):
        self._s.gravity = g

    def post_collision_f(self):
        """f1 (post-collision populations, Bundle C, isothermal.pyx:1532-1704)
        unpacked to (19, NI, NJ, NK) in lattice-vector order."""
        s = self._s
        NI, NJ, NK, NC = s.NI, s.NJ, s.NK, s.NCNT
        f1 = np.asarray(s.f1)
        n = np.arange(NC)
        out = np.empty((19, NC))
        for l in range(19):
            if l < 5:
                idx = 5 * n + l
            elif l < 14:
                idx = 5 * NC + 9 * n + (l - 5)
            else:
                idx = 14 * NC + 5 * n + (l - 14)
            out[l] = f1[idx]
        return out.reshape(19, NI, NJ, NK)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--only", nargs="*")
    args = ap.parse_args()
    work = tempfile.mkdtemp(prefix="jyulb_ref_build_", dir="/tmp")
    mod = build_reference(work)
    mk = RefSolver(mod)
    scs = scenarios.small_scenarios() + scenarios.example_scenarios()
    for sc in scs:
        if args.only and sc["name"] not in args.only:
            continue
        out = scenarios.run_scenario(mk, sc, record_f=not sc["name"].startswith("ex_"))
        out["scenario"] = np.array(json.dumps(sc))
        path = os.path.join(HERE, sc["name"] + ".npz")
        np.savez_compressed(path, **out)
        print(f"wrote {path}: nfluid={int(out['nfluid'])} calls={sc['calls']}", flush=True)
    shutil.rmtree(work, ignore_errors=True)


if __name__ == "__main__":
    main()
