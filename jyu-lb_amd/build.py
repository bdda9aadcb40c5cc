"""Build the HIP engine (libjyulb_hip.so) for gfx950, in-tree.

Usage: python jyu-lb_amd/build.py [--verbose]

The library lands in jyu-lb_amd/jyulb/_lib/ next to the Python package that
loads it. Flags:
  -ffp-contract=off   no FMA contraction: the FP64 kernel keeps the
                      reference's rounding, term by term
  -packed-fp32-ops    disabled (target feature), per the platform rules for
                      this GPU pool
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc", "lb_engine.hip")
OUT = os.path.join(HERE, "jyulb", "_lib", "libjyulb_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def sources():
    csrc = os.path.join(HERE, "csrc")
    inc = os.path.join(os.path.dirname(HERE), "include")
    out = [os.path.join(csrc, f) for f in os.listdir(csrc)]
    out += [os.path.join(inc, f) for f in os.listdir(inc)]
    return out


def build(verbose=False, out=OUT, defines=()):
    os.makedirs(os.path.dirname(out), exist_ok=True)
    tmp = out + ".tmp"
    cmd = [HIPCC] + FLAGS + ["-D" + d for d in defines] + ["-o", tmp, SRC]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libjyulb_hip.so")
    if verbose and r.stderr.strip():
        sys.stderr.write(r.stderr)
    os.replace(tmp, out)
    return out


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
