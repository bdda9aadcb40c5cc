"""Build the HIP engine (libjyulb_hip.so) for gfx950, in-tree.

Usage: python jyu-lb_amd/build.py [--verbose]

The library lands in jyu-lb_amd/jyulb/_lib/ next to the Python package that
loads it. It is linked from five translation units compiled in parallel:
csrc/lb_engine.hip (host code, set-up and reduction kernels) and
csrc/lb_inst.hip once per storage variant (LB_INST_P = 0..3: the fused step,
adjoint and initialisation kernels of that variant). Flags:
  -ffp-contract=off   no FMA contraction: the FP64 kernel keeps the
                      reference's rounding, term by term
  -packed-fp32-ops    disabled (target feature), per the platform rules for
                      this GPU pool
"""
import os
import shutil
import subprocess
import sys
import tempfile

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "jyulb", "_lib", "libjyulb_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]

# storage variants of lb_inst.hip: LB_FP64, LB_FP32_SHIFTED, LB_FP16_STORAGE, LB_FP32_STORAGE
VARIANTS = (0, 1, 2, 3)


def sources():
    inc = os.path.join(os.path.dirname(HERE), "include")
    out = [os.path.join(CSRC, f) for f in os.listdir(CSRC)]
    out += [os.path.join(inc, f) for f in os.listdir(inc)]
    return out


def _units(defines):
    d = ["-D" + x for x in defines]
    units = [("engine", [os.path.join(CSRC, "lb_engine.hip")] + d)]
    for v in VARIANTS:
        units.append((f"inst{v}", [os.path.join(CSRC, "lb_inst.hip"), f"-DLB_INST_P={v}"] + d))
    return units


def build(verbose=False, out=OUT, defines=()):
    os.makedirs(os.path.dirname(out), exist_ok=True)
    work = tempfile.mkdtemp(prefix="jyulb_build_")
    try:
        procs = []
        for name, args in _units(defines):
            obj = os.path.join(work, name + ".o")
            cmd = [HIPCC] + FLAGS + ["-c", "-o", obj] + args
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((name, obj, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                                      text=True)))
        failed = False
        for name, obj, p in procs:
            so, se = p.communicate()
            if p.returncode != 0:
                sys.stderr.write(f"[{name}]\n" + so + se)
                failed = True
            elif verbose and se.strip():
                sys.stderr.write(f"[{name}]\n" + se)
        if failed:
            raise RuntimeError("hipcc failed building libjyulb_hip.so")
        tmp = out + ".tmp"
        cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + [o for _, o, _ in procs]
        if verbose:
            print(" ".join(cmd), flush=True)
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("linking libjyulb_hip.so failed")
        os.replace(tmp, out)
    finally:
        shutil.rmtree(work, ignore_errors=True)
    return out


if __name__ == "__main__":
    print(build(verbose="--verbose" in sys.argv))
