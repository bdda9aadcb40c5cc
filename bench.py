"""Throughput benchmark of the D3Q19 evolve path on one MI355X.

Headline workload: the porous medium of 768 x 384 x 384 cells (113M cells,
90.6M fluid sites): seeded random spheres of radius 6..12 up to 20 % solid
(jyulb.geometry.porous_spheres, seed 5), D3Q19 TRT (magic 3/16), tau = 0.6,
gravity gx = 1e-6, periodic on every face. A "step" is one fused
stream+collide pass over the lattice (an AA step in place, or a pull step into
the second population array), with the populations resident in HBM, timed with
HIP events on the stream the kernels run on. `value` is FP64 MLUPS
(fluid-site updates) in the Solver's default layout ("auto", which picks the
sparse fluid-only layout for this medium); FP32 on shifted populations, FP16
storage / FP32 arithmetic and FP32 storage / FP64 arithmetic are in
`variants`, the dense rows layout in `layout_compare`.

Beside it:
  secondary_workload  lid-driven cavity 256^3 (TRT, lid ux = 0.05), every precision
  precision_workload  Zou-He duct 512x128x128 (regularized, 5000 steps) in every
                      precision, velocity error against FP64
  adjoint             porous channel 128x64x64 (1000 steps, checkpoint every 50,
                      and one segment), and the 768x384x384 medium at checkpoint
                      interval 25
  solver_evolve       Solver.evolve(500), host fields in and out, against the
                      engine's own evolve(500), on the cavity and the medium
  cpu_baseline        the FP64 CPU oracle (a C/OpenMP restatement of the
                      reference's step) on the same medium for a few steps

Prints one JSON line (see --help).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "jyu-lb_amd"))

METRIC = "MLUPS D3Q19 AA fp64 vs fp32-shifted; adjoint/forward step time ratio"
PEAK_HBM_GBS = 8000.0
BYTES_PER_SITE = {"fp64": 2 * 19 * 8, "fp32": 2 * 19 * 4, "fp16": 2 * 19 * 2, "fp32_storage": 2 * 19 * 4}
C5_LAT = (768, 384, 384)
C5_NAME = ("porous medium 768x384x384 (113M cells; spheres r 6..12, 20% solid, seed 5), D3Q19 TRT "
           "(magic 3/16), tau=0.6, gx=1e-6, periodic")


def step_kernel(prec, layout, rlx="1", pull=False):
    """rocprof name pattern of the bulk step kernel (both parities; with the
    sparse layout's pull steps, the one kernel of every step after the
    first of a call)."""
    p = {"fp64": 0, "fp32": 1, "fp16": 2, "fp32_storage": 3}[prec]
    if layout == "sparse":
        return f"k_step_sp_pull<{p},{rlx},false>" if pull else f"k_step_sp<{p},{rlx},*,false>"
    if prec == "fp16" and rlx == "1":
        return "k_step_h<1,true,true>" if pull else "k_step_h<1,*>"
    return f"k_step<{p},{rlx},true,true>" if pull else f"k_step<{p},{rlx},*>"


# ---------------------------------------------------------------------------
# workloads
# ---------------------------------------------------------------------------
def cavity_inputs(n):
    from jyulb.defs import FLUID, SOLID
    phase = np.full((n, n, n), FLUID, dtype=np.uint8)
    phase[:, :, n - 1] = SOLID
    vel = np.zeros((n, n, n, 3))
    vel[:, :, n - 1, 0] = 0.05
    den = np.ones((n, n, n))
    return phase, den, vel


def make_cavity(n, precision, layout="auto"):
    """Lid-driven cavity: TRT, tau = 0.6, bounce-back walls on every face, the
    lid a SOLID layer at the top z plane moving with ux = 0.05."""
    from jyulb.isothermal import Solver
    from jyulb.defs import TRT, WALL
    kvisc = (0.6 - 0.5) / 3.0
    s = Solver(np.array([n, n, n], dtype=np.int32), TRT, kvisc, np.zeros(3),
               np.full(6, WALL, dtype=np.int32), precision=precision, layout=layout)
    phase, den, vel = cavity_inputs(n)
    s.get_phase()[...] = phase
    s.get_den()[...] = den
    s.get_vel()[...] = vel
    nfl = s.finalize_init()
    return s, nfl


def duct(precision, lat=(512, 128, 128), peak=0.04, tau=0.55, layout="auto"):
    """Precision workload: Zou-He duct, regularized collision, tau = 0.55,
    parabolic inlet u_x = peak (1 - (2 eta_y)^2)(1 - (2 eta_z)^2) with
    eta = (idx + 0.5) / n - 0.5 at 1X (FIXED_VEL), outlet density 1 at X1
    (FIXED_DEN), bounce-back walls on both y and both z faces."""
    from jyulb.isothermal import Solver
    from jyulb.defs import REGULARIZATION, WALL, FIXED_DEN, FIXED_VEL
    nx, ny, nz = lat
    s = Solver(np.array(lat, dtype=np.int32), REGULARIZATION, (tau - 0.5) / 3.0, np.zeros(3),
               np.array([FIXED_VEL, FIXED_DEN, WALL, WALL, WALL, WALL], dtype=np.int32), precision=precision,
               layout=layout)
    s.get_phase()[...] = 1
    ey = (np.arange(ny) + 0.5) / ny - 0.5
    ez = (np.arange(nz) + 0.5) / nz - 0.5
    s.get_vel()[0, :, :, 0] = peak * (1 - (2 * ey[:, None]) ** 2) * (1 - (2 * ez[None, :]) ** 2)
    nfl = s.finalize_init()
    return s, nfl


def porous_channel(precision, lat=(128, 64, 64), seed=4, peak=0.02, layout="auto"):
    """Adjoint workload: porous channel of seeded spheres (radius 4..8, 10 %
    solid, none in 8-layer buffers behind the inlet and outlet faces), BGK
    tau = 0.7, gravity gx = 1e-5, parabolic inlet velocity face (peak 0.02) at
    1X, outlet density 1 at X1, bounce-back walls in y and z."""
    from jyulb.isothermal import Solver
    from jyulb.geometry import porous_spheres
    from jyulb.defs import BGK, WALL, FIXED_DEN, FIXED_VEL, FACE_1X, FACE_X1
    nx, ny, nz = lat
    s = Solver(np.array(lat, dtype=np.int32), BGK, 0.2 / 3.0, np.array([1e-5, 0.0, 0.0]),
               np.array([FIXED_VEL, FIXED_DEN, WALL, WALL, WALL, WALL], dtype=np.int32), precision=precision,
               layout=layout)
    solid = porous_spheres(lat, 4, 8, 0.10, seed, open_faces=(FACE_1X, FACE_X1), open_depth=8)
    s.get_phase()[...] = np.where(solid, 0, 1)
    ey = (np.arange(ny) + 0.5) / ny - 0.5
    ez = (np.arange(nz) + 0.5) / nz - 0.5
    s.get_vel()[0, :, :, 0] = peak * (1 - (2 * ey[:, None]) ** 2) * (1 - (2 * ez[None, :]) ** 2)
    nfl = s.finalize_init()
    return s, nfl


def porous_medium(precision, lat=C5_LAT, seed=5, layout="auto", radius=(6, 12), solid=None):
    """Throughput workload: seeded random-sphere porous medium (radius 6..12,
    20 % solid), TRT tau = 0.6, gravity gx = 1e-6, periodic on all faces.
    `solid` reuses a generated geometry."""
    from jyulb.isothermal import Solver
    from jyulb.geometry import porous_spheres
    from jyulb.defs import TRT, PERIODIC
    s = Solver(np.array(lat, dtype=np.int32), TRT, (0.6 - 0.5) / 3.0, np.array([1e-6, 0.0, 0.0]),
               np.full(6, PERIODIC, dtype=np.int32), precision=precision, layout=layout)
    if solid is None:
        solid = porous_spheres(lat, radius[0], radius[1], 0.20, seed)
    s.get_phase()[...] = np.where(solid, 0, 1)
    nfl = s.finalize_init()
    return s, nfl


# ---------------------------------------------------------------------------
# timing
# ---------------------------------------------------------------------------
def time_device(solver, steps, warmup):
    """HIP events around `steps` engine steps on torch's current stream, which
    the engine is handed (the kernels run on it), after `warmup` steps that
    also hold the launch tuning. Returns (seconds, wall seconds)."""
    import torch
    eng = solver.engine
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    # what Solver.evolve does on entry: Zou-He face data, gravity, host fields
    solver._set_bc_data()
    solver._load_inputs()
    eng.evolve(warmup)
    eng.synchronize()
    torch.cuda.synchronize()
    t0 = torch.cuda.Event(enable_timing=True)
    t1 = torch.cuda.Event(enable_timing=True)
    w0 = time.perf_counter()
    t0.record()
    eng.evolve(steps)
    t1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - w0
    return t0.elapsed_time(t1) / 1e3, wall


def copy_ceiling():
    """The practical HBM rate of a pure streaming kernel on this GPU: the
    engine's device copy (lb_stream_copy, one 8-byte element per thread) of
    2 GiB, read + write bytes over its time (HIP events, 10 copies)."""
    import torch
    from jyulb import _native
    a = torch.empty(2 ** 28, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    a.fill_(1.0)
    st = torch.cuda.current_stream().cuda_stream
    _native.stream_copy(b.data_ptr(), a.data_ptr(), 2 ** 31, st)
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    n = 10
    for _ in range(n):
        _native.stream_copy(b.data_ptr(), a.data_ptr(), 2 ** 31, st)
    ev[1].record()
    torch.cuda.synchronize()
    t = ev[0].elapsed_time(ev[1]) / 1e3 / n
    assert torch.equal(a, b)
    del a, b
    return round(2 * 2 ** 31 / t / 1e9, 1)


def headline(precs, steps, warmup, layout, solid):
    """The porous medium in each precision: MLUPS over the fluid sites and
    over all cells, achieved algorithmic bandwidth, and the velocity field
    after the run (for the errors against FP64)."""
    import torch
    res, fields = {}, {}
    ncell = int(np.prod(C5_LAT))
    for prec in precs:
        s, nfl = porous_medium(prec, layout=layout, solid=solid)
        lay = s.engine.layout()
        print(f"[bench] {prec} {lay}: {C5_LAT} medium, {nfl} fluid sites; warmup {warmup}, timing {steps} steps",
              file=sys.stderr, flush=True)
        t, wall = time_device(s, steps, warmup)
        fields[prec] = s.engine.download_den_vel()[1]
        gbs = BYTES_PER_SITE[prec] * nfl * steps / t / 1e9
        res[prec] = dict(mlups=round(nfl * steps / t / 1e6, 1), mlups_cells=round(ncell * steps / t / 1e6, 1),
                         ms_per_step=round(1e3 * t / steps, 5), achieved_gbs=round(gbs, 1),
                         frac=round(gbs / PEAK_HBM_GBS, 4), layout=lay,
                         kernel=step_kernel(prec, lay, pull=s.engine.pull_steps()), pull=s.engine.pull_steps(),
                         nfluid=nfl, device_bytes=int(s.engine.info()["device_bytes"]),
                         fp16_scale=s.engine.info()["fp16_scale"])
        if lay == "sparse":
            res[prec]["link_codes"] = s.engine.sparse_codes()
        print(f"[bench] {prec} {lay}: {res[prec]['mlups']} MLUPS, {res[prec]['ms_per_step']} ms/step, "
              f"{gbs:.0f} GB/s algorithmic", file=sys.stderr, flush=True)
        del s
        torch.cuda.synchronize()
    errors = {}
    if "fp64" in fields:
        r = fields["fp64"][~solid]
        for p, v in fields.items():
            if p == "fp64":
                continue
            d = v[~solid] - r
            errors[p] = dict(vel_rel_l2=float(np.sqrt(np.sum(d * d) / np.sum(r * r))),
                             vel_rel_linf=float(np.max(np.abs(d)) / np.max(np.abs(r))))
    return res, errors


def cavity_workload(n, steps, warmup, precs):
    """The cavity in each precision. The warm-up call is long enough (60
    steps at least) to hold the launch tuning of a dense lattice, which times
    up to 8 launch candidates and the pull steps on the call's own steps."""
    import torch
    out = {}
    for prec in precs:
        s, nfl = make_cavity(n, prec)
        t, _ = time_device(s, steps, max(warmup, 60))
        lay = s.engine.layout()
        gbs = BYTES_PER_SITE[prec] * nfl * steps / t / 1e9
        out[prec] = dict(mlups=round(nfl * steps / t / 1e6, 1), ms_per_step=round(1e3 * t / steps, 5),
                         achieved_gbs=round(gbs, 1), frac=round(gbs / PEAK_HBM_GBS, 4), layout=lay,
                         kernel=step_kernel(prec, lay, pull=s.engine.pull_steps()), pull=s.engine.pull_steps())
        print(f"[bench] cavity {n}^3 {prec} {lay}: {out[prec]['mlups']} MLUPS", file=sys.stderr, flush=True)
        del s
        torch.cuda.synchronize()
    return dict(workload=f"lid-driven cavity {n}^3, D3Q19 TRT (magic 3/16), tau=0.6, lid ux=0.05, "
                         "bounce-back walls", steps=steps, fluid_sites=n * n * (n - 1), variants=out)


def duct_precision(steps, precs):
    """The duct in each precision over `steps` steps: MLUPS (HIP events over
    the whole run, one evolve call) and the velocity error against FP64."""
    out, fields = {}, {}
    for prec in precs:
        s, nfl = duct(prec)
        print(f"[bench] duct {prec}: {steps} steps", file=sys.stderr, flush=True)
        t, _ = time_device(s, steps, 0)
        fields[prec] = np.array(s.engine.download_den_vel()[1])
        out[prec] = dict(mlups=round(nfl * steps / t / 1e6, 1), layout=s.engine.layout())
        del s
    ref = fields.get("fp64")
    if ref is not None:
        r = ref.reshape(-1, 3)
        for prec in precs:
            if prec == "fp64":
                continue
            d = fields[prec].reshape(-1, 3) - r
            out[prec]["vel_rel_l2"] = float(np.sqrt(np.sum(d * d) / np.sum(r * r)))
            out[prec]["vel_rel_linf"] = float(np.max(np.abs(d)) / np.max(np.abs(r)))
    return dict(workload="Zou-He duct 512x128x128, regularized, tau=0.55, inlet parabola peak 0.04, "
                         "outlet rho=1, bounce-back side walls", steps=steps, variants=out)


def time_adjoint(make, precision, steps, interval, grad_face=True):
    """Forward evolve vs the adjoint (forward pass with checkpoints, then the
    reverse pass with per-segment recomputation) over the same steps, all on
    torch's current stream, timed with HIP events. The forward pass records
    (rho, u*) over its last segment, so with interval >= steps the reverse
    pass recomputes nothing."""
    import torch
    from jyulb.defs import FACE_1X, FACE_X1, X
    from jyulb.adjoint import _prepare
    s, nfl = make(precision)
    eng = s.engine
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    _prepare(s)
    eng.snapshot()
    ga = torch.empty(tuple(s.get_acc().shape), dtype=torch.float64, device="cuda")
    gface = FACE_1X if grad_face else -1
    gf = torch.empty(eng.face_shape(FACE_1X) + (3,), dtype=torch.float64, device="cuda") if grad_face else None
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def run():
        eng.restore()
        torch.cuda.synchronize()
        ev[0].record()
        eng.evolve(steps)
        ev[1].record()
        eng.restore()
        J = eng.adj_forward(steps, interval, FACE_X1, X)
        ev[2].record()
        eng.adj_backward_device(gface, ga.data_ptr(), gf.data_ptr() if gf is not None else 0)
        ev[3].record()
        torch.cuda.synchronize()
        return J

    run()  # allocates checkpoints and adjoint buffers
    J = run()
    t_fwd = ev[0].elapsed_time(ev[1]) / 1e3
    t_af = ev[1].elapsed_time(ev[2]) / 1e3
    t_ab = ev[2].elapsed_time(ev[3]) / 1e3
    out = dict(precision=precision, layout=eng.layout(), steps=steps, checkpoint_interval=interval,
               fluid_sites=nfl, J=J, device_bytes=eng.info()["device_bytes"],
               forward_ms_per_step=round(1e3 * t_fwd / steps, 5),
               adjoint_forward_ms_per_step=round(1e3 * t_af / steps, 5),
               adjoint_reverse_ms_per_step=round(1e3 * t_ab / steps, 5),
               adjoint_over_forward=round(t_ab / t_fwd, 3),
               gradient_over_forward=round((t_af + t_ab) / t_fwd, 3))
    del s
    return out


def solver_evolve(make, steps):
    """Solver.evolve(steps) end to end (host fields read before and written
    after the steps, as the reference's API does) against engine.evolve(steps)
    on the same state, both wall-clock with the device idle before and after."""
    import torch
    s, nfl = make()
    eng = s.engine
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    s.evolve(100)  # first call: launch tuning, buffers
    torch.cuda.synchronize()
    w = time.perf_counter()
    eng.evolve(steps)
    eng.synchronize()
    t_eng = time.perf_counter() - w
    w = time.perf_counter()
    s.evolve(steps)
    t_sol = time.perf_counter() - w
    del s
    return dict(steps=steps, fluid_sites=nfl, engine_s=round(t_eng, 4), solver_s=round(t_sol, 4),
                solver_over_engine=round(t_sol / t_eng, 4))


def cpu_baseline(steps, solid):
    """FP64 CPU oracle on the same porous medium (full lattice) for `steps`
    steps after its initialisation, on all host threads. If the host lacks
    the memory (two population arrays of 17.4 GB each plus the fields), the
    largest lattice of half the length in x that fits is timed instead."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import jyulb_oracle as jo
    from jyulb.defs import TRT, PERIODIC
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    lat = list(C5_LAT)
    avail = None
    try:
        for line in open("/proc/meminfo"):
            if line.startswith("MemAvailable:"):
                avail = int(line.split()[1]) * 1024
    except OSError:
        pass
    note = ""
    while avail is not None and (lat[0] + 2) * (lat[1] + 2) * (lat[2] + 2) * (2 * 19 * 8 + 57) * 1.3 > avail:
        lat[0] //= 2
        note = f"; host memory ({avail / 2**30:.0f} GiB available) holds at most {lat[0]} of 768 x-layers"
    sol = solid[:lat[0]]
    s = jo.OracleSolver(np.array(lat), TRT, (0.6 - 0.5) / 3.0, np.array([1e-6, 0.0, 0.0]),
                        np.full(6, PERIODIC))
    s.get_phase()[...] = np.where(sol, 0, 1)
    nfl = s.finalize_init()
    t = time.perf_counter()
    s.evolve(steps)
    dt = time.perf_counter() - t
    del s
    return dict(value=round(nfl * steps / dt / 1e6, 2), unit="MLUPS", cores=cores, kind="port",
                sample=f"FP64 CPU oracle (C/OpenMP restatement of the reference step) on the porous medium "
                       f"{lat[0]}x{lat[1]}x{lat[2]} ({nfl} fluid sites), {steps} steps after init, {cores} "
                       f"threads, evolve() only{note}")


def load_traffic(prec, layout):
    path = os.path.join(ROOT, "profiles", "k_step_porous768_traffic.json")
    try:
        d = json.load(open(path))
        return d.get(f"{prec}:{layout}", {}).get("hbm_bytes_per_launch")
    except (OSError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100, help="timed steps of the headline workload")
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--precisions", default="fp64,fp32,fp16,fp32_storage")
    ap.add_argument("--layout", default="auto", help="population layout of the headline workload")
    ap.add_argument("--no-layout-compare", action="store_true", help="skip the FP64 rows-layout comparison")
    ap.add_argument("--cavity-n", type=int, default=256)
    ap.add_argument("--cavity-steps", type=int, default=200, help="0: skip the cavity")
    ap.add_argument("--duct-steps", type=int, default=5000, help="steps of the precision workload (0: skip)")
    ap.add_argument("--adj-steps", type=int, default=1000, help="steps of the adjoint channel (0: skip)")
    ap.add_argument("--adj-interval", type=int, default=50)
    ap.add_argument("--adj768-steps", type=int, default=50,
                    help="steps of the adjoint on the 768x384x384 medium, checkpoint interval 25 (0: skip)")
    ap.add_argument("--solver-steps", type=int, default=500, help="steps of the Solver.evolve leg (0: skip)")
    ap.add_argument("--cpu-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: runs on one GPU only; multi-GPU runs are not supported (the path would "
              "shard as 1-D slabs along x, one process per GPU, exchanging the 5 populations that "
              "cross each slab face per step).", file=sys.stderr)
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print("bench.py: no GPU available", file=sys.stderr)
        sys.exit(1)
    from jyulb.geometry import porous_spheres
    ceiling = copy_ceiling()
    print(f"[bench] 2 GiB device copy: {ceiling} GB/s", file=sys.stderr, flush=True)
    precs = [p for p in args.precisions.split(",") if p]
    solid = porous_spheres(C5_LAT, 6, 12, 0.20, 5)
    res, errors = headline(precs, args.steps, args.warmup, args.layout, solid)
    for r in res.values():
        r["frac_of_copy"] = round(r["achieved_gbs"] / ceiling, 4)
    compare = None
    if not args.no_layout_compare and "fp64" in res and res["fp64"]["layout"] != "rows":
        rr, _ = headline(["fp64"], args.steps, args.warmup, "rows", solid)
        compare = dict(rows=rr["fp64"], default_over_rows=round(res["fp64"]["mlups"] / rr["fp64"]["mlups"], 3))
    cav = cavity_workload(args.cavity_n, args.cavity_steps, args.warmup, precs) if args.cavity_steps > 0 else None
    duct_res = duct_precision(args.duct_steps, precs) if args.duct_steps > 0 else None
    adj = None
    if args.adj_steps > 0:
        print(f"[bench] adjoint: {args.adj_steps} steps, checkpoint every {args.adj_interval}",
              file=sys.stderr, flush=True)
        adj = dict(workload="porous channel 128x64x64 (spheres r 4..8, 10% solid, 8-layer open buffers at "
                            "inlet/outlet), BGK tau=0.7, gx=1e-5, Zou-He inlet (peak 0.02) / outlet; "
                            "J = mean outlet x-flux; gradients w.r.t. the body force and the inlet velocity")
        adj.update(time_adjoint(porous_channel, "fp32", args.adj_steps, args.adj_interval))
        print(f"[bench] adjoint/forward {adj['adjoint_over_forward']} (reverse pass), "
              f"{adj['gradient_over_forward']} (whole gradient)", file=sys.stderr, flush=True)
        if args.adj_interval < args.adj_steps:
            one = time_adjoint(porous_channel, "fp32", args.adj_steps, args.adj_steps)
            adj["one_segment"] = {k: one[k] for k in ("checkpoint_interval", "device_bytes", "forward_ms_per_step",
                                                      "adjoint_forward_ms_per_step", "adjoint_reverse_ms_per_step",
                                                      "adjoint_over_forward", "gradient_over_forward")}
    if args.adj768_steps > 0:
        print(f"[bench] adjoint on the medium: {args.adj768_steps} steps, checkpoint every 25",
              file=sys.stderr, flush=True)
        a5 = time_adjoint(lambda p: porous_medium(p, solid=solid), "fp32", args.adj768_steps, 25, grad_face=False)
        a5["workload"] = C5_NAME + "; J = mean x-flux through the X1 layer; gradient w.r.t. the body force"
        adj = adj or {}
        adj["porous_medium"] = a5
        print(f"[bench] medium adjoint/forward {a5['adjoint_over_forward']} (reverse pass), "
              f"{a5['gradient_over_forward']} (whole gradient)", file=sys.stderr, flush=True)
    sol = None
    if args.solver_steps > 0:
        sol = {}
        for name, mk in (("cavity256", lambda: make_cavity(256, "fp64")),
                         ("porous768", lambda: porous_medium("fp64", solid=solid))):
            print(f"[bench] Solver.evolve({args.solver_steps}) on the {name} workload", file=sys.stderr, flush=True)
            sol[name] = solver_evolve(mk, args.solver_steps)
    cpu = None
    if not args.no_cpu:
        print(f"[bench] cpu baseline: {args.cpu_steps} oracle steps", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.cpu_steps, solid)
    main_p = "fp64" if "fp64" in res else precs[0]
    m = res[main_p]
    out = {
        "metric": METRIC,
        "value": m["mlups"],
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_p == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": C5_NAME, "lattice": list(C5_LAT), "lattice_cells": int(np.prod(C5_LAT)),
                   "fluid_sites": m["nfluid"], "layout": m["layout"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": m["achieved_gbs"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": m["frac"], "traffic": load_traffic(main_p, m["layout"]), "kernel": m["kernel"],
                     "bytes_per_site": BYTES_PER_SITE[main_p], "copy_ceiling_gbs": ceiling,
                     "frac_of_copy": m["frac_of_copy"]},
        "variants": res,
        "precision_error_vs_fp64": errors,
        "layout_compare": compare,
        "secondary_workload": cav,
        "precision_workload": duct_res,
        "adjoint": adj,
        "solver_evolve": sol,
        "cpu_baseline": cpu,
    }
    if "fp32" in res and "fp64" in res:
        out["fp32_over_fp64"] = round(res["fp32"]["mlups"] / res["fp64"]["mlups"], 3)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
