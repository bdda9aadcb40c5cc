"""Throughput benchmark of the D3Q19 evolve path on one MI355X.

Workload (the headline configuration): lid-driven cavity, 256^3 lattice,
D3Q19 TRT (magic parameter 3/16), tau = 0.6, rho0 = 1, bounce-back walls on
all faces, the lid a SOLID layer at the top z plane moving with ux = 0.05.
A "step" is one fused AA stream+collide pass over the lattice, with the
populations resident in HBM. The same workload runs in FP64 (the reference's
numerics), FP32 on shifted populations, FP16 storage / FP32 arithmetic and
FP32 storage / FP64 arithmetic; the reduced-precision velocity fields are
compared with the FP64 one after the same number of steps. Beside it: the
precision workload (Zou-He duct 512x128x128, regularized, 5000 steps) in each
precision with its velocity error against FP64, and the adjoint workload
(porous channel 128x64x64, 1000 steps, checkpoint every 50) with the
adjoint/forward step-time ratio.

Prints one JSON line (see --help). The CPU baseline is the FP64 CPU oracle
(a C/OpenMP restatement of the reference algorithm) on a bounded sample of the
same workload, timed on this host's cores.
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
KERNEL = {"fp64": "k_step<0,1,*>", "fp32": "k_step<1,1,*>", "fp16": "k_step_h<1,*>", "fp32_storage": "k_step<3,1,*>"}


def cavity_inputs(n):
    from jyulb.defs import FLUID, SOLID
    phase = np.full((n, n, n), FLUID, dtype=np.uint8)
    phase[:, :, n - 1] = SOLID
    vel = np.zeros((n, n, n, 3))
    vel[:, :, n - 1, 0] = 0.05
    den = np.ones((n, n, n))
    return phase, den, vel


def make_cavity(n, precision, layout=0):
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


def time_device(solver, steps, warmup):
    import torch
    eng = solver.engine
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    # what Solver.evolve does on entry: Zou-He face data, gravity, acc, walls
    solver._set_bc_data()
    solver._sync_inputs()
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


def duct(precision, lat=(512, 128, 128), peak=0.04, tau=0.55):
    """Precision workload: Zou-He duct, regularized collision, tau = 0.55,
    parabolic inlet (peak 0.04) at 1X, outlet density 1 at X1, bounce-back
    walls in y and z."""
    from jyulb.isothermal import Solver
    from jyulb.defs import REGULARIZATION, WALL, FIXED_DEN, FIXED_VEL
    nx, ny, nz = lat
    s = Solver(np.array(lat, dtype=np.int32), REGULARIZATION, (tau - 0.5) / 3.0, np.zeros(3),
               np.array([FIXED_VEL, FIXED_DEN, WALL, WALL, WALL, WALL], dtype=np.int32), precision=precision)
    s.get_phase()[...] = 1
    ey = (np.arange(ny) + 0.5) / ny - 0.5
    ez = (np.arange(nz) + 0.5) / nz - 0.5
    s.get_vel()[0, :, :, 0] = peak * (1 - (2 * ey[:, None]) ** 2) * (1 - (2 * ez[None, :]) ** 2)
    nfl = s.finalize_init()
    return s, nfl


def duct_precision(steps, precs):
    """The duct in each precision over `steps` steps: MLUPS (HIP events over
    the whole run, one evolve call) and the velocity error against FP64."""
    out, fields, fl = {}, {}, None
    for prec in precs:
        s, nfl = duct(prec)
        print(f"[bench] duct {prec}: {steps} steps", file=sys.stderr, flush=True)
        t, _ = time_device(s, steps, 0)
        fields[prec] = np.array(s.engine.download_den_vel()[1])
        out[prec] = dict(mlups=round(nfl * steps / t / 1e6, 1))
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


def porous_channel(precision, lat=(128, 64, 64), seed=4, peak=0.02):
    """Adjoint workload: porous channel of seeded spheres (radius 4..8, 10 %
    solid, none in 8-layer buffers behind the inlet and outlet faces), BGK
    tau = 0.7, gravity gx = 1e-5, parabolic inlet velocity face (peak 0.02) at
    1X, outlet density 1 at X1, bounce-back walls in y and z."""
    from jyulb.isothermal import Solver
    from jyulb.geometry import porous_spheres
    from jyulb.defs import BGK, WALL, FIXED_DEN, FIXED_VEL, FACE_1X, FACE_X1
    nx, ny, nz = lat
    s = Solver(np.array(lat, dtype=np.int32), BGK, 0.2 / 3.0, np.array([1e-5, 0.0, 0.0]),
               np.array([FIXED_VEL, FIXED_DEN, WALL, WALL, WALL, WALL], dtype=np.int32), precision=precision)
    solid = porous_spheres(lat, 4, 8, 0.10, seed, open_faces=(FACE_1X, FACE_X1), open_depth=8)
    s.get_phase()[...] = np.where(solid, 0, 1)
    ey = (np.arange(ny) + 0.5) / ny - 0.5
    ez = (np.arange(nz) + 0.5) / nz - 0.5
    s.get_vel()[0, :, :, 0] = peak * (1 - (2 * ey[:, None]) ** 2) * (1 - (2 * ez[None, :]) ** 2)
    nfl = s.finalize_init()
    return s, nfl


def porous_medium(precision, lat=(768, 384, 384), seed=5, layout="rows", radius=(6, 12), solid=None):
    """Throughput workload: seeded random-sphere porous medium (radius 6..12,
    20 % solid), TRT tau = 0.6, gravity gx = 1e-6, periodic on all faces
    (113M lattice cells). `solid` reuses a generated geometry."""
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


def sparse_workload(steps, precs, lat=(384, 192, 192)):
    """Dense ROWS against the sparse fluid-only layout on a porous medium of
    384x192x192 cells (spheres of radius 3..6 up to 20 % solid, seed 5: the
    768x384x384 medium at half scale), TRT tau = 0.6, periodic."""
    import torch
    from jyulb.geometry import porous_spheres
    solid = porous_spheres(lat, 3, 6, 0.20, 5)
    ncell = int(np.prod(lat))
    out = {}
    for prec in precs:
        row = {}
        for layout in ("rows", "sparse"):
            s, nfl = porous_medium(prec, lat=lat, layout=layout, solid=solid)
            t, _ = time_device(s, steps, 40)  # the warmup holds the launch tuning
            row[layout] = dict(mlups=round(nfl * steps / t / 1e6, 1), mlups_cells=round(ncell * steps / t / 1e6, 1),
                               ms_per_step=round(1e3 * t / steps, 4))
            del s
            torch.cuda.synchronize()
        row["sparse_over_rows"] = round(row["sparse"]["mlups"] / row["rows"]["mlups"], 3)
        print(f"[bench] porous {lat[0]}x{lat[1]}x{lat[2]} {prec}: rows {row['rows']['mlups']} MLUPS, "
              f"sparse {row['sparse']['mlups']} MLUPS", file=sys.stderr, flush=True)
        out[prec] = row
    return dict(workload=f"porous medium {lat[0]}x{lat[1]}x{lat[2]} (spheres r 3..6, 20% solid, seed 5), "
                         "TRT tau=0.6, gx=1e-6, periodic; dense ROWS vs sparse fluid-only layout",
                steps=steps, lattice_cells=ncell, fluid_sites=int(ncell - solid.sum()), variants=out)


def throughput_workload(steps, precs, layouts=("rows",)):
    """The 768x384x384 porous medium in each precision and layout: MLUPS over
    the fluid sites and over all lattice cells (HIP events around one evolve
    call). Entries are keyed by precision, with ":<layout>" for layouts other
    than rows."""
    out = {}
    ncell = 768 * 384 * 384
    for layout, prec in [(lay, pr) for lay in layouts for pr in precs]:
        s, nfl = porous_medium(prec, layout=layout)
        key = prec if layout == "rows" else f"{prec}:{layout}"
        print(f"[bench] porous 768x384x384 {key}: {nfl} fluid sites, {steps} steps", file=sys.stderr, flush=True)
        t, _ = time_device(s, steps, 40)  # the warmup holds the launch tuning
        out[key] = dict(mlups=round(nfl * steps / t / 1e6, 1), mlups_cells=round(ncell * steps / t / 1e6, 1),
                         ms_per_step=round(1e3 * t / steps, 4),
                         achieved_gbs=round(BYTES_PER_SITE[prec] * nfl * steps / t / 1e9, 1),
                         device_bytes=int(s.engine.info()["device_bytes"]))
        del s
    adj = {}
    for layout in (layouts if "fp32" in precs else ()):
        # adjoint step cost at checkpoint interval 25 (50 steps: 3 checkpoints)
        import torch
        from jyulb.defs import FACE_X1, X
        from jyulb.adjoint import _prepare
        s, nfl = porous_medium("fp32", layout=layout)
        eng = s.engine
        eng.set_stream(torch.cuda.current_stream().cuda_stream)
        _prepare(s)
        eng.snapshot()
        ga = torch.empty(tuple(s.get_acc().shape), dtype=torch.float64, device="cuda")
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        n_adj, interval = 50, 25
        for _ in range(2):  # the first pass allocates the checkpoints and adjoint buffers
            eng.restore()
            torch.cuda.synchronize()
            ev[0].record()
            eng.evolve(n_adj)
            ev[1].record()
            eng.restore()
            eng.adj_forward(n_adj, interval, FACE_X1, X)
            ev[2].record()
            eng.adj_backward_device(-1, ga.data_ptr(), 0)
            ev[3].record()
            torch.cuda.synchronize()
        t_f, t_af, t_ab = (ev[0].elapsed_time(ev[1]), ev[1].elapsed_time(ev[2]), ev[2].elapsed_time(ev[3]))
        print(f"[bench] porous 768x384x384 adjoint fp32 {layout}: reverse/forward {t_ab / t_f:.3f}",
              file=sys.stderr, flush=True)
        adj["adjoint" if layout == "rows" else f"adjoint:{layout}"] = dict(
                   precision="fp32", layout=layout, steps=n_adj, checkpoint_interval=interval,
                   forward_ms_per_step=round(t_f / n_adj, 4), adjoint_reverse_ms_per_step=round(t_ab / n_adj, 4),
                   adjoint_over_forward=round(t_ab / t_f, 3), gradient_over_forward=round((t_af + t_ab) / t_f, 3),
                   device_bytes=int(eng.info()["device_bytes"]))
        del s
    return dict(workload="porous medium 768x384x384 (spheres r 6..12, 20% solid), TRT tau=0.6, gx=1e-6, "
                         "periodic", steps=steps, lattice_cells=ncell, variants=out, **adj)


def time_adjoint(precision, steps, interval):
    """Forward evolve vs the adjoint (forward pass with checkpoints, then the
    reverse pass with per-segment recomputation) over the same steps, all on
    torch's current stream, timed with HIP events. The forward pass records
    (rho, u*) over its last segment, so with interval >= steps the reverse
    pass recomputes nothing."""
    import torch
    from jyulb.defs import FACE_1X, FACE_X1, X
    from jyulb.adjoint import _prepare
    s, nfl = porous_channel(precision)
    eng = s.engine
    eng.set_stream(torch.cuda.current_stream().cuda_stream)
    _prepare(s)
    eng.snapshot()
    ga = torch.empty(tuple(s.get_acc().shape), dtype=torch.float64, device="cuda")
    gf = torch.empty(eng.face_shape(FACE_1X) + (3,), dtype=torch.float64, device="cuda")
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]

    def run(timed):
        eng.restore()
        torch.cuda.synchronize()
        ev[0].record()
        eng.evolve(steps)
        ev[1].record()
        eng.restore()
        J = eng.adj_forward(steps, interval, FACE_X1, X)
        ev[2].record()
        eng.adj_backward_device(FACE_1X, ga.data_ptr(), gf.data_ptr())
        ev[3].record()
        torch.cuda.synchronize()
        return J

    run(False)  # allocates checkpoints and adjoint buffers
    J = run(True)
    t_fwd = ev[0].elapsed_time(ev[1]) / 1e3
    t_af = ev[1].elapsed_time(ev[2]) / 1e3
    t_ab = ev[2].elapsed_time(ev[3]) / 1e3
    del s
    return dict(workload="porous channel 128x64x64 (spheres r 4..8, 10% solid, 8-layer open buffers at inlet/outlet), BGK tau=0.7, gx=1e-5, "
                         "Zou-He inlet (peak 0.02) / outlet; J = mean outlet x-flux",
                precision=precision, steps=steps, checkpoint_interval=interval, fluid_sites=nfl, J=J,
                device_bytes=eng.info()["device_bytes"],  # lattice, checkpoints and (rho, u*) record
                forward_ms_per_step=round(1e3 * t_fwd / steps, 5),
                adjoint_forward_ms_per_step=round(1e3 * t_af / steps, 5),
                adjoint_reverse_ms_per_step=round(1e3 * t_ab / steps, 5),
                adjoint_over_forward=round(t_ab / t_fwd, 3),
                gradient_over_forward=round((t_af + t_ab) / t_fwd, 3))


def cpu_baseline(n, steps):
    """FP64 CPU oracle on the same cavity (full lattice, `steps` steps)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import jyulb_oracle as jo
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from jyulb.defs import TRT, WALL
    s = jo.OracleSolver(np.array([n, n, n]), TRT, (0.6 - 0.5) / 3.0, np.zeros(3), np.full(6, WALL))
    phase, den, vel = cavity_inputs(n)
    s.get_phase()[...] = phase
    s.get_den()[...] = den
    s.get_vel()[...] = vel
    nfl = s.finalize_init()
    t = time.perf_counter()
    s.evolve(steps)
    dt = time.perf_counter() - t
    return dict(value=nfl * steps / dt / 1e6, unit="MLUPS", cores=cores, kind="port",
                sample=f"FP64 CPU oracle (C/OpenMP restatement of the reference step), same "
                       f"{n}^3 TRT cavity, {steps} steps after init, {cores} threads, evolve() only")


def load_traffic(prec):
    path = os.path.join(ROOT, "profiles", "k_step_cavity256_traffic.json")
    if not os.path.exists(path):
        return None
    try:
        d = json.load(open(path))
        return d.get(prec, {}).get("hbm_bytes_per_launch")
    except Exception:
        return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=256, help="cavity edge length")
    ap.add_argument("--precisions", default="fp64,fp32,fp16,fp32_storage")
    ap.add_argument("--cpu-steps", type=int, default=60)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--layout", type=int, default=0, help="population layout (LB_LAYOUT_*)")
    ap.add_argument("--adj-steps", type=int, default=1000, help="steps of the adjoint workload (0: skip)")
    ap.add_argument("--duct-steps", type=int, default=5000, help="steps of the precision workload (0: skip)")
    ap.add_argument("--porous768-steps", type=int, default=0,
                    help="steps of the 768x384x384 porous-medium throughput workload (0: skip, the default)")
    ap.add_argument("--porous768-layouts", default="rows",
                    help="population layouts of the porous-medium workload, e.g. rows,sparse")
    ap.add_argument("--sparse-steps", type=int, default=200,
                    help="steps of the rows-vs-sparse porous workload, 384x192x192 (0: skip)")
    ap.add_argument("--adj-interval", type=int, default=50)
    ap.add_argument("--calibrate", action="store_true",
                    help="also run a plain device copy of 2 GiB (torch copy_), the reference "
                         "kernel for scaling rocprof FETCH_SIZE/WRITE_SIZE counters")
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
    if args.calibrate:
        a = torch.empty(2 ** 28, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1.0)
        torch.cuda.synchronize()
        b.copy_(a)
        torch.cuda.synchronize()
        print("[bench] calibration copy: 2 GiB read + 2 GiB written", file=sys.stderr, flush=True)
        del a, b
    precs = [p for p in args.precisions.split(",") if p]
    res = {}
    fields = {}
    for prec in precs:
        s, nfl = make_cavity(args.n, prec, args.layout)
        print(f"[bench] {prec}: {args.n}^3 cavity, {nfl} fluid sites; warmup {args.warmup}, "
              f"timing {args.steps} steps", file=sys.stderr, flush=True)
        t, wall = time_device(s, args.steps, args.warmup)
        den, vel = s.engine.download_den_vel()
        fields[prec] = vel
        mlups = nfl * args.steps / t / 1e6
        gbs = BYTES_PER_SITE[prec] * nfl * args.steps / t / 1e9
        res[prec] = dict(mlups=mlups, ms_per_step=1e3 * t / args.steps, wall_s=wall, nfluid=nfl,
                         achieved_gbs=gbs, scale=s.engine.info()["fp16_scale"])
        print(f"[bench] {prec}: {mlups:.1f} MLUPS, {1e3 * t / args.steps:.4f} ms/step, "
              f"{gbs:.0f} GB/s algorithmic", file=sys.stderr, flush=True)
        del s
        torch.cuda.synchronize()
    errors = {}
    if "fp64" in fields:
        ref = fields["fp64"]
        fl = np.ones(ref.shape[:3], dtype=bool)
        fl[:, :, -1] = False
        r = ref[fl]
        nref = np.sqrt(np.sum(r * r))
        for p, v in fields.items():
            if p == "fp64":
                continue
            d = v[fl] - r
            errors[p] = dict(vel_rel_l2=float(np.sqrt(np.sum(d * d)) / nref),
                             vel_linf=float(np.max(np.abs(d))),
                             vel_rel_linf=float(np.max(np.abs(d)) / np.max(np.abs(r))))
    duct_res = duct_precision(args.duct_steps, precs) if args.duct_steps > 0 else None
    big = (throughput_workload(args.porous768_steps, precs, args.porous768_layouts.split(","))
           if args.porous768_steps > 0 else None)
    sparse = sparse_workload(args.sparse_steps, precs) if args.sparse_steps > 0 else None
    adj = None
    if args.adj_steps > 0:
        print(f"[bench] adjoint: {args.adj_steps} steps, checkpoint every {args.adj_interval}",
              file=sys.stderr, flush=True)
        adj = time_adjoint("fp32", args.adj_steps, args.adj_interval)
        print(f"[bench] adjoint/forward {adj['adjoint_over_forward']} (reverse pass), "
              f"{adj['gradient_over_forward']} (whole gradient)", file=sys.stderr, flush=True)
        if args.adj_interval < args.adj_steps:
            # one segment: the forward pass records every step and nothing is recomputed
            one = time_adjoint("fp32", args.adj_steps, args.adj_steps)
            adj["one_segment"] = {k: one[k] for k in ("checkpoint_interval", "device_bytes", "forward_ms_per_step",
                                                      "adjoint_forward_ms_per_step", "adjoint_reverse_ms_per_step",
                                                      "adjoint_over_forward", "gradient_over_forward")}
            print(f"[bench] one segment: adjoint/forward {one['adjoint_over_forward']} (reverse pass), "
                  f"{one['gradient_over_forward']} (whole gradient)", file=sys.stderr, flush=True)
    main_p = "fp64" if "fp64" in res else precs[0]
    m = res[main_p]
    cpu = None
    if not args.no_cpu:
        print(f"[bench] cpu baseline: {args.cpu_steps} oracle steps", file=sys.stderr, flush=True)
        cpu = cpu_baseline(args.n, args.cpu_steps)
    out = {
        "metric": METRIC,
        "value": round(m["mlups"], 2),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(m["ms_per_step"], 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if main_p == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": f"lid-driven cavity {args.n}^3, D3Q19 TRT (magic 3/16), tau=0.6, "
                               f"lid ux=0.05, bounce-back walls", "lattice": [args.n] * 3,
                   "fluid_sites": m["nfluid"], "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": round(m["achieved_gbs"], 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(m["achieved_gbs"] / PEAK_HBM_GBS, 4),
                     "traffic": load_traffic(main_p), "kernel": KERNEL[main_p],
                     "bytes_per_site": BYTES_PER_SITE[main_p]},
        "variants": {p: {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
                     for p, r in res.items()},
        "precision_error_vs_fp64": errors,
        "precision_workload": duct_res,
        "throughput_workload": big,
        "sparse_workload": sparse,
        "cpu_baseline": cpu,
        "adjoint": adj,
    }
    if "fp32" in res and "fp64" in res:
        out["fp32_over_fp64"] = round(res["fp32"]["mlups"] / res["fp64"]["mlups"], 3)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
