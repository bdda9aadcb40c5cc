"""Gradients of flow objectives through the evolve path (reverse mode).

No counterpart in the reference: this differentiates `steps` evolve steps of
a Solver with respect to the per-site acceleration field (get_acc(),
isothermal.pyx:1293-1295) and the velocity data of a Zou-He face
(bc_data_vel_face_*, isothermal.pyx:476-502), for the objective

    J = mean over the fluid sites of a face layer of rho * u*_axis

(e.g. face X1, axis X: the mean outlet x-flux), evaluated on the density and
half-force shifted velocity of the last step, or the dissipation objective

    J = mean over the fluid sites of Pi_neq : Pi_neq

with Pi_neq = sum_l c_l c_l f_l - rho u* u* - rho/3 I the non-equilibrium
momentum flux of the last step's collision input (the viscous stress is
-(1 - omega/2) Pi_neq). The reverse pass is the
hand-written adjoint AA step of libjyulb_hip.so with checkpoints every
`interval` steps and recomputation (include/jyulb_gpu.h: lb_adj_forward,
lb_adj_backward). The forward pass records its last segment, so with
interval >= steps nothing is recomputed, at 16 bytes per site and step in FP32
(32 in FP64). Parameters are held fixed over the steps; the populations at
the start are the Solver's current state. Walls must be stationary; the
arithmetic is FP64 or FP32 (the Solver's precision).
"""
import os

import numpy as np

from .defs import FACE_X1, X, FIXED_DEN, FIXED_VEL

_NAMES = ('1X', 'X1', '1Y', 'Y1', '1Z', 'Z1')
OBJ_DISSIPATION = 6  # LB_OBJ_DISSIPATION of include/jyulb_gpu.h


def _prepare(solver):
    if not solver.init_done or solver.fluid_lsite_domain is None:
        raise RuntimeError('Finalize solver initialization before differentiating evolve')
    solver._set_bc_data()
    solver._load_inputs()


def _finish(solver, steps):
    solver._pull_fields()
    solver.tot_tsteps = (solver.tot_tsteps + int(steps)) & 0xFFFFFFFF


def checkpoint_interval(solver, steps, budget_bytes=None, objective="flux"):
    """The checkpoint interval with the least recomputation whose device
    memory fits in `budget_bytes` (default: half of the free device memory).

    With n segments of length i = ceil(steps / n) the reverse pass recomputes
    all but the last segment; the fewest segments that fit are the fastest
    choice, and i = steps (one segment) recomputes nothing. Memory for
    interval i: (ceil(steps / i) + 1) population arrays (19 values per site)
    for the checkpoints, min(i, steps) x 4 values per site for the (rho, u*)
    record, one population array of cotangents (two where the adjoint steps
    take their pull form: from 1 GiB of cotangents, or as LB_ADJ_PULL says)
    and 3 values per site of the
    acceleration cotangent, plus (Q + 4) values per site of seeds for the
    dissipation objective; nx * ny * round_up(nz, 64) sites and 8 bytes per
    value in FP64, 4 in FP32. Raises MemoryError if even the smallest
    footprint does not fit.
    """
    from ._native import LB_FP64
    eng = solver.engine
    nx, ny, nz = eng.lat
    nq = int(nx) * int(ny) * ((int(nz) + 63) // 64 * 64)
    vb = 8 if eng.precision == LB_FP64 else 4
    fb, rb = 19 * nq * vb, 4 * nq * vb
    env = os.environ.get('LB_ADJ_PULL')
    if env is not None:
        try:
            pull = int(env) != 0
        except ValueError:
            pull = False
    else:
        pull = fb >= 1 << 30
    extra = (2 if pull else 1) * fb + 3 * nq * vb
    extra += 23 * nq * vb if objective == "dissipation" else 0
    steps = int(steps)
    if steps < 1:
        raise ValueError('steps must be >= 1')
    if budget_bytes is None:
        import torch
        budget_bytes = torch.cuda.mem_get_info()[0] // 2

    def need(i):
        return (-(-steps // i) + 1) * fb + min(i, steps) * rb + extra

    for nseg in range(1, steps + 1):
        i = -(-steps // nseg)
        if need(i) <= budget_bytes:
            return i
    raise MemoryError(f'the adjoint of {steps} steps needs at least '
                      f'{min(need(i) for i in range(1, steps + 1))} bytes; {budget_bytes} available')


def _interval(solver, steps, interval, objective="flux"):
    return checkpoint_interval(solver, steps, objective=objective) if interval == 'auto' else int(interval)


def flux_gradient(solver, steps, interval=50, face=FACE_X1, axis=X, grad_face=None):
    """Advance `solver` by `steps` steps and return (J, grad_acc, grad_face_vel).

    grad_acc has the shape of get_acc(); grad_face_vel has the shape of
    solver.bc_data_vel_face_<grad_face> (None if grad_face is None). The
    density/velocity views hold the state after the steps, as after evolve.
    interval='auto' takes checkpoint_interval(solver, steps).
    """
    _prepare(solver)
    eng = solver.engine
    J = eng.adj_forward(steps, _interval(solver, steps, interval), face, axis)
    ga, gf = eng.adj_backward(-1 if grad_face is None else int(grad_face))
    _finish(solver, steps)
    return J, ga, gf


def dissipation_gradient(solver, steps, interval=50, grad_face=None):
    """Like flux_gradient, for the dissipation objective."""
    _prepare(solver)
    eng = solver.engine
    J = eng.adj_forward(steps, _interval(solver, steps, interval, 'dissipation'), OBJ_DISSIPATION, 0)
    ga, gf = eng.adj_backward(-1 if grad_face is None else int(grad_face))
    _finish(solver, steps)
    return J, ga, gf


def fixed_face(solver):
    """The FIXED_VEL face of the solver (or the first FIXED_DEN one), or None."""
    bc = solver.bc
    for kind in (FIXED_VEL, FIXED_DEN):
        for f in range(6):
            if bc[f] == kind:
                return f
    return None


def _torch_function():
    import torch

    class FluxObjective(torch.autograd.Function):
        """J(acc, face_vel) for one run of `steps` evolve steps on the GPU.

        acc: float64 tensor of shape get_acc() on the GPU; face_vel: float64
        tensor with the shape of bc_data_vel_face_<inlet_face> (or None).
        backward() runs the adjoint and returns dJ/dacc and dJ/dface_vel.
        """

        @staticmethod
        def forward(ctx, acc, face_vel, solver, steps, interval, face, axis, inlet_face):
            eng = solver.engine
            eng.set_stream(torch.cuda.current_stream().cuda_stream)
            _prepare(solver)
            acc_c = acc.detach().to(torch.float64).contiguous()
            eng.upload_acc_device(acc_c.data_ptr())
            den_face = None
            if face_vel is not None:
                den_face = torch.as_tensor(getattr(solver, 'bc_data_den_face_' + _NAMES[inlet_face]),
                                           dtype=torch.float64, device=acc.device).contiguous()
                fv = face_vel.detach().to(torch.float64).contiguous()
                eng.set_face_data_device(inlet_face, den_face.data_ptr(), fv.data_ptr())
            J = eng.adj_forward(steps, interval, face, axis)
            _finish(solver, steps)
            ctx.solver = solver
            ctx.gen = eng.adj_gen  # this forward pass; backward() must follow it
            ctx.inlet_face = inlet_face if face_vel is not None else -1
            ctx.acc_shape = acc.shape
            ctx.face_shape = None if face_vel is None else face_vel.shape
            ctx.device = acc.device
            return torch.tensor(J, dtype=torch.float64, device=acc.device)

        @staticmethod
        def backward(ctx, gJ):
            eng = ctx.solver.engine
            if eng.adj_gen != ctx.gen:
                raise RuntimeError('the solver ran another adjoint forward pass after this objective was '
                                   'evaluated: its reverse pass needs its own forward pass (one objective per '
                                   'solver per backward)')
            eng.set_stream(torch.cuda.current_stream().cuda_stream)
            ga = torch.empty(ctx.acc_shape, dtype=torch.float64, device=ctx.device)
            gf = None
            if ctx.face_shape is not None:
                gf = torch.empty(ctx.face_shape, dtype=torch.float64, device=ctx.device)
            eng.adj_backward_device(ctx.inlet_face, ga.data_ptr(), 0 if gf is None else gf.data_ptr())
            ga = ga * gJ
            if gf is not None:
                gf = gf * gJ
            return ga, gf, None, None, None, None, None, None

    return FluxObjective


_FN = None


def flux_objective(acc, face_vel, solver, steps, interval=50, face=FACE_X1, axis=X, inlet_face=None):
    """torch entry point: J as a differentiable function of the acceleration
    field `acc` and (optionally) the Zou-He face velocity data `face_vel`.
    interval='auto' takes checkpoint_interval(solver, steps)."""
    global _FN
    if _FN is None:
        _FN = _torch_function()
    if face_vel is not None and inlet_face is None:
        inlet_face = fixed_face(solver)
    obj = "dissipation" if int(face) == OBJ_DISSIPATION else "flux"
    return _FN.apply(acc, face_vel, solver, int(steps), _interval(solver, steps, interval, obj), int(face),
                     int(axis), -1 if inlet_face is None else int(inlet_face))


def dissipation_objective(acc, face_vel, solver, steps, interval=50, inlet_face=None):
    """torch entry point for the dissipation objective (flux_objective with
    face = OBJ_DISSIPATION): mean Pi_neq : Pi_neq of the last step as a
    differentiable function of `acc` and (optionally) `face_vel`."""
    return flux_objective(acc, face_vel, solver, steps, interval, OBJ_DISSIPATION, 0, inlet_face)
