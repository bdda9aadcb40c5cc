"""Lattice-Boltzmann solver for isothermal (single-phase) fluid flows, D3Q19.

Drop-in for the reference's `jyulb.isothermal.Solver` (isothermal.pyx):
same constructors, property, methods, exceptions, halo-padded host arrays and
aliasing interior views. The per-site work of `finalize_init` and `evolve`
(streaming with halfway bounce-back and Ladd's moving walls, Zou-He faces,
BGK / TRT / regularized collision with forcing, the density and velocity
moments) runs on the GPU in the fused AA-pattern kernel of libjyulb_hip.so;
this module keeps the host-side semantics: argument checks, the halo passes
of the density/velocity/phase fields, `_set_bc_data` and the hand-off of the
mutable host arrays on every `evolve` call.

All quantities are in lattice units (isothermal.pyx:1-36).

Additions, keyword-only so that reference call sites are unaffected:
  precision  "fp64" (default; the reference's numerics), "fp32" (FP32
             arithmetic on shifted populations f - w*rho0), "fp16" (FP16
             storage of the scaled shifted populations, FP32 arithmetic) or
             "fp32_storage" (FP32 storage of the shifted populations, FP64
             arithmetic: the FP32 traffic with the rounding of storage only)
  layout     population layout on the device: "auto" (default: chosen at
             finalize_init, "sparse" when at least 5 % of the cells are solid,
             "rows" otherwise), "rows", "soa", "soa_padded", or "sparse"
             (fluid sites only, with a table of the odd-step locations: less
             traffic on porous media). Results are identical in every layout.
"""
import numpy as np

from . import _native
from .D3Q19 import D3Q19_H10, LVECS
from .defs import (X, Y, Z, SOLID, FACE_1X, FACE_X1, FACE_1Y, FACE_Y1, FACE_1Z, FACE_Z1,
                   WALL, PERIODIC, FIXED_DEN, FIXED_VEL, SYMMETRIC, BGK, TRT, REGULARIZATION)

_FIXED = (FIXED_DEN, FIXED_VEL)


class Solver:
    """Isothermal D3Q19 solver (isothermal.pyx:132-1706).

    Solver(lat_size, rlx_scheme=BGK, kvisc=1/6, g=None, bc=None)
    """

    # ------------------------------------------------------------------
    # constructors
    # ------------------------------------------------------------------
    def __init__(self, lat_size, rlx_scheme=BGK, kvisc=1.0 / 6.0, g=None, bc=None, *,
                 precision="fp64", rho0=1.0, fp16_scale=0.0, layout="auto", keep_f2=False):
        """isothermal.pyx:156-298"""
        self.error_msg = ''
        if lat_size is None:
            raise TypeError("Argument 'lat_size' must not be None")
        lat_size = np.asarray(lat_size)
        if lat_size[X] < 1 or lat_size[Y] < 1 or lat_size[Z] < 1:
            self.error_msg = 'Invalid lattice size input.'
            raise ValueError(self.error_msg)
        if rlx_scheme not in (BGK, TRT, REGULARIZATION):
            self.error_msg = 'Invalid relaxation scheme input.'
            raise ValueError(self.error_msg)
        if kvisc <= 0.0:
            self.error_msg = 'Invalid kinematic viscosity input.'
            raise ValueError(self.error_msg)
        self._g = np.zeros(3) if g is None else np.array([g[X], g[Y], g[Z]], dtype=np.float64)
        if bc is None:
            self._bc = [PERIODIC] * 6
        else:
            self._check_bc(bc)
            self._bc = [int(bc[f]) for f in range(6)]

        NI, NJ, NK = int(lat_size[X]) + 2, int(lat_size[Y]) + 2, int(lat_size[Z]) + 2
        self.NI, self.NJ, self.NK, self.NCNT = NI, NJ, NK, NI * NJ * NK

        # device state (allocated now, like the reference's f1/f2)
        self.precision = precision
        self._eng = _native.Engine((NI - 2, NJ - 2, NK - 2), rlx_scheme, kvisc, self._g,
                                   self._bc, precision=precision, rho0=rho0, fp16_scale=fp16_scale,
                                   layout=layout)
        self._keep_f2 = bool(keep_f2)
        if self._keep_f2:
            self._eng.keep_previous(True)

        # flow field, isothermal.pyx:243-248, in pinned host memory that the
        # engine reads and writes in place on every evolve call
        self._phase = _native.host_array((NI, NJ, NK), np.uint8, SOLID)
        self.fluid_lsite_domain = None
        self._den = _native.host_array((NI, NJ, NK), np.float64, 1.0)
        self._vel = _native.host_array((NI, NJ, NK, 3), np.float64, 0.0)
        self._acc = _native.host_array((NI, NJ, NK, 3), np.float64, 0.0)

        # boundary condition data, isothermal.pyx:253-263
        self.bc_data_den_face_1X = np.ones((NJ, NK))
        self.bc_data_den_face_X1 = np.ones((NJ, NK))
        self.bc_data_den_face_1Y = np.ones((NI, NK))
        self.bc_data_den_face_Y1 = np.ones((NI, NK))
        self.bc_data_den_face_1Z = np.ones((NI, NJ))
        self.bc_data_den_face_Z1 = np.ones((NI, NJ))
        self.bc_data_vel_face_1X = np.zeros((NJ, NK, 3))
        self.bc_data_vel_face_X1 = np.zeros((NJ, NK, 3))
        self.bc_data_vel_face_1Y = np.zeros((NI, NK, 3))
        self.bc_data_vel_face_Y1 = np.zeros((NI, NK, 3))
        self.bc_data_vel_face_1Z = np.zeros((NI, NJ, 3))
        self.bc_data_vel_face_Z1 = np.zeros((NI, NJ, 3))

        # velocity set and relaxation parameters, isothermal.pyx:266-287
        self.dvs = D3Q19_H10()
        self.kvisc = kvisc
        tau_e = 0.5 + self.dvs.AS2 * kvisc
        inv_tau_e = 1.0 / tau_e
        self.rlx_scheme = rlx_scheme
        self._rlx_prm = [0.0] * LVECS
        self._rlx_prm[0] = inv_tau_e
        if rlx_scheme == TRT:
            self._rlx_prm[1] = 8.0 * (2.0 - inv_tau_e) / (8.0 - inv_tau_e)
        # neighbour index offsets, isothermal.pyx:290-294
        self._njump = [self.dvs.LI[l] * NJ * NK + self.dvs.LJ[l] * NK + self.dvs.LK[l]
                       for l in range(LVECS)]
        self.tot_tsteps = 0
        self.init_done = 0
        self._fluid = None

    @classmethod
    def with_iflow(cls, phase, iden, ivel, rlx_scheme=BGK, kvisc=1.0 / 6.0, g=None, bc=None, **kw):
        """Solver with the initial flow field prescribed (isothermal.pyx:301-360)."""
        if phase is None or iden is None or ivel is None:
            raise TypeError("Arguments 'phase', 'iden' and 'ivel' must not be None")
        psx, psy, psz = phase.shape[X], phase.shape[Y], phase.shape[Z]
        for ax, n, name in ((X, psx, 'x'), (Y, psy, 'y'), (Z, psz, 'z')):
            if iden.shape[ax] != n or ivel.shape[ax] != n:
                raise ValueError('Invalid input for the initial flow field: '
                                 f'array sizes do not match ({name}-dimension)!')
        slvr = cls(np.array(phase.shape, dtype=np.int32), rlx_scheme, kvisc, g, bc, **kw)
        slvr.phase[1:slvr.NI - 1, 1:slvr.NJ - 1, 1:slvr.NK - 1] = phase
        slvr.den[1:slvr.NI - 1, 1:slvr.NJ - 1, 1:slvr.NK - 1] = iden
        slvr.vel[1:slvr.NI - 1, 1:slvr.NJ - 1, 1:slvr.NK - 1, :] = ivel
        return slvr

    # ------------------------------------------------------------------
    # read-only attributes of the reference (isothermal.pxd:136-176)
    # ------------------------------------------------------------------
    @property
    def phase(self):
        return self._phase

    @property
    def den(self):
        return self._den

    @property
    def vel(self):
        return self._vel

    @property
    def acc(self):
        return self._acc

    @property
    def bc(self):
        return list(self._bc)

    @property
    def g(self):
        return [float(v) for v in self._g]

    @property
    def rlx_prm(self):
        return list(self._rlx_prm)

    @property
    def njump(self):
        return list(self._njump)

    @property
    def gravity(self):
        """isothermal.pyx:365-373"""
        return [float(v) for v in self._g]

    @gravity.setter
    def gravity(self, g):
        self._g[X], self._g[Y], self._g[Z] = g[X], g[Y], g[Z]

    @property
    def f1(self):
        """Newest post-collision populations in the reference's Bundle-C
        layout (isothermal.pyx:1532-1704), flat 19*NCNT; zeros at halo and
        SOLID sites."""
        return self._bundle_c(self.post_collision_f())

    @property
    def f2(self):
        """The reference's f2 (isothermal.pxd:150-157): after evolve, the
        post-collision populations of the step before the last one, in f1's
        layout (zeros before the first evolve call). The engine holds one
        population array, so f2 is kept only for a Solver created with
        keep_f2=True, which copies the state before the last step of every
        evolve call."""
        if not self._keep_f2:
            raise AttributeError("f2 is kept only by a Solver created with keep_f2=True "
                                 "(the engine holds one population array)")
        out = np.zeros((LVECS, self.NI, self.NJ, self.NK))
        if self.init_done and self._fluid is not None:
            prev = self._eng.download_previous_post_f()
            if prev is not None:
                out[:, 1:-1, 1:-1, 1:-1] = prev
        return self._bundle_c(out)

    def _bundle_c(self, post):
        """(19, NI, NJ, NK) populations in the Bundle-C order of
        isothermal.pyx:1536-1704, flat 19*NCNT."""
        post = post.reshape(LVECS, -1)
        out = np.zeros(LVECS * self.NCNT)
        n = np.arange(self.NCNT)
        for l in range(LVECS):
            if l < 5:
                idx = 5 * n + l
            elif l < 14:
                idx = 5 * self.NCNT + 9 * n + (l - 5)
            else:
                idx = 14 * self.NCNT + 5 * n + (l - 14)
            out[idx] = post[l]
        return out

    # ------------------------------------------------------------------
    # configuration of the boundary conditions
    # ------------------------------------------------------------------
    def _check_bc(self, bc):
        """isothermal.pyx:378-426"""
        for s in range(6):
            if bc[s] not in (WALL, PERIODIC, FIXED_DEN, FIXED_VEL, SYMMETRIC):
                self.error_msg = 'Invalid boundary condition input.'
                raise ValueError(self.error_msg)
        for lo, hi, ax in ((FACE_1X, FACE_X1, 'x'), (FACE_1Y, FACE_Y1, 'y'), (FACE_1Z, FACE_Z1, 'z')):
            if (bc[lo] == PERIODIC or bc[hi] == PERIODIC) and bc[lo] != bc[hi]:
                self.error_msg = f'Invalid periodic bc setup (in {ax}-dir.).'
                raise ValueError(self.error_msg)
        fx = [bc[2 * a] in _FIXED or bc[2 * a + 1] in _FIXED for a in range(3)]
        if (fx[0] and (fx[1] or fx[2])) or (fx[1] and (fx[0] or fx[2])) or (fx[2] and (fx[0] or fx[1])):
            self.error_msg = 'Invalid fixed variable bc setup.'
            raise ValueError(self.error_msg)

    def _face_arrays(self, face):
        name = ('1X', 'X1', '1Y', 'Y1', '1Z', 'Z1')[face]
        return getattr(self, 'bc_data_den_face_' + name), getattr(self, 'bc_data_vel_face_' + name)

    def _set_bc_data(self):
        """isothermal.pyx:476-502, including its face-Z1 condition, which
        tests bc[FACE_1Z] for FIXED_VEL (isothermal.pyx:500)."""
        bc, NI, NJ, NK = self._bc, self.NI, self.NJ, self.NK
        if bc[FACE_1X] in _FIXED:
            self.bc_data_den_face_1X[:, :] = self.den[1, :, :]
            self.bc_data_vel_face_1X[:, :, :] = self.vel[1, :, :, :]
        if bc[FACE_X1] in _FIXED:
            self.bc_data_den_face_X1[:, :] = self.den[NI - 2, :, :]
            self.bc_data_vel_face_X1[:, :, :] = self.vel[NI - 2, :, :, :]
        if bc[FACE_1Y] in _FIXED:
            self.bc_data_den_face_1Y[:, :] = self.den[:, 1, :]
            self.bc_data_vel_face_1Y[:, :, :] = self.vel[:, 1, :, :]
        if bc[FACE_Y1] in _FIXED:
            self.bc_data_den_face_Y1[:, :] = self.den[:, NJ - 2, :]
            self.bc_data_vel_face_Y1[:, :, :] = self.vel[:, NJ - 2, :, :]
        if bc[FACE_1Z] in _FIXED:
            self.bc_data_den_face_1Z[:, :] = self.den[:, :, 1]
            self.bc_data_vel_face_1Z[:, :, :] = self.vel[:, :, 1, :]
        if bc[FACE_Z1] == FIXED_DEN or bc[FACE_1Z] == FIXED_VEL:
            self.bc_data_den_face_Z1[:, :] = self.den[:, :, NK - 2]
            self.bc_data_vel_face_Z1[:, :, :] = self.vel[:, :, NK - 2, :]

    # ------------------------------------------------------------------
    # halo passes of the host fields
    # ------------------------------------------------------------------
    def _enforce_phase_per_sym_bc(self):
        """isothermal.pyx:509-577"""
        p, bc, NI, NJ, NK = self.phase, self._bc, self.NI, self.NJ, self.NK
        if bc[FACE_1X] == PERIODIC:
            p[NI - 1, 1:NJ - 1, 1:NK - 1] = p[1, 1:NJ - 1, 1:NK - 1]
        elif bc[FACE_1X] == SYMMETRIC:
            p[0, 1:NJ - 1, 1:NK - 1] = p[1, 1:NJ - 1, 1:NK - 1]
        if bc[FACE_X1] == PERIODIC:
            p[0, 1:NJ - 1, 1:NK - 1] = p[NI - 2, 1:NJ - 1, 1:NK - 1]
        elif bc[FACE_X1] == SYMMETRIC:
            p[NI - 1, 1:NJ - 1, 1:NK - 1] = p[NI - 2, 1:NJ - 1, 1:NK - 1]
        if bc[FACE_1Y] == PERIODIC:
            p[:, NJ - 1, 1:NK - 1] = p[:, 1, 1:NK - 1]
        elif bc[FACE_1Y] == SYMMETRIC:
            p[:, 0, 1:NK - 1] = p[:, 1, 1:NK - 1]
        if bc[FACE_Y1] == PERIODIC:
            p[:, 0, 1:NK - 1] = p[:, NJ - 2, 1:NK - 1]
        elif bc[FACE_Y1] == SYMMETRIC:
            p[:, NJ - 1, 1:NK - 1] = p[:, NJ - 2, 1:NK - 1]
        if bc[FACE_1Z] == PERIODIC:
            p[:, :, NK - 1] = p[:, :, 1]
        elif bc[FACE_1Z] == SYMMETRIC:
            p[:, :, 0] = p[:, :, 1]
        if bc[FACE_Z1] == PERIODIC:
            p[:, :, 0] = p[:, :, NK - 2]
        elif bc[FACE_Z1] == SYMMETRIC:
            p[:, :, NK - 1] = p[:, :, NK - 2]

    def _enforce_den_vel_per_sym_bc(self):
        """isothermal.pyx:582-686"""
        d, v, bc, NI, NJ, NK = self.den, self.vel, self._bc, self.NI, self.NJ, self.NK
        J, K = slice(1, NJ - 1), slice(1, NK - 1)

        def cp(dst, src, neg):
            d[dst] = d[src]
            v[dst] = v[src]
            if neg is not None:
                v[dst + (neg,)] = -v[src + (neg,)]

        if bc[FACE_1X] == PERIODIC:
            cp((NI - 1, J, K), (1, J, K), None)
        elif bc[FACE_1X] == SYMMETRIC:
            cp((0, J, K), (1, J, K), X)
        if bc[FACE_X1] == PERIODIC:
            cp((0, J, K), (NI - 2, J, K), None)
        elif bc[FACE_X1] == SYMMETRIC:
            cp((NI - 1, J, K), (NI - 2, J, K), X)
        A = slice(None)
        if bc[FACE_1Y] == PERIODIC:
            cp((A, NJ - 1, K), (A, 1, K), None)
        elif bc[FACE_1Y] == SYMMETRIC:
            cp((A, 0, K), (A, 1, K), Y)
        if bc[FACE_Y1] == PERIODIC:
            cp((A, 0, K), (A, NJ - 2, K), None)
        elif bc[FACE_Y1] == SYMMETRIC:
            cp((A, NJ - 1, K), (A, NJ - 2, K), Y)
        if bc[FACE_1Z] == PERIODIC:
            cp((A, A, NK - 1), (A, A, 1), None)
        elif bc[FACE_1Z] == SYMMETRIC:
            cp((A, A, 0), (A, A, 1), Z)
        if bc[FACE_Z1] == PERIODIC:
            cp((A, A, 0), (A, A, NK - 2), None)
        elif bc[FACE_Z1] == SYMMETRIC:
            cp((A, A, NK - 1), (A, A, NK - 2), Z)

    # ------------------------------------------------------------------
    # host <-> device hand-off
    # ------------------------------------------------------------------
    def _interior(self, a):
        return a[1:self.NI - 1, 1:self.NJ - 1, 1:self.NK - 1]

    def _sync_inputs(self):
        """Mutable inputs of evolve that live outside the bound host fields:
        gravity and the Zou-He face data (isothermal.pyx:476-502, 1293-1295).
        The acceleration field, the wall velocities and the densities that
        Ladd's term reads in the first step come from the host fields
        themselves, which the engine reads at every evolve call."""
        e = self._eng
        e.set_gravity(self._g)
        for f in range(6):
            if self._bc[f] in _FIXED:
                e.set_face_data(f, *self._face_arrays(f))

    def _load_inputs(self):
        """_sync_inputs plus the host fields as the inputs of the next steps
        (the set-up of an engine-level run such as the adjoint)."""
        self._sync_inputs()
        self._eng.load_host_fields()

    def _pull_fields(self, halo_pass=True):
        """den/vel of the fluid sites after the last device step into the host
        fields, then (halo_pass) the den/vel halo pass."""
        self._eng.store_host_fields(halo_pass)

    # ------------------------------------------------------------------
    # update methods
    # ------------------------------------------------------------------
    def finalize_init(self):
        """isothermal.pyx:1333-1365"""
        if self.error_occured():
            raise RuntimeError('Invalid solver setup: finalization aborted!')
        phaux = self._interior(self.phase)
        n = int(np.count_nonzero(phaux))
        if n >= 1:
            self.fluid_lsite_domain = np.array(np.nonzero(phaux), dtype=np.uint32)
            self._fluid = phaux != SOLID
            self._enforce_phase_per_sym_bc()
            self._enforce_den_vel_per_sym_bc()
            e = self._eng
            e.upload_phase(phaux)
            e.upload_den_vel(self._interior(self.den), self._interior(self.vel))
            n_dev = e.finalize()
            if n_dev != n:
                raise RuntimeError(f'device fluid-site count {n_dev} != {n}')
            e.bind_host_fields(self._phase, self._den, self._vel, self._acc)
        self.init_done = 1
        return n

    def evolve(self, tsteps):
        """Execute `tsteps` time steps; returns the total step count
        (isothermal.pyx:1370-1418)."""
        if self.error_occured():
            raise RuntimeError('Invalid solver setup: evolve aborted!')
        self._set_bc_data()
        if self.init_done == 0:
            raise RuntimeError('Finalize solver initialization before calling evolve\n'
                               '(i.e. call first the method finalize_init)!')
        if self.fluid_lsite_domain is None or self.fluid_lsite_domain.shape[1] < 1:
            return self.tot_tsteps
        tsteps = int(tsteps) & 0xFFFFFFFF
        if tsteps > 0:
            self._sync_inputs()
            # On the device: the host fields are read as they are now (acc,
            # wall velocities; a phase edit since finalize_init raises), the
            # steps run, and den/vel of the fluid sites come back in place,
            # with the halo den/vel of step T-1 as the reference leaves them
            # (isothermal.pyx:1408-1415).
            self._eng.evolve_host(tsteps)
        self.tot_tsteps = (self.tot_tsteps + tsteps) & 0xFFFFFFFF
        return self.tot_tsteps

    # ------------------------------------------------------------------
    # status, material relations, getters
    # ------------------------------------------------------------------
    def error_occured(self):
        return len(self.error_msg)

    def eos_den2p(self, den):
        """isothermal.pyx:1431-1443"""
        return self.dvs.INV_AS2 * den

    def eos_p2den(self, p):
        """isothermal.pyx:1446-1458"""
        return self.dvs.AS2 * p

    def get_phase(self):
        return self.phase[1:self.NI - 1, 1:self.NJ - 1, 1:self.NK - 1]

    def get_acc(self):
        return self.acc[1:self.NI - 1, 1:self.NJ - 1, 1:self.NK - 1, :]

    def get_vel(self):
        return self.vel[1:self.NI - 1, 1:self.NJ - 1, 1:self.NK - 1, :]

    def get_den(self):
        return self.den[1:self.NI - 1, 1:self.NJ - 1, 1:self.NK - 1]

    def get_fluid_lat_sites(self):
        return self.fluid_lsite_domain

    # ------------------------------------------------------------------
    # additions (not in the reference)
    # ------------------------------------------------------------------
    def post_collision_f(self):
        """Newest post-collision populations, (19, NI, NJ, NK) f64 with zero
        halo and SOLID entries: the reference's f1 in lattice-vector order."""
        out = np.zeros((LVECS, self.NI, self.NJ, self.NK))
        if self.init_done and self._fluid is not None:
            out[:, 1:-1, 1:-1, 1:-1] = self._eng.download_post_f()
        return out

    def flow_info(self):
        """calc_flow_info (include/solver/solver.cpp:64-129) on the GPU, in
        FP64, over the fluid sites of the last step: dict with the fields of
        IsothermalFlowInfo (include/solver/solver.h:13-15) -- tot_jx/jy/jz,
        tot_j2 (sum den |u|), tot_den, ave_ux/uy/uz, ave_u2 (mean |u|), max_u2
        (max |u|), max_den, min_den (extrema starting from rho0, as in the
        reference) -- and fluid_sites."""
        return self._eng.flow_info()

    def convergence(self):
        """Relative L2 change of the velocity field since the previous call
        (FP64 on the GPU); 1.0 on the first call."""
        return self._eng.convergence()

    @property
    def engine(self):
        """The device engine (for device-resident loops and tensor hand-off)."""
        return self._eng
