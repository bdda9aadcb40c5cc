"""Python face of the CPU FP64 oracle -- TEST INFRASTRUCTURE, never the product.

`OracleSolver` mirrors the public surface of the reference's
`jyulb.isothermal.Solver` (isothermal.pyx:156-1527) -- constructor checks,
halo-padded arrays with aliasing interior views, finalize_init, evolve,
gravity -- and runs the step loop in the C restatement `jyulb_oracle.c`
(built by oracle/Makefile into oracle/liboracle.so). Only tests, the smoke
check and bench.py's CPU-baseline leg use it, always as the checker.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

X, Y, Z = 0, 1, 2
WALL, PERIODIC, FIXED_DEN, FIXED_VEL, SYMMETRIC = 0, 1, 2, 3, 4
BGK, TRT, REGULARIZATION = 0, 1, 2
SOLID = 0


class _State(ctypes.Structure):
    _fields_ = [
        ("NI", ctypes.c_int32), ("NJ", ctypes.c_int32), ("NK", ctypes.c_int32),
        ("rlx", ctypes.c_int32), ("bc", ctypes.c_int32 * 6),
        ("g", ctypes.c_double * 3), ("prm", ctypes.c_double * 2),
        ("phase", ctypes.c_void_p), ("den", ctypes.c_void_p), ("vel", ctypes.c_void_p),
        ("acc", ctypes.c_void_p), ("f1", ctypes.c_void_p), ("f2", ctypes.c_void_p),
        ("bcden", ctypes.c_void_p * 6), ("bcvel", ctypes.c_void_p * 6),
        ("nfluid", ctypes.c_int64), ("sites", ctypes.c_void_p),
    ]


def lib():
    """Load oracle/liboracle.so (building it with make if it is missing)."""
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-C", _HERE, "liboracle.so"],
                                  stdout=subprocess.DEVNULL)
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(_State)
        for fn in ("orc_set_bc_data", "orc_enforce_phase", "orc_enforce_den_vel",
                   "orc_enforce_distr", "orc_init_f", "orc_update"):
            getattr(L, fn).argtypes = [P]
            getattr(L, fn).restype = None
        L.orc_evolve_steps.argtypes = [P, ctypes.c_uint32]
        L.orc_evolve_steps.restype = None
        dp = ctypes.POINTER(ctypes.c_double)
        L.orc_project.argtypes = [dp, dp]
        L.orc_f_to_h.argtypes = [dp, dp]
        L.orc_collide.argtypes = [ctypes.c_int, dp, dp, dp, dp]
        L.orc_collide.restype = ctypes.c_double
        L.orc_zouhe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, dp, dp]
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class OracleSolver:
    """CPU restatement of jyulb.isothermal.Solver (isothermal.pyx:156-298)."""

    def __init__(self, lat_size, rlx_scheme=BGK, kvisc=1.0 / 6.0, g=None, bc=None):
        self.error_msg = ""
        lat_size = np.asarray(lat_size)
        # argument checks, isothermal.pyx:196-206
        if lat_size[X] < 1 or lat_size[Y] < 1 or lat_size[Z] < 1:
            self.error_msg = "Invalid lattice size input."
            raise ValueError(self.error_msg)
        if rlx_scheme not in (BGK, TRT, REGULARIZATION):
            self.error_msg = "Invalid relaxation scheme input."
            raise ValueError(self.error_msg)
        if kvisc <= 0.0:
            self.error_msg = "Invalid kinematic viscosity input."
            raise ValueError(self.error_msg)
        self.g = np.zeros(3) if g is None else np.array([g[0], g[1], g[2]], dtype=np.float64)
        self.bc = np.full(6, PERIODIC, dtype=np.int32) if bc is None else np.array(bc[:6], dtype=np.int32)
        NI, NJ, NK = int(lat_size[X]) + 2, int(lat_size[Y]) + 2, int(lat_size[Z]) + 2
        self.NI, self.NJ, self.NK, self.NCNT = NI, NJ, NK, NI * NJ * NK
        # fields, isothermal.pyx:243-248
        self.phase = np.full((NI, NJ, NK), SOLID, dtype=np.uint8)
        self.den = np.ones((NI, NJ, NK))
        self.vel = np.zeros((NI, NJ, NK, 3))
        self.acc = np.zeros((NI, NJ, NK, 3))
        # face data, isothermal.pyx:253-263
        fs = [(NJ, NK), (NJ, NK), (NI, NK), (NI, NK), (NI, NJ), (NI, NJ)]
        self.bc_den = [np.ones(s) for s in fs]
        self.bc_vel = [np.zeros(s + (3,)) for s in fs]
        self.f1 = np.zeros(19 * self.NCNT)
        self.f2 = np.zeros(19 * self.NCNT)
        # relaxation parameters, isothermal.pyx:273-287
        self.kvisc = kvisc
        inv_tau_e = 1.0 / (0.5 + 3.0 * kvisc)
        self.rlx_scheme = rlx_scheme
        self.rlx_prm = np.zeros(2)
        self.rlx_prm[0] = inv_tau_e
        if rlx_scheme == TRT:
            self.rlx_prm[1] = 8.0 * (2.0 - inv_tau_e) / (8.0 - inv_tau_e)
        self.tot_tsteps = 0
        self.init_done = 0
        self.fluid_lsite_domain = None

    def _state(self):
        s = _State()
        s.NI, s.NJ, s.NK = self.NI, self.NJ, self.NK
        s.rlx = self.rlx_scheme
        for f in range(6):
            s.bc[f] = int(self.bc[f])
            s.bcden[f] = _ptr(self.bc_den[f])
            s.bcvel[f] = _ptr(self.bc_vel[f])
        for d in range(3):
            s.g[d] = float(self.g[d])
        s.prm[0], s.prm[1] = float(self.rlx_prm[0]), float(self.rlx_prm[1])
        s.phase, s.den, s.vel, s.acc = _ptr(self.phase), _ptr(self.den), _ptr(self.vel), _ptr(self.acc)
        s.f1, s.f2 = _ptr(self.f1), _ptr(self.f2)
        if self.fluid_lsite_domain is not None:
            s.nfluid = self.fluid_lsite_domain.shape[1]
            s.sites = _ptr(self.fluid_lsite_domain)
        return s

    @property
    def gravity(self):
        return self.g

    @gravity.setter
    def gravity(self, g):
        self.g[X], self.g[Y], self.g[Z] = g[X], g[Y], g[Z]

    def error_occured(self):
        return len(self.error_msg)

    def eos_den2p(self, den):
        return den / 3.0

    def eos_p2den(self, p):
        return 3.0 * p

    def get_phase(self):
        return self.phase[1:-1, 1:-1, 1:-1]

    def get_den(self):
        return self.den[1:-1, 1:-1, 1:-1]

    def get_vel(self):
        return self.vel[1:-1, 1:-1, 1:-1, :]

    def get_acc(self):
        return self.acc[1:-1, 1:-1, 1:-1, :]

    def get_fluid_lat_sites(self):
        return self.fluid_lsite_domain

    def finalize_init(self):
        """Solver.finalize_init, isothermal.pyx:1333-1365"""
        if self.error_occured():
            raise RuntimeError("Invalid solver setup: finalization aborted!")
        ph = self.get_phase()
        n = int(np.count_nonzero(ph))
        if n >= 1:
            self.fluid_lsite_domain = np.ascontiguousarray(np.array(np.nonzero(ph), dtype=np.uint32))
            s = self._state()
            L = lib()
            L.orc_enforce_phase(ctypes.byref(s))
            L.orc_enforce_den_vel(ctypes.byref(s))
            L.orc_init_f(ctypes.byref(s))
        self.init_done = 1
        return n

    def evolve(self, tsteps):
        """Solver.evolve, isothermal.pyx:1370-1418"""
        if self.error_occured():
            raise RuntimeError("Invalid solver setup: evolve aborted!")
        s = self._state()
        L = lib()
        L.orc_set_bc_data(ctypes.byref(s))
        if self.init_done == 0:
            raise RuntimeError("Finalize solver initialization before calling evolve\n"
                               "(i.e. call first the method finalize_init)!")
        if self.fluid_lsite_domain is None or self.fluid_lsite_domain.shape[1] < 1:
            return self.tot_tsteps
        L.orc_evolve_steps(ctypes.byref(s), int(tsteps))
        if tsteps % 2 == 1:
            self.f1, self.f2 = self.f2, self.f1
        self.tot_tsteps += int(tsteps)
        return self.tot_tsteps

    def post_collision_f(self):
        """f1 as (19, NI, NJ, NK): the newest post-collision populations."""
        return self.f1.reshape(19, self.NI, self.NJ, self.NK).copy()


# ---- single-site helpers for closed-form tests ------------------------------

def project(h):
    f = np.zeros(19)
    lib().orc_project(_dp(np.ascontiguousarray(h, dtype=np.float64)), _dp(f))
    return f


def f_to_h(f):
    h = np.zeros(10)
    lib().orc_f_to_h(_dp(np.ascontiguousarray(f, dtype=np.float64)), _dp(h))
    return h


def collide(rlx, f, acc, prm):
    f = np.array(f, dtype=np.float64)
    vel = np.zeros(3)
    den = lib().orc_collide(rlx, _dp(f), _dp(np.asarray(acc, dtype=np.float64)),
                            _dp(np.asarray(prm, dtype=np.float64)), _dp(vel))
    return den, vel, f


def zouhe(face, kind, den, vel, f):
    f = np.array(f, dtype=np.float64)
    lib().orc_zouhe(face, kind, den, _dp(np.asarray(vel, dtype=np.float64)), _dp(f))
    return f
