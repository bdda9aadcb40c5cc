"""ctypes binding of libjyulb_hip.so (C ABI: include/jyulb_gpu.h).

The library is built in-tree by jyu-lb_amd/build.py. There is no fallback:
if the library or a HIP device is missing, using the solver raises.
"""
import ctypes
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libjyulb_hip.so")

LB_OK, LB_EINVAL, LB_ESTATE, LB_EHIP, LB_ENOMEM = 0, 1, 2, 3, 4
LB_FP64, LB_FP32_SHIFTED, LB_FP16_STORAGE, LB_FP32_STORAGE = 0, 1, 2, 3

PRECISIONS = {"fp64": LB_FP64, "fp32": LB_FP32_SHIFTED, "fp32_shifted": LB_FP32_SHIFTED,
              "fp16": LB_FP16_STORAGE, "fp16_storage": LB_FP16_STORAGE, "fp32_storage": LB_FP32_STORAGE,
              LB_FP64: LB_FP64, LB_FP32_SHIFTED: LB_FP32_SHIFTED, LB_FP16_STORAGE: LB_FP16_STORAGE,
              LB_FP32_STORAGE: LB_FP32_STORAGE}


LB_LAYOUT_ROWS, LB_LAYOUT_SOA, LB_LAYOUT_SOA_PADDED, LB_LAYOUT_SPARSE, LB_LAYOUT_AUTO = 0, 1, 2, 3, 4
LAYOUTS = {"rows": LB_LAYOUT_ROWS, "soa": LB_LAYOUT_SOA, "soa_padded": LB_LAYOUT_SOA_PADDED,
           "sparse": LB_LAYOUT_SPARSE, "auto": LB_LAYOUT_AUTO}
LAYOUT_NAMES = {v: k for k, v in LAYOUTS.items()}
LAYOUTS.update({v: v for v in list(LAYOUTS.values())})


class LbConfig(ctypes.Structure):
    _fields_ = [
        ("lat", ctypes.c_int32 * 3),
        ("rlx_scheme", ctypes.c_int32),
        ("kvisc", ctypes.c_double),
        ("g", ctypes.c_double * 3),
        ("bc", ctypes.c_int32 * 6),
        ("precision", ctypes.c_int32),
        ("rho0", ctypes.c_double),
        ("fp16_scale", ctypes.c_double),
        ("layout", ctypes.c_int32),
    ]


class LbFlowInfo(ctypes.Structure):
    """lb_flow_info_t: the fields of the reference's IsothermalFlowInfo
    (include/solver/solver.h:13-15) plus the number of fluid sites."""
    _fields_ = [(n, ctypes.c_double) for n in (
        "tot_jx", "tot_jy", "tot_jz", "tot_j2", "tot_den", "ave_ux", "ave_uy", "ave_uz", "ave_u2",
        "max_u2", "max_den", "min_den", "fluid_sites")]


_lib = None


def lib():
    """Load the engine; raises ImportError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not found: build it with `python jyu-lb_amd/build.py`")
    # One HIP runtime per process: torch ships its own libamdhip64 with the
    # same soname as ROCm's. Loaded first, it is the one the engine binds to,
    # so torch streams and device pointers are valid handles here; loaded
    # after the engine, a second runtime would start and torch would see no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    H = ctypes.c_void_p
    dp = ctypes.POINTER(ctypes.c_double)
    sig = {
        "lb_last_error": ([], ctypes.c_char_p),
        "lb_device_count": ([], ctypes.c_int),
        "lb_create": ([ctypes.POINTER(LbConfig), ctypes.POINTER(H)], ctypes.c_int),
        "lb_destroy": ([H], ctypes.c_int),
        "lb_set_stream": ([H, ctypes.c_void_p], ctypes.c_int),
        "lb_set_gravity": ([H, dp], ctypes.c_int),
        "lb_upload_phase": ([H, ctypes.c_void_p], ctypes.c_int),
        "lb_upload_den_vel": ([H, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
        "lb_upload_acc": ([H, ctypes.c_void_p], ctypes.c_int),
        "lb_set_face_data": ([H, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
        "lb_finalize": ([H, ctypes.POINTER(ctypes.c_int64)], ctypes.c_int),
        "lb_evolve": ([H, ctypes.c_uint32], ctypes.c_int),
        "lb_refresh_walls": ([H], ctypes.c_int),
        "lb_download_den_vel": ([H, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
        "lb_download_post_f": ([H, ctypes.c_void_p], ctypes.c_int),
        "lb_keep_previous": ([H, ctypes.c_int], ctypes.c_int),
        "lb_download_previous_post_f": ([H, ctypes.c_void_p], ctypes.c_int),
        "lb_reduce": ([H, dp], ctypes.c_int),
        "lb_convergence": ([H, dp], ctypes.c_int),
        "lb_upload_acc_device": ([H, ctypes.c_void_p], ctypes.c_int),
        "lb_download_den_vel_device": ([H, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
        "lb_info": ([H, ctypes.POINTER(ctypes.c_uint64), ctypes.POINTER(ctypes.c_int),
                     ctypes.POINTER(ctypes.c_uint64), dp], ctypes.c_int),
        "lb_sparse_codes": ([H, ctypes.POINTER(ctypes.c_uint64)], ctypes.c_int),
        "lb_pull_steps": ([H, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "lb_synchronize": ([H], ctypes.c_int),
        "lb_set_face_data_device": ([H, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
        "lb_objective": ([H, ctypes.c_int, ctypes.c_int, dp], ctypes.c_int),
        "lb_adj_forward": ([H, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_int, ctypes.c_int, dp], ctypes.c_int),
        "lb_adj_backward": ([H, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
        "lb_adj_backward_device": ([H, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p], ctypes.c_int),
        "lb_snapshot": ([H], ctypes.c_int),
        "lb_restore": ([H], ctypes.c_int),
        "lb_flow_info": ([H, ctypes.POINTER(LbFlowInfo)], ctypes.c_int),
        "lb_host_alloc": ([ctypes.c_size_t, ctypes.POINTER(ctypes.c_void_p)], ctypes.c_int),
        "lb_host_free": ([ctypes.c_void_p], ctypes.c_int),
        "lb_bind_host_fields": ([H, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p],
                                ctypes.c_int),
        "lb_load_host_fields": ([H], ctypes.c_int),
        "lb_evolve_host": ([H, ctypes.c_uint32], ctypes.c_int),
        "lb_store_host_fields": ([H, ctypes.c_int], ctypes.c_int),
        "lb_layout": ([H, ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
        "lb_stream_copy": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p], ctypes.c_int),
        "lb_flow_info_fields_noalloc": ([ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                                         ctypes.POINTER(ctypes.c_int32), ctypes.c_double, ctypes.c_void_p,
                                         ctypes.c_void_p], ctypes.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc):
    """Map a C status code to the reference's exception types."""
    if rc == LB_OK:
        return
    msg = lib().lb_last_error().decode(errors="replace")
    if rc == LB_EINVAL:
        raise ValueError(msg)
    raise RuntimeError(msg)


def ptr(a):
    return ctypes.c_void_p(a.ctypes.data)


def stream_copy(dst_ptr, src_ptr, nbytes, stream=0):
    """lb_stream_copy: device-to-device copy on `stream` (device pointers)."""
    check(lib().lb_stream_copy(ctypes.c_void_p(dst_ptr), ctypes.c_void_p(src_ptr), ctypes.c_size_t(int(nbytes)),
                               ctypes.c_void_p(stream)))


class _Pinned:
    """Owner of one pinned host allocation (lb_host_alloc); numpy arrays made
    from it keep it alive through their base, and it is freed with the last."""

    def __init__(self, nbytes):
        p = ctypes.c_void_p()
        check(lib().lb_host_alloc(ctypes.c_size_t(int(nbytes)), ctypes.byref(p)))
        self.ptr = p.value
        self.__array_interface__ = {"shape": (int(nbytes),), "typestr": "|u1", "data": (self.ptr, False),
                                    "version": 3}

    def __del__(self):
        if getattr(self, "ptr", None):
            try:
                lib().lb_host_free(ctypes.c_void_p(self.ptr))
            except Exception:
                pass
            self.ptr = None


def host_array(shape, dtype, fill):
    """C-contiguous array of `shape` filled with `fill`, in pinned host memory
    when it can be had (plain numpy memory otherwise: same results, slower
    copies)."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape)) * dtype.itemsize
    try:
        a = np.asarray(_Pinned(max(n, 1)))[:n].view(dtype).reshape(shape)
    except (RuntimeError, ValueError, MemoryError):
        return np.full(shape, fill, dtype=dtype)
    a[...] = fill
    return a


class Engine:
    """Owner of one device solver handle (lb_solver*)."""

    def __init__(self, lat, rlx, kvisc, g, bc, precision=LB_FP64, rho0=1.0, fp16_scale=0.0, layout=LB_LAYOUT_AUTO):
        if precision not in PRECISIONS:
            raise ValueError(f"unknown precision {precision!r}; one of "
                             f"{sorted(k for k in PRECISIONS if isinstance(k, str))}")
        if layout not in LAYOUTS:
            raise ValueError(f"unknown layout {layout!r}; one of {sorted(k for k in LAYOUTS if isinstance(k, str))}")
        L = lib()
        if L.lb_device_count() < 1:
            raise RuntimeError("jyulb: no HIP device available (the evolve path runs on the GPU only)")
        cfg = LbConfig()
        for a in range(3):
            cfg.lat[a] = int(lat[a])
            cfg.g[a] = float(g[a])
        for f in range(6):
            cfg.bc[f] = int(bc[f])
        cfg.rlx_scheme = int(rlx)
        cfg.kvisc = float(kvisc)
        cfg.precision = int(PRECISIONS[precision])
        cfg.rho0 = float(rho0)
        cfg.fp16_scale = float(fp16_scale)
        cfg.layout = int(LAYOUTS[layout])
        h = ctypes.c_void_p()
        check(L.lb_create(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        self.lat = tuple(int(v) for v in lat)
        self.precision = cfg.precision

    def close(self):
        if getattr(self, "h", None) is not None and self.h.value:
            lib().lb_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def call(self, name, *args):
        check(getattr(lib(), name)(self.h, *args))

    def set_stream(self, stream_handle):
        self.call("lb_set_stream", ctypes.c_void_p(stream_handle))

    def set_gravity(self, g):
        a = np.ascontiguousarray(g[:3], dtype=np.float64)
        self.call("lb_set_gravity", a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))

    def upload_phase(self, phase):
        a = np.ascontiguousarray(phase, dtype=np.uint8)
        self.call("lb_upload_phase", ptr(a))

    def upload_den_vel(self, den, vel):
        d = np.ascontiguousarray(den, dtype=np.float64)
        v = np.ascontiguousarray(vel, dtype=np.float64)
        self.call("lb_upload_den_vel", ptr(d), ptr(v))

    def upload_acc(self, acc):
        if acc is None:
            self.call("lb_upload_acc", None)
        else:
            a = np.ascontiguousarray(acc, dtype=np.float64)
            self.call("lb_upload_acc", ptr(a))

    def set_face_data(self, face, den_face, vel_face):
        d = np.ascontiguousarray(den_face, dtype=np.float64)
        v = np.ascontiguousarray(vel_face, dtype=np.float64)
        self.call("lb_set_face_data", int(face), ptr(d), ptr(v))

    def finalize(self):
        n = ctypes.c_int64()
        self.call("lb_finalize", ctypes.byref(n))
        return int(n.value)

    def evolve(self, steps):
        self.call("lb_evolve", ctypes.c_uint32(int(steps)))

    def refresh_walls(self):
        self.call("lb_refresh_walls")

    def download_den_vel(self):
        nx, ny, nz = self.lat
        den = np.empty((nx, ny, nz))
        vel = np.empty((nx, ny, nz, 3))
        self.call("lb_download_den_vel", ptr(den), ptr(vel))
        return den, vel

    def download_den_vel_device(self, den_ptr, vel_ptr):
        """den/vel into device arrays (interior C order), on the handle's stream."""
        self.call("lb_download_den_vel_device", ctypes.c_void_p(den_ptr), ctypes.c_void_p(vel_ptr))

    def download_post_f(self):
        nx, ny, nz = self.lat
        f = np.empty((19, nx, ny, nz))
        self.call("lb_download_post_f", ptr(f))
        return f

    def keep_previous(self, on=True):
        """Keep the state before the last step of every evolve call (lb_keep_previous)."""
        self.call("lb_keep_previous", int(bool(on)))

    def download_previous_post_f(self):
        """Post-collision populations of the step before the last one (f2), or
        None if no evolve call kept them."""
        nx, ny, nz = self.lat
        f = np.empty((19, nx, ny, nz))
        if lib().lb_download_previous_post_f(self.h, ptr(f)) != LB_OK:
            return None
        return f

    def flow_info(self):
        """calc_flow_info on the device (lb_flow_info): dict of the
        IsothermalFlowInfo fields plus fluid_sites."""
        info = LbFlowInfo()
        self.call("lb_flow_info", ctypes.byref(info))
        return {n: float(getattr(info, n)) for n, _ in LbFlowInfo._fields_}

    def bind_host_fields(self, phase, den, vel, acc):
        """Bind the Solver's halo-padded host arrays (kept referenced here, so
        that they outlive the binding)."""
        for a, dt in ((phase, np.uint8), (den, np.float64), (vel, np.float64), (acc, np.float64)):
            if a.dtype != dt or not a.flags.c_contiguous:
                raise ValueError("host fields must be C-contiguous arrays of the Solver's dtypes")
        self.call("lb_bind_host_fields", ptr(phase), ptr(den), ptr(vel), ptr(acc))
        self._bound = (phase, den, vel, acc)

    def load_host_fields(self):
        self.call("lb_load_host_fields")

    def evolve_host(self, steps):
        self.call("lb_evolve_host", ctypes.c_uint32(int(steps)))

    def store_host_fields(self, halo_pass=True):
        self.call("lb_store_host_fields", int(bool(halo_pass)))

    def reduce(self):
        out = np.zeros(8)
        self.call("lb_reduce", out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        return out

    def convergence(self):
        out = ctypes.c_double()
        self.call("lb_convergence", ctypes.byref(out))
        return out.value

    def info(self):
        st, odd, by, sc = ctypes.c_uint64(), ctypes.c_int(), ctypes.c_uint64(), ctypes.c_double()
        self.call("lb_info", ctypes.byref(st), ctypes.byref(odd), ctypes.byref(by), ctypes.byref(sc))
        return dict(steps=st.value, next_odd=odd.value, device_bytes=by.value, fp16_scale=sc.value)

    def sparse_codes(self):
        """Odd-step links of the sparse layout by code kind: near offsets, the
        wave's alternate location, bounce-back, full 32-bit locations."""
        c = (ctypes.c_uint64 * 4)()
        self.call("lb_sparse_codes", c)
        return dict(zip(("near", "alt", "bounce", "far"), (int(x) for x in c)))

    def pull_steps(self):
        """True when the solver steps by pull steps (a second population array)."""
        v = ctypes.c_int()
        self.call("lb_pull_steps", ctypes.byref(v))
        return bool(v.value)

    def layout(self):
        """Name of the population layout in use ("auto" until finalize resolves it)."""
        v = ctypes.c_int()
        self.call("lb_layout", ctypes.byref(v))
        return LAYOUT_NAMES[v.value]

    def synchronize(self):
        self.call("lb_synchronize")

    def face_shape(self, face):
        nx, ny, nz = self.lat
        return (ny + 2, nz + 2) if face < 2 else (nx + 2, nz + 2) if face < 4 else (nx + 2, ny + 2)

    def set_face_data_device(self, face, den_ptr, vel_ptr):
        self.call("lb_set_face_data_device", int(face), ctypes.c_void_p(den_ptr), ctypes.c_void_p(vel_ptr))

    def objective(self, face, axis):
        out = ctypes.c_double()
        self.call("lb_objective", int(face), int(axis), ctypes.byref(out))
        return out.value

    adj_gen = 0  # number of adjoint forward passes run on this handle

    def adj_forward(self, steps, interval, face, axis):
        self.adj_gen += 1
        out = ctypes.c_double()
        self.call("lb_adj_forward", ctypes.c_uint32(int(steps)), ctypes.c_uint32(int(interval)),
                  int(face), int(axis), ctypes.byref(out))
        return out.value

    def adj_backward(self, grad_face=-1):
        """Host gradients: (grad_acc (nx,ny,nz,3), grad_face_vel (face shape,3) or None)."""
        nx, ny, nz = self.lat
        ga = np.empty((nx, ny, nz, 3))
        gf = np.empty(self.face_shape(grad_face) + (3,)) if grad_face >= 0 else None
        self.call("lb_adj_backward", int(grad_face), ptr(ga), ptr(gf) if gf is not None else None)
        return ga, gf

    def adj_backward_device(self, grad_face, ga_ptr, gf_ptr):
        self.call("lb_adj_backward_device", int(grad_face), ctypes.c_void_p(ga_ptr),
                  ctypes.c_void_p(gf_ptr) if gf_ptr else None)

    def upload_acc_device(self, acc_ptr):
        self.call("lb_upload_acc_device", ctypes.c_void_p(acc_ptr))

    def snapshot(self):
        self.call("lb_snapshot")

    def restore(self):
        self.call("lb_restore")
