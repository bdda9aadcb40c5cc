"""The allocation-free entry point lb_flow_info_fields_noalloc against
lb_flow_info, through the C ABI: the same calc_flow_info
(include/solver/solver.cpp:64-129) of the same fields, one computed by the
solver handle on its own fields (library-allocated partials), the other on
device copies of them with the partials in the library's fixed scratch
array. Two calls in a row, the second on a smaller lattice, so that the
partials the first call left in the array must not leak into the second.
Sums run in a different order (interior C order against padded rows):
1e-12 relative."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELDS = ("tot_jx", "tot_jy", "tot_jz", "tot_j2", "tot_den", "ave_ux", "ave_uy", "ave_uz", "ave_u2",
          "max_u2", "max_den", "min_den", "fluid_sites")


def _noalloc_flow_info(s, stream):
    import torch
    from jyulb import _native
    eng = s.engine
    nx, ny, nz = eng.lat
    den = torch.empty((nx, ny, nz), dtype=torch.float64, device="cuda")
    vel = torch.empty((nx, ny, nz, 3), dtype=torch.float64, device="cuda")
    eng.download_den_vel_device(den.data_ptr(), vel.data_ptr())
    phase = torch.tensor(np.ascontiguousarray(s.get_phase()), dtype=torch.uint8, device="cuda")
    out = torch.full((13,), float("nan"), dtype=torch.float64, device="cuda")
    lat = (ctypes.c_int32 * 3)(nx, ny, nz)
    _native.check(_native.lib().lb_flow_info_fields_noalloc(
        ctypes.c_void_p(phase.data_ptr()), ctypes.c_void_p(den.data_ptr()), ctypes.c_void_p(vel.data_ptr()),
        lat, ctypes.c_double(1.0), ctypes.c_void_p(out.data_ptr()), ctypes.c_void_p(stream)))
    torch.cuda.synchronize()
    return dict(zip(FIELDS, out.cpu().numpy().tolist()))


def test_flow_info_fields_noalloc_matches_flow_info():
    import torch
    import bench
    stream = torch.cuda.current_stream().cuda_stream
    for make, lat, steps in ((bench.porous_channel, (40, 24, 30), 120), (bench.porous_medium, (20, 12, 10), 60)):
        print(f"no-alloc flow info on {make.__name__} {lat}, {steps} steps", flush=True)
        s, nfl = make("fp64", lat=lat) if make is bench.porous_channel else make("fp64", lat=lat, radius=(2, 4))
        s.engine.set_stream(stream)
        s.evolve(steps)
        ref = s.flow_info()
        got = _noalloc_flow_info(s, stream)
        assert got["fluid_sites"] == nfl == ref["fluid_sites"]
        for k in FIELDS:
            scale = max(abs(ref[k]), abs(ref["tot_j2"]) if k.startswith("tot_j") else abs(ref["ave_u2"])
                        if k.startswith("ave_u") else 0.0)
            assert abs(got[k] - ref[k]) <= 1e-12 * scale, (lat, k, got[k], ref[k])
        del s


def test_flow_info_fields_noalloc_refuses_oversized_lattice():
    from jyulb import _native
    lat = (ctypes.c_int32 * 3)(2048, 1024, 1024)
    rc = _native.lib().lb_flow_info_fields_noalloc(ctypes.c_void_p(1), ctypes.c_void_p(1), ctypes.c_void_p(1), lat,
                                                   ctypes.c_double(1.0), ctypes.c_void_p(1), None)
    assert rc == _native.LB_EINVAL
    assert "exceed its limit of 2^31 - 1" in _native.lib().lb_last_error().decode()
