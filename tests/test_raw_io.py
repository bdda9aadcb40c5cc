"""Raw-format lattice files of the reference's C++ front-end (jyulb.raw_io).

The geometry fixtures are two of the reference's example inputs
(examples_cpp/isothermal/channel/channel_x100_y12_z24.geom.in.raw, 28,800
bytes, and examples_cpp/isothermal/tube/tube_x23_y23_z101.geom.in.raw, 53,429
bytes). Their shapes pin the node order: the channel's walls are the two z
layers, the tube runs along z with a cross-section symmetric in x and y; a
transposed or y-fastest reading puts solid nodes elsewhere. The GPU test runs
the reference's tube example setup (examples_cpp/isothermal/tube/tube.in:
TRT, nu = 0.15, gz = 1e-5, walls on the x faces, periodic y and z) through the
HIP path and compares it with the CPU oracle.
"""
import os

import numpy as np
import pytest

from jyulb import raw_io
from jyulb.defs import SOLID, FLUID

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_channel_geometry():
    base = os.path.join(GOLD, "channel_x100_y12_z24")
    ph = raw_io.read_geom(base, (100, 12, 24))
    assert ph.shape == (100, 12, 24)
    assert int((ph == FLUID).sum()) == 26400
    assert (ph[:, :, 0] == SOLID).all() and (ph[:, :, -1] == SOLID).all()
    assert (ph[:, :, 1:-1] == FLUID).all()
    with pytest.raises(ValueError):
        raw_io.read_geom(base, (100, 12, 25))


def test_tube_geometry():
    base = os.path.join(GOLD, "tube_x23_y23_z101")
    ph = raw_io.read_geom(base, (23, 23, 101))
    assert int((ph == FLUID).sum()) == 35249 == 101 * 349
    sec = ph[:, :, 0]
    assert all((ph[:, :, k] == sec).all() for k in range(101))
    np.testing.assert_array_equal(sec, sec[::-1, :])
    np.testing.assert_array_equal(sec, sec[:, ::-1])
    np.testing.assert_array_equal(sec, sec.T)
    assert sec[11, 11] == FLUID and sec[0, 0] == SOLID


def test_round_trips_and_byte_layout(tmp_path):
    rng = np.random.default_rng(0)
    lat = (5, 4, 3)
    ph = (rng.random(lat) < 0.7).astype(np.uint8)
    base = os.fspath(tmp_path / "case")
    raw_io.write_geom(base, ph, suffix=".geom.in.raw")
    np.testing.assert_array_equal(raw_io.read_geom(base, lat), ph)
    raw = np.fromfile(base + ".geom.in.raw", dtype=np.uint8)
    for (i, j, k) in [(0, 0, 0), (4, 1, 2), (2, 3, 1)]:
        assert raw[k * 5 * 4 + j * 5 + i] == (255 if ph[i, j, k] else 0)

    den = 1.0 + rng.random(lat)
    raw_io.write_den(base, den, ph, iden=0.5, suffix=".den.in.raw")
    got = raw_io.read_den(base, lat)
    np.testing.assert_array_equal(got[ph == 1], den[ph == 1])
    assert (got[ph == 0] == 0.5).all()

    vel = rng.random(lat + (3,))
    raw_io.write_vel(base, vel, ph, ivel=(0.1, 0.2, 0.3), suffix=".vel.in.raw")
    got = raw_io.read_vel(base, lat)
    np.testing.assert_array_equal(got[ph == 1], vel[ph == 1])
    np.testing.assert_array_equal(got[ph == 0], np.broadcast_to([0.1, 0.2, 0.3], got[ph == 0].shape))
    flat = np.fromfile(base + ".vel.in.raw", dtype="<f8")
    for (i, j, k, c) in [(4, 3, 2, 2), (1, 0, 1, 0), (3, 2, 0, 1)]:
        assert flat[k * 3 * 5 * 4 + j * 3 * 5 + i * 3 + c] == got[i, j, k, c]

    frc = rng.random(lat + (3,))
    raw_io.write_frc(base, frc, ph, suffix=".force.in.raw")
    got = raw_io.read_frc(base, lat)
    np.testing.assert_array_equal(got[ph == 1], frc[ph == 1])
    assert (got[ph == 0] == 0).all()


@pytest.mark.gpu
def test_tube_example_on_gpu_matches_oracle():
    import jyulb_oracle as jo
    from jyulb.isothermal import Solver
    from jyulb.defs import TRT, WALL, PERIODIC
    lat = np.array([23, 23, 101], dtype=np.int32)
    ph = raw_io.read_geom(os.path.join(GOLD, "tube_x23_y23_z101"), lat)
    bc = np.array([WALL, WALL, PERIODIC, PERIODIC, PERIODIC, PERIODIC], dtype=np.int32)
    g = np.array([0.0, 0.0, 1e-5])
    out = []
    for cls in (jo.OracleSolver, Solver):
        s = cls(lat, TRT, 0.15, g, bc)
        s.get_phase()[...] = ph
        s.finalize_init()
        s.evolve(300)
        out.append((np.array(s.get_den()), np.array(s.get_vel())))
    fl = ph != SOLID
    for a, b in zip(out[0], out[1]):
        scale = np.max(np.abs(a[fl]))
        assert np.max(np.abs(a[fl] - b[fl])) <= 1e-12 * scale
    uz = out[1][1][fl][:, 2]
    assert uz.mean() > 0
