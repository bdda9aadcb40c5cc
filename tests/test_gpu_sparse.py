"""Sparse layout (fluid sites only) against the dense layouts and the oracle.

The sparse layout stores the populations of the fluid sites only, compacted in
ascending dense order, and runs the odd AA step through a table of the 18
locations per site that the dense resolver names (bounce-back, periodic,
z-wrapped and mirrored links): the fluid-only arrays of the reference's C++
kernel (include/kernel/kernel.cpp:227-271), here with 32-bit entries computed
on the GPU at finalize_init. The per-site arithmetic is the dense kernel's, so
every precision reproduces the dense ROWS layout bit for bit. FP64 against
all reference fixtures runs in test_gpu_parity.py.
"""
import json
import os

import numpy as np
import pytest

import scenarios

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
# every boundary kind: periodic, walls + symmetric, symmetric, Zou-He on x / y
# / negative faces, moving walls (Couette), a long TRT run
NAMES = ["bgk_periodic", "trt_symwall", "reg_symmetric", "reg_zouhe_x", "trt_zouhe_y", "bgk_zouhe_neg",
         "ex_couette", "trt_long"]


def _run_fixture(name, **kw):
    from jyulb.isothermal import Solver
    z = np.load(os.path.join(GOLD, name + ".npz"))
    sc = json.loads(str(z["scenario"]))

    def mk(*a):
        return Solver(*a, **kw)

    return scenarios.run_scenario(mk, sc, record_f=True)


@pytest.mark.parametrize("precision", ["fp32", "fp16", "fp32_storage"])
@pytest.mark.parametrize("name", NAMES)
def test_sparse_equals_rows_bitwise(name, precision):
    print(f"scenario {name} precision {precision}", flush=True)
    a = _run_fixture(name, precision=precision, layout="rows")
    b = _run_fixture(name, precision=precision, layout="sparse")
    assert a.keys() == b.keys()
    for k in a:
        np.testing.assert_array_equal(b[k], a[k], err_msg=f"{name} {precision} {k}")


def _porous_run(Solver, lat, steps, **kw):
    from jyulb.defs import TRT, PERIODIC, SYMMETRIC, SOLID, FLUID
    from jyulb.geometry import porous_spheres
    lat = np.array(lat, dtype=np.int32)
    bc = np.array([PERIODIC, PERIODIC, SYMMETRIC, SYMMETRIC, PERIODIC, PERIODIC], dtype=np.int32)
    s = Solver(lat, TRT, 1.0 / 30.0, np.array([1e-5, 0.0, 2e-6]), bc, **kw)
    solid = porous_spheres(lat, 3, 6, 0.25, seed=11)
    ph = np.where(solid, SOLID, FLUID).astype(np.uint8)
    s.get_phase()[...] = ph
    rng = np.random.default_rng(4)
    acc = np.asarray(s.get_acc())
    acc[...] = 1e-6 * rng.standard_normal(acc.shape) * (ph != SOLID)[..., None]
    nf = s.finalize_init()
    out = [nf]
    for st in steps:
        s.evolve(st)
        out += [np.array(s.get_den(), copy=True), np.array(s.get_vel(), copy=True)]
    return out, ph


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16"])
def test_sparse_porous_medium_equals_rows(precision):
    """Porous medium, 25% solid, periodic x/z and symmetric y faces, gravity
    plus a per-site force, nz not a multiple of 64; odd and even step counts
    over three evolve calls."""
    from jyulb.isothermal import Solver
    lat, steps = (72, 40, 90), (7, 40, 53)
    print(f"porous {lat} {precision}", flush=True)
    a, ph = _porous_run(Solver, lat, steps, precision=precision, layout="rows")
    b, _ = _porous_run(Solver, lat, steps, precision=precision, layout="sparse")
    assert a[0] == b[0] == int((ph != 0).sum())
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_array_equal(y, x)


@pytest.mark.parametrize("xcd_order", ["0", "18446744073709551615"])
@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16", "fp32_storage"])
def test_sparse_large_lattice_both_tile_orders(precision, xcd_order, monkeypatch):
    """A lattice of 3000+ tiles of 256 compact sites whose last wave holds 3
    sites (fewer than the 18 links), in the XCD tile order and in the plain
    one, against the rows layout, bit for bit."""
    from jyulb.isothermal import Solver
    monkeypatch.setenv("LB_XCD_REMAP_CELLS", xcd_order)
    lat, steps = (144, 80, 96), (5, 8)
    print(f"large sparse {lat} {precision} xcd_cells={xcd_order}", flush=True)
    a, ph = _porous_run(Solver, lat, steps, precision=precision, layout="rows")
    b, _ = _porous_run(Solver, lat, steps, precision=precision, layout="sparse")
    assert a[0] == b[0] == int((ph != 0).sum()) > 3000 * 256 and a[0] % 64 == 3
    for x, y in zip(a[1:], b[1:]):
        np.testing.assert_array_equal(y, x)


def test_sparse_code_kinds():
    """The odd-step table of a porous medium with periodic x/z and symmetric
    y faces holds every kind of code: near offsets for most links, the wave's
    alternate location (rows that wrap around the periodic z face), bounce-
    back at the spheres, and full 32-bit locations (mirrored links at the
    symmetric faces); the kinds add up to 18 links per fluid site."""
    from jyulb.isothermal import Solver
    from jyulb.defs import TRT, PERIODIC, SYMMETRIC
    from jyulb.geometry import porous_spheres
    lat = np.array([48, 40, 90], dtype=np.int32)
    bc = np.array([PERIODIC, PERIODIC, SYMMETRIC, SYMMETRIC, PERIODIC, PERIODIC], dtype=np.int32)
    s = Solver(lat, TRT, 1.0 / 30.0, np.zeros(3), bc, layout="sparse")
    s.get_phase()[...] = np.where(porous_spheres(lat, 3, 6, 0.25, seed=11), 0, 1)
    nf = s.finalize_init()
    k = s.engine.sparse_codes()
    print("code kinds", k, "fluid sites", nf, flush=True)
    assert sum(k.values()) == 18 * nf
    assert min(k.values()) > 0
    assert k["near"] > 0.9 * 18 * nf and k["far"] < 0.02 * 18 * nf


@pytest.mark.parametrize("precision", ["fp64", "fp32", "fp16", "fp32_storage"])
def test_sparse_pull_steps_equal_in_place_steps(precision, monkeypatch):
    """The sparse layout's pull steps (a second population array, aligned
    stores) against its odd AA steps in place, bit for bit, over evolve calls
    of odd and even lengths."""
    from jyulb.isothermal import Solver
    lat, steps = (72, 40, 90), (7, 40, 53)
    out = []
    for pull in ("1", "0"):
        monkeypatch.setenv("LB_SP_PULL", pull)
        print(f"pull {pull} {precision}", flush=True)
        a, ph = _porous_run(lambda *x, **k: Solver(*x, layout="sparse", **k), lat, steps, precision=precision)
        out.append(a)
    for x, y in zip(out[0], out[1]):
        np.testing.assert_array_equal(x, y)


@pytest.mark.parametrize("name", ["trt_symwall", "reg_zouhe_x", "bgk_zouhe_neg", "ex_couette", "trt_long"])
def test_sparse_pull_steps_on_reference_scenarios(name, monkeypatch):
    """Pull steps forced on small lattices (walls, symmetric faces, Zou-He
    faces, moving walls) against the rows layout, FP64, bit for bit."""
    monkeypatch.setenv("LB_SP_PULL", "1")
    print(f"pull scenario {name}", flush=True)
    a = _run_fixture(name, layout="rows")
    b = _run_fixture(name, layout="sparse")
    for k in a:
        np.testing.assert_array_equal(b[k], a[k], err_msg=f"{name} {k}")


def test_sparse_porous_medium_matches_oracle():
    import jyulb_oracle as jo
    from jyulb.isothermal import Solver

    def sp(*a):
        return Solver(*a, layout="sparse")

    lat, steps = (40, 24, 30), (5, 12)
    ref, ph = _porous_run(jo.OracleSolver, lat, steps)
    got, _ = _porous_run(sp, lat, steps)
    fl = ph != 0
    assert ref[0] == got[0]
    for x, y in zip(ref[1:], got[1:]):
        np.testing.assert_array_equal(y[fl], x[fl])  # interior views


def test_sparse_fewer_fluid_sites_than_a_wave():
    """27 fluid sites: the compact arrays are padded to one wave of 64."""
    import jyulb_oracle as jo
    from jyulb.isothermal import Solver
    from jyulb.defs import TRT, PERIODIC
    lat = np.array([3, 3, 3], dtype=np.int32)
    bc = np.array([PERIODIC] * 6, dtype=np.int32)
    res = []
    for mk in (jo.OracleSolver, lambda *a: Solver(*a, layout="sparse")):
        s = mk(lat, TRT, 0.1, np.array([1e-5, -2e-5, 3e-5]), bc)
        s.get_phase()[...] = 1
        np.asarray(s.get_vel())[..., 0] = 0.001 * np.arange(27).reshape(3, 3, 3)
        assert s.finalize_init() == 27
        s.evolve(9)
        res.append((np.array(s.get_den()), np.array(s.get_vel())))
    for x, y in zip(res[0], res[1]):
        np.testing.assert_array_equal(y, x)


def test_sparse_all_solid_lattice():
    """No fluid site: evolve does nothing and, like the reference
    (isothermal.pyx:1405-1406), returns the unchanged step count."""
    from jyulb.isothermal import Solver
    from jyulb.defs import BGK
    s = Solver(np.array([8, 4, 5], dtype=np.int32), BGK, 0.1, layout="sparse")
    assert s.finalize_init() == 0
    assert s.evolve(4) == 0
    assert np.all(np.asarray(s.get_den()) == 1.0)


def test_sparse_device_memory_follows_fluid_sites():
    """Population bytes scale with the fluid sites, not the lattice."""
    from jyulb.isothermal import Solver
    lat = (64, 32, 64)
    a, ph = _porous_run(Solver, lat, (2,))
    nf = a[0]
    rows = Solver(np.array(lat, dtype=np.int32), 0, 0.1, layout="rows")
    rows.get_phase()[...] = ph
    rows.finalize_init()
    sp = Solver(np.array(lat, dtype=np.int32), 0, 0.1, layout="sparse")
    sp.get_phase()[...] = ph
    sp.finalize_init()
    b_rows = rows.engine.info()["device_bytes"]
    b_sp = sp.engine.info()["device_bytes"]
    nfp = (nf + 63) // 64 * 64
    ncell = 64 * 32 * 64
    # rows: 19 f64 per cell; sparse: per padded fluid site 19 f64 (one array:
    # pull steps only from 1 GiB of populations), its site index, the 18 u32
    # locations and their 4-bit codes (two u32 and a byte), per wave of 64 the
    # 18 u32 bases and 18 alternates, and a compact index per cell
    sparse = (19 * 8 + 4 + 18 * 4 + 9) * nfp + 36 * 4 * (nfp // 64) + 4 * ncell
    assert b_rows - b_sp == 19 * 8 * ncell - sparse


@pytest.mark.parametrize("layout", ["rows", "sparse"])
def test_lattice_beyond_32_bit_site_index_is_refused(layout):
    """Sites are indexed with 32-bit integers and counted with ints; a lattice
    of 2^31 or more padded cells (here 2^32) is refused by lb_create before
    it allocates anything (the engine handle is created directly: the Solver
    would first build its host arrays)."""
    from jyulb import _native
    from jyulb.defs import BGK, PERIODIC
    with pytest.raises(ValueError, match="limit of 2\\^31 padded cells"):
        _native.Engine((4096, 1024, 1024), BGK, 0.1, np.zeros(3), [PERIODIC] * 6, precision="fp16",
                       layout=layout)
