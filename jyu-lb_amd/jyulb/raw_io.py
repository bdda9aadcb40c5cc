"""Raw-format lattice I/O of the reference's C++ front-end, for the jyulb arrays.

Same files, suffixes and byte layouts as IsothermalRawIO
(include/io/raw_io.cpp:23-275):

  <base>.geom.in.raw / .geom.out.raw    u8 per node, 0 = solid, nonzero = fluid
                                         (the writer uses 255 for fluid)
  <base>.den.in.raw  / .den.out.raw     f64 per node
  <base>.vel.in.raw  / .vel.out.raw     3 x f64 per node, components interleaved
  <base>.force.in.raw / .force.out.raw  3 x f64 per node, components interleaved

Nodes are ordered with x fastest: n = k*nx*ny + j*nx + i (raw_ijk2n,
raw_io.cpp:252-255); vector components follow each node (raw_ijkl2n,
raw_io.cpp:257-260). The arrays exchanged with the Solver are its interior
views in C order, i.e. indexed [i, j, k] (and [i, j, k, c]); the functions
here transpose between the two orders. As in the reference, the writers fill
non-fluid nodes of den/vel with the given initial values and the readers
return full arrays (the caller applies them to fluid sites).
"""
import os

import numpy as np

from .defs import SOLID, FLUID


def _path(base, ext):
    return os.fspath(base) + ext


def _check(a, lat, ncomp, what):
    nx, ny, nz = (int(v) for v in lat)
    n = nx * ny * nz * ncomp
    if a.size != n:
        raise ValueError(f"{what}: file holds {a.size} values, lattice {nx}x{ny}x{nz} needs {n}")


def read_geom(base, lat, suffix=".geom.in.raw"):
    """Phase field [i, j, k] (SOLID=0 / FLUID=1) from <base>.geom.in.raw
    (raw_io.cpp:23-44)."""
    nx, ny, nz = (int(v) for v in lat)
    raw = np.fromfile(_path(base, suffix), dtype=np.uint8)
    _check(raw, lat, 1, "read_geom")
    xyz = raw.reshape(nz, ny, nx).transpose(2, 1, 0)
    return np.where(xyz != 0, FLUID, SOLID).astype(np.uint8)


def write_geom(base, phase, suffix=".geom.out.raw"):
    """Phase field [i, j, k] to <base>.geom.out.raw as 0 (solid) / 255
    (fluid) (raw_io.cpp:46-68)."""
    ph = np.asarray(phase)
    raw = np.where(ph.transpose(2, 1, 0) != SOLID, 255, 0).astype(np.uint8)
    raw.tofile(_path(base, suffix))


def read_den(base, lat, suffix=".den.in.raw"):
    """Density [i, j, k] from <base>.den.in.raw (raw_io.cpp:70-94)."""
    nx, ny, nz = (int(v) for v in lat)
    raw = np.fromfile(_path(base, suffix), dtype="<f8")
    _check(raw, lat, 1, "read_den")
    return np.ascontiguousarray(raw.reshape(nz, ny, nx).transpose(2, 1, 0))


def write_den(base, den, phase, iden=1.0, suffix=".den.out.raw"):
    """Density [i, j, k] of the fluid nodes to <base>.den.out.raw; other
    nodes get `iden` (raw_io.cpp:96-123)."""
    d = np.where(np.asarray(phase) != SOLID, np.asarray(den, dtype=np.float64), float(iden))
    np.ascontiguousarray(d.transpose(2, 1, 0)).astype("<f8").tofile(_path(base, suffix))


def _read_vec(base, lat, suffix, what):
    nx, ny, nz = (int(v) for v in lat)
    raw = np.fromfile(_path(base, suffix), dtype="<f8")
    _check(raw, lat, 3, what)
    return np.ascontiguousarray(raw.reshape(nz, ny, nx, 3).transpose(2, 1, 0, 3))


def _write_vec(base, vec, phase, ivec, suffix):
    v = np.array(vec, dtype=np.float64, copy=True)
    v[np.asarray(phase) == SOLID] = np.asarray(ivec, dtype=np.float64)
    np.ascontiguousarray(v.transpose(2, 1, 0, 3)).astype("<f8").tofile(_path(base, suffix))


def read_vel(base, lat, suffix=".vel.in.raw"):
    """Velocity [i, j, k, 3] from <base>.vel.in.raw (raw_io.cpp:125-149)."""
    return _read_vec(base, lat, suffix, "read_vel")


def write_vel(base, vel, phase, ivel=(0.0, 0.0, 0.0), suffix=".vel.out.raw"):
    """Velocity [i, j, k, 3] of the fluid nodes to <base>.vel.out.raw; other
    nodes get `ivel` (raw_io.cpp:151-184)."""
    _write_vec(base, vel, phase, ivel, suffix)


def read_frc(base, lat, suffix=".force.in.raw"):
    """Body force [i, j, k, 3] from <base>.force.in.raw (raw_io.cpp:186-210)."""
    return _read_vec(base, lat, suffix, "read_frc")


def write_frc(base, frc, phase, ifrc=(0.0, 0.0, 0.0), suffix=".force.out.raw"):
    """Body force [i, j, k, 3] of the fluid nodes to <base>.force.out.raw;
    other nodes get `ifrc` (raw_io.cpp:212-245)."""
    _write_vec(base, frc, phase, ifrc, suffix)
