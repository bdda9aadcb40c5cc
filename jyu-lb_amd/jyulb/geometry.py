"""Synthetic porous geometries (no counterpart in the reference package).

The reference reads porous media from raw files (source/isothermal/raw_io.cpp:
23-44); for synthetic workloads of a given solid fraction this module places
seeded random spheres instead.
"""
import numpy as np


def porous_spheres(lat, rmin, rmax, solid_fraction, seed, open_faces=(), open_depth=1):
    """Boolean SOLID mask of shape `lat`: spheres with radius uniform in
    [rmin, rmax] and centres uniform in the box (numpy default_rng(seed)) are
    added until the solid fraction reaches `solid_fraction`. Spheres do not
    wrap around the box. `open_faces` lists faces (0..5 = 1X, X1, 1Y, Y1, 1Z,
    Z1) whose `open_depth` boundary layers are kept free of solid, e.g. the
    buffers in front of Zou-He faces (a velocity face directly against solid
    pushes fluid into a dead end and the density diverges)."""
    lat = tuple(int(n) for n in lat)
    rng = np.random.default_rng(seed)
    solid = np.zeros(lat, dtype=bool)
    axes = [np.arange(n, dtype=np.float64) for n in lat]
    keep = np.ones(lat, dtype=bool)
    for f in open_faces:
        sl = [slice(None)] * 3
        n = lat[f // 2]
        sl[f // 2] = slice(0, open_depth) if f % 2 == 0 else slice(max(n - open_depth, 0), n)
        keep[tuple(sl)] = False
    target = solid_fraction * solid.size
    if target > int(keep.sum()):
        raise ValueError(f"solid fraction {solid_fraction} needs {int(np.ceil(target))} solid cells; only "
                         f"{int(keep.sum())} lie outside the open boundary layers")
    count = 0
    while count < target:
        c = rng.random(3) * np.array(lat)
        r = rng.uniform(rmin, rmax)
        lo = [max(int(np.floor(c[a] - r)), 0) for a in range(3)]
        hi = [min(int(np.ceil(c[a] + r)) + 1, lat[a]) for a in range(3)]
        d2 = ((axes[0][lo[0]:hi[0], None, None] - c[0]) ** 2
              + (axes[1][None, lo[1]:hi[1], None] - c[1]) ** 2
              + (axes[2][None, None, lo[2]:hi[2]] - c[2]) ** 2)
        box = (slice(lo[0], hi[0]), slice(lo[1], hi[1]), slice(lo[2], hi[2]))
        add = (d2 <= r * r) & keep[box] & ~solid[box]
        solid[box] |= add
        count += int(add.sum())
    return solid
