"""The D3Q19 discrete velocity set and its H10 Hermite moment space.

Mirrors the reference module (jyulb/D3Q19.pxd, jyulb/D3Q19.pyx:66-275): the
lattice-vector and moment enumerations and the constants that the reference
exposes as readonly attributes of `D3Q19_H10` (reachable as `solver.dvs`).
The per-site update procedures of that module (forw_euler_BGK/TRT/regul,
D3Q19.pyx:655-940) run on the GPU inside the fused step kernel
(csrc/lb_physics.h); they are not callable from Python, as in the reference
where they are cdef functions.
"""
import math

from .defs import FACE_1X, FACE_X1, FACE_1Y, FACE_Y1, FACE_1Z, FACE_Z1

# Lattice index vectors (D3Q19.pxd:59-79): L_1I_1J_K0 = (-1,-1,0), ...
L_1I_1J_K0 = 0
L_1I_J0_1K = 1
L_1I_J0_K0 = 2
L_1I_J0_K1 = 3
L_1I_J1_K0 = 4
L_I0_1J_1K = 5
L_I0_1J_K0 = 6
L_I0_1J_K1 = 7
L_I0_J0_1K = 8
L_I0_J0_K0 = 9
L_I0_J0_K1 = 10
L_I0_J1_1K = 11
L_I0_J1_K0 = 12
L_I0_J1_K1 = 13
L_I1_1J_K0 = 14
L_I1_J0_1K = 15
L_I1_J0_K0 = 16
L_I1_J0_K1 = 17
L_I1_J1_K0 = 18
LVECS = 19

# Hermite moments up to rank 2 (D3Q19.pxd:86-97)
H0 = 0
H1_X = 1
H1_Y = 2
H1_Z = 3
H2_XX = 4
H2_XY = 5
H2_XZ = 6
H2_YY = 7
H2_YZ = 8
H2_ZZ = 9
HMOMS = 10

_LI = (-1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1)
_LJ = (-1, 0, 0, 0, 1, -1, -1, -1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 0, 1)
_LK = (0, -1, 0, 1, 0, -1, 0, 1, -1, 0, 1, -1, 0, 1, 0, -1, 0, 1, 0)


class D3Q19_H10:
    """Constants of the D3Q19 velocity set with the H10 moment representation
    (D3Q19.pyx:66-275). All moments are sums over the lattice vectors."""

    def __init__(self):
        AS = math.sqrt(3.0)
        self.AS, self.AS2, self.AS4 = AS, 3.0, 9.0
        self.INV_AS, self.INV_AS2, self.INV_AS4 = 1.0 / AS, 1.0 / 3.0, 1.0 / 9.0
        self.W0, self.W1, self.W2 = 12.0 / 36.0, 2.0 / 36.0, 1.0 / 36.0
        self.B1 = [1.0, 0.0, 0.0]
        self.B2 = [0.0, 1.0, 0.0]
        self.B3 = [0.0, 0.0, 1.0]
        self.LI, self.LJ, self.LK = list(_LI), list(_LJ), list(_LK)
        self.LX = [float(v) for v in _LI]
        self.LY = [float(v) for v in _LJ]
        self.LZ = [float(v) for v in _LK]
        self.OPP_LVEC = [LVECS - 1 - l for l in range(LVECS)]
        nz = [abs(a) + abs(b) + abs(c) for a, b, c in zip(_LI, _LJ, _LK)]
        self.LW = [self.W0 if n == 0 else self.W1 if n == 1 else self.W2 for n in nz]
        self.FDIR_LVECS = 5
        self.FDIR_LVEC = [[0] * 5 for _ in range(6)]
        self.FDIR_LVEC[FACE_1X] = [L_1I_1J_K0, L_1I_J0_1K, L_1I_J0_K0, L_1I_J0_K1, L_1I_J1_K0]
        self.FDIR_LVEC[FACE_X1] = [L_I1_1J_K0, L_I1_J0_1K, L_I1_J0_K0, L_I1_J0_K1, L_I1_J1_K0]
        self.FDIR_LVEC[FACE_1Y] = [L_1I_1J_K0, L_I0_1J_1K, L_I0_1J_K0, L_I0_1J_K1, L_I1_1J_K0]
        self.FDIR_LVEC[FACE_Y1] = [L_1I_J1_K0, L_I0_J1_1K, L_I0_J1_K0, L_I0_J1_K1, L_I1_J1_K0]
        self.FDIR_LVEC[FACE_1Z] = [L_1I_J0_1K, L_I0_1J_1K, L_I0_J0_1K, L_I0_J1_1K, L_I1_J0_1K]
        self.FDIR_LVEC[FACE_Z1] = [L_1I_J0_K1, L_I0_1J_K1, L_I0_J0_K1, L_I0_J1_K1, L_I1_J0_K1]
        self.LRANGE = 1
        self.LSPAN = 3
        X, Y, Z = self.LX, self.LY, self.LZ
        i3 = self.INV_AS2
        self.HPOL2_XX = [X[l] * X[l] - i3 for l in range(LVECS)]
        self.HPOL2_XY = [X[l] * Y[l] for l in range(LVECS)]
        self.HPOL2_XZ = [X[l] * Z[l] for l in range(LVECS)]
        self.HPOL2_YY = [Y[l] * Y[l] - i3 for l in range(LVECS)]
        self.HPOL2_YZ = [Y[l] * Z[l] for l in range(LVECS)]
        self.HPOL2_ZZ = [Z[l] * Z[l] - i3 for l in range(LVECS)]
        w = self.LW
        self.K1_X = [self.AS2 * w[l] * X[l] for l in range(LVECS)]
        self.K1_Y = [self.AS2 * w[l] * Y[l] for l in range(LVECS)]
        self.K1_Z = [self.AS2 * w[l] * Z[l] for l in range(LVECS)]
        k2 = 0.5 * self.AS4
        self.K2_XX = [k2 * w[l] * self.HPOL2_XX[l] for l in range(LVECS)]
        self.K2_XY = [k2 * w[l] * self.HPOL2_XY[l] for l in range(LVECS)]
        self.K2_XZ = [k2 * w[l] * self.HPOL2_XZ[l] for l in range(LVECS)]
        self.K2_YY = [k2 * w[l] * self.HPOL2_YY[l] for l in range(LVECS)]
        self.K2_YZ = [k2 * w[l] * self.HPOL2_YZ[l] for l in range(LVECS)]
        self.K2_ZZ = [k2 * w[l] * self.HPOL2_ZZ[l] for l in range(LVECS)]
