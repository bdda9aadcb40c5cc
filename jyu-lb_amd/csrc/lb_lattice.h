// D3Q19 lattice tables and lattice-wide constants shared by host and device.
//
// Lattice vector numbering follows the reference (jyulb/D3Q19.pxd:59-79):
// l -> (ci,cj,ck) with ci the slowest-varying component, opp(l) = 18 - l,
// rest vector l = 9. Weights 1/3, 1/18, 1/36 (D3Q19.pyx:192-211).
#pragma once
#include <stdint.h>

#ifndef __HIPCC__
#define __host__
#define __device__
#endif

namespace lb {

constexpr int Q = 19;
constexpr int REST = 9;

constexpr int CI[Q] = {-1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1};
constexpr int CJ[Q] = {-1, 0, 0, 0, 1, -1, -1, -1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 0, 1};
constexpr int CK[Q] = {0, -1, 0, 1, 0, -1, 0, 1, -1, 0, 1, -1, 0, 1, 0, -1, 0, 1, 0};

__host__ __device__ constexpr int opp(int l) { return 18 - l; }
__host__ __device__ constexpr int ccomp(int l, int axis) {
    return axis == 0 ? CI[l] : axis == 1 ? CJ[l] : CK[l];
}
// number of nonzero components: 0 rest, 1 axis, 2 diagonal
__host__ __device__ constexpr int nnz(int l) {
    return (CI[l] != 0) + (CJ[l] != 0) + (CK[l] != 0);
}
// the lattice vector with component `axis` negated (specular mirror)
__host__ __device__ constexpr int mirror(int l, int axis) {
    // index = (ci+1)*? is not linear for D3Q19, so search the table
    int ci = CI[l], cj = CJ[l], ck = CK[l];
    if (axis == 0) ci = -ci;
    if (axis == 1) cj = -cj;
    if (axis == 2) ck = -ck;
    int r = -1;
    for (int m = 0; m < Q; ++m)
        if (CI[m] == ci && CJ[m] == cj && CK[m] == ck) r = m;
    return r;
}

// Boundary-condition and enumeration ids, identical to jyulb/defs.pxd:17-103.
enum Face { F1X = 0, FX1 = 1, F1Y = 2, FY1 = 3, F1Z = 4, FZ1 = 5 };
enum Bc { WALL = 0, PERIODIC = 1, FIXED_DEN = 2, FIXED_VEL = 3, SYMMETRIC = 4 };
enum Rlx { BGK = 0, TRT = 1, REGUL = 2 };
enum Phase { SOLID = 0, FLUID = 1 };

// Per-cell link mask (u32), built once per geometry by the setup kernel:
//   bit l (l != 9)  the pulled neighbour x - c_l resolves to a SOLID site
//                   (halfway bounce-back link)
//   FLUID           the cell is an active fluid site
//   BND             at least one fluid link crosses a periodic or symmetric
//                   face, so the odd AA step resolves its addresses the slow way
//   MOVING          at least one solid link ends on a solid site whose
//                   velocity is nonzero (Ladd moving-wall term is needed)
//   ZH(f)           the cell lies on a FIXED_DEN/FIXED_VEL face f (Zou-He)
constexpr uint32_t M_FLUID = 1u << 19;
constexpr uint32_t M_BND = 1u << 20;
constexpr uint32_t M_MOVING = 1u << 21;
constexpr uint32_t M_ZH0 = 1u << 22;  // faces 0..5 -> bits 22..27
constexpr uint32_t M_ZH_ANY = 0x3Fu << 22;
constexpr uint32_t M_LINKS = (1u << 19) - 1;
// the site lies on a periodic z face (k = 0 or nz - 1) and crosses no other
// periodic/symmetric face: its z-crossing links wrap within the lattice row,
// which the bulk kernel does with a uniform row-start address
constexpr uint32_t M_KWRAP = 1u << 28;
// sites handled by the special-site kernel on odd steps; even steps touch only
// the site's own slots, and send only the Zou-He sites there (see lb_engine.hip)
constexpr uint32_t M_SPECIAL = M_BND | M_ZH_ANY;

}  // namespace lb
