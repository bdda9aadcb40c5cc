// Definitions shared by the host code and the kernel translation units of
// the D3Q19 engine: storage variants, lattice geometry, the parameter blocks
// of the step and adjoint kernels, address resolution and the launchers
// that each per-precision translation unit (lb_inst.hip) instantiates.
#pragma once
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <math.h>
#include <stdint.h>
#include <utility>

#include "../../include/jyulb_gpu.h"
#include "lb_lattice.h"
#include "lb_physics.h"

namespace lb {

// ---------------------------------------------------------------------------
// storage variants
// ---------------------------------------------------------------------------
template <int P> struct Var;
template <> struct Var<LB_FP64> { using TS = double; using TC = double; static constexpr bool SH = false; };
template <> struct Var<LB_FP32_SHIFTED> { using TS = float; using TC = float; static constexpr bool SH = true; };
template <> struct Var<LB_FP16_STORAGE> { using TS = _Float16; using TC = float; static constexpr bool SH = true; };
template <> struct Var<LB_FP32_STORAGE> { using TS = float; using TC = double; static constexpr bool SH = true; };

template <class TC> LBD TC ldp(const double* p) { return (TC)(*p); }

// Cache policy of the population loads and stores: every population is
// touched once per step, so streaming (non-temporal) accesses keep L2 for the
// lines that neighbouring waves share. Separate switches per step parity and
// direction (bit 0 even loads, 1 even stores, 2 odd loads, 3 odd stores;
// bit 4: in odd steps stream the populations without a k component and
// apply bits 2/3 only to the k-shifted ones).
#ifndef LB_NT_MASK
#define LB_NT_MASK 3
#endif
template <bool NT, class T> LBD T ld_stream(const T* p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT, class T> LBD void st_stream(T* p, T v) {
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

template <int P, bool NT>
LBD typename Var<P>::TC ldf(const typename Var<P>::TS* p, float inv_scale) {
    if constexpr (P == LB_FP16_STORAGE) {
        return (float)ld_stream<NT>(p) * inv_scale;
    } else {
        (void)inv_scale;
        return ld_stream<NT>(p);
    }
}
template <int P, bool NT>
LBD void stf(typename Var<P>::TS* p, typename Var<P>::TC v, float scale) {
    if constexpr (P == LB_FP16_STORAGE)
        st_stream<NT>(p, (_Float16)(v * scale));
    else {
        (void)scale;
        st_stream<NT>(p, (typename Var<P>::TS)v);
    }
}

// ---------------------------------------------------------------------------
// geometry helpers
// ---------------------------------------------------------------------------
struct Dims {
    int nx, ny, nz, nzp;
    int bc[6];
};

// lattice vector l with component a negated (specular mirror), as a table
struct MirrorTable {
    int8_t m[3][Q];
};
__host__ __device__ constexpr MirrorTable make_mirror_table() {
    MirrorTable t{};
    for (int a = 0; a < 3; ++a)
        for (int l = 0; l < Q; ++l) t.m[a][l] = (int8_t)mirror(l, a);
    return t;
}
__device__ constexpr MirrorTable c_mirror_tab = make_mirror_table();

// Resolve a possibly out-of-range site (i,j,k) reached along lattice vector l
// to the interior site whose data it carries, undoing the reference's halo
// passes in reverse order (z, then y, then x: the x pass covers interior j,k
// only, the y pass all i and interior k, the z pass all i and j,
// isothermal.pyx:509-837). Periodic faces wrap, symmetric faces mirror the
// site onto the adjacent layer and reflect l and the velocity component
// (flip bit per axis). Returns false for a WALL/FIXED halo: a SOLID site at
// rest.
LBD bool resolve(const Dims& d, int& i, int& j, int& k, int& l, int& flip) {
    int c[3] = {i, j, k};
    const int n[3] = {d.nx, d.ny, d.nz};
#pragma unroll
    for (int a = 2; a >= 0; --a) {
        if (c[a] < 0 || c[a] >= n[a]) {
            const bool lo = c[a] < 0;
            const int t = d.bc[2 * a + (lo ? 0 : 1)];
            if (t == PERIODIC) {
                c[a] += lo ? n[a] : -n[a];
            } else if (t == SYMMETRIC) {
                c[a] = lo ? 0 : n[a] - 1;
                l = c_mirror_tab.m[a][l];
                flip ^= 1 << a;
            } else {
                return false;
            }
        }
    }
    i = c[0];
    j = c[1];
    k = c[2];
    return true;
}

// A pointer known to be equal across the wave. Marking it keeps the compiler
// from folding per-lane offsets into it early: the address of a load then
// stays "scalar base + 32-bit lane offset" instead of one 64-bit vector
// address per population held in registers until the store.
template <class T>
LBD T* uniform_ptr(T* p) {
    const uint64_t v = (uint64_t)p;
    const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
    const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
    return (T*)(((uint64_t)hi << 32) | lo);
}

LBD void decompose(uint32_t c, int ny, int nzp, int& i, int& j, int& k) {
    uint32_t row = c / (uint32_t)nzp;
    k = (int)(c - row * (uint32_t)nzp);
    i = (int)(row / (uint32_t)ny);
    j = (int)(row - (uint32_t)i * (uint32_t)ny);
}

// ---------------------------------------------------------------------------
// step kernel
// ---------------------------------------------------------------------------
// Sparse-layout location codes of the odd step (see step_sparse): SP_BITS
// bits per link, SP_PER_WORD links per 32-bit word. Per site SP_WORDS words
// (link slots 0-15) and one byte (slots 16 and 17): SP_CODE_BYTES bytes, in
// SP_WORDS rows of 32-bit words and a row of bytes. A code below SP_ALT is the
// offset of the location from the wave's base for the link plus the site's
// lane, SP_ALT names the wave's alternate location for the link, SP_BOUNCE
// the site's own slot l, SP_FAR a location read from the 32-bit table. Per
// wave: SP_WAVE_WORDS words, the 18 bases, then the 18 alternates.
constexpr int SP_BITS = 4, SP_PER_WORD = 8, SP_WORDS = 2;
constexpr int SP_CODE_BYTES = 4 * SP_WORDS + 1;
constexpr int SP_WAVE_WORDS = 36;
constexpr uint32_t SP_FAR = 15u, SP_BOUNCE = 14u, SP_ALT = 13u;
static_assert(SP_PER_WORD * SP_BITS == 32 && SP_PER_WORD * SP_WORDS + 8 / SP_BITS == 18, "code packing");
// code of link slot sl from w[0..SP_WORDS-1] and the site's byte in w[SP_WORDS]
LBD uint32_t sp_code(const uint32_t* w, int sl) {
    return (w[sl / SP_PER_WORD] >> (SP_BITS * (sl % SP_PER_WORD))) & ((1u << SP_BITS) - 1u);
}

// lattices of at least this many cells deal the step tiles to the XCDs in
// contiguous ranges (tile_index)
#ifndef LB_XCD_REMAP_CELLS
#define LB_XCD_REMAP_CELLS (1u << 25)
#endif

constexpr uint32_t XCD_CONTIGUOUS = 0xffffffffu;

struct StepParams {
    void* f;
    void* f2;  // sparse layout: destination of a pull step (k_step_sp_pull), null otherwise
    const uint32_t* mask;
    double* den;
    double* vel;  // 3 * nq, SoA
    const void* acc;  // 3 * nq of the compute type, SoA, or null
    const double* fden[6];
    const double* fvel[6];
    double g[3];
    double prm[2];
    double rho0;
    float scale, inv_scale;
    Dims d;
    int NJ, NK;  // halo-padded sizes (face-data indexing)
    uint32_t nq;
    uint32_t ntiles;
    int store_dv;
    int moving;
    int64_t rs;  // element stride between rows (i,j) of the population array
    int64_t ps;  // element stride between populations
    const uint8_t* wclass;  // per wave of 64 sites: 0 = plain (see k_mask)
    uint32_t xcd;           // tile order (tile_index): 0 round-robin, XCD_CONTIGUOUS, or a chunk of tiles per XCD
    uint32_t lds[2];        // dynamic LDS per workgroup of the bulk step launch (even, odd): caps occupancy
    uint32_t tile;          // threads per workgroup of the launch that reads ntiles (see TILE)
    const uint32_t* mov_idx;  // per site: index of a moving site in ladd_u
    const double* ladd_u;     // per moving site: the 18 wall-velocity sums of Ladd's term
    uint32_t nwaves;
    // sites with M_SPECIAL: those on a Zou-He face first (nzh of them), then
    // the rest; ascending within each wave
    const uint32_t* special;
    uint32_t nspecial, nzh;
    void* rec;  // if set: (rho, u*) of this step, 4 rows of the compute type (adjoint recompute):
                // nq values by cell in the dense layouts, ps by compact site in the sparse layout
    // sparse layout (LB_LAYOUT_SPARSE), null otherwise: compact index of every
    // site (fluid sites only), site of every compact index, and the odd-step
    // location of link l of compact site n at tab[link_slot(l) * ps + n]
    const uint32_t* cidx;
    const uint32_t* site;
    const uint32_t* tab;
    const uint32_t* tabc;   // compact form of tab (see step_sparse): SP_WORDS rows of ps words
    const uint8_t* tabb;    // ... and its row of ps bytes (link slots 16 and 17), after them
    const uint32_t* tabw;   // its base and alternate locations, SP_WAVE_WORDS per wave of 64 compact sites
    uint32_t nf;  // fluid sites = compact sites the step processes
    // dissipation objective (k_diss): per site 19 population seeds, 3
    // acceleration seeds and the site's Pi^neq:Pi^neq, each an nq-long row of
    // the compute type; dscale = 1 / (fluid sites)
    void* dseed;
    double dscale;
};

// expanded index of padded site c in the population array: population l of
// site (i,j,k) lives at (i*ny + j)*rs + k + l*ps
LBD int64_t eidx(const StepParams& p, uint32_t c) {
    if (p.cidx) return p.cidx[c];  // sparse: population l of site c at cidx[c] + l*ps
    if (p.rs == p.d.nzp) return c;
    const uint32_t row = c / (uint32_t)p.d.nzp;
    return (int64_t)row * p.rs + (c - row * (uint32_t)p.d.nzp);
}

// Threads per workgroup: TILE for the adjoint and special-site kernels; the
// bulk step kernels run p.tile threads (a multiple of 64, at most
// STEP_TILE_MAX), chosen per precision at lb_create.
constexpr int TILE = 256;
constexpr int STEP_TILE_MAX = 512;
constexpr int RED_BLOCK_ADJ = 256;

// element offset of the location an odd step reads f_l from (and writes
// f*_opp(l) back to)
template <bool SPECIAL = true>
LBD int64_t odd_loc(const StepParams& p, uint32_t m, int64_t e, int l, int i, int j, int k) {
    if (l == REST) return e + (int64_t)REST * p.ps;
    if (m & (1u << l)) return e + (int64_t)l * p.ps;
    if (!SPECIAL || !(m & (M_BND | M_KWRAP)))
        return e - ((int64_t)CI[l] * p.d.ny + CJ[l]) * p.rs - CK[l] + (int64_t)opp(l) * p.ps;
    int yi = i - CI[l], yj = j - CJ[l], yk = k - CK[l], lp = l, fl = 0;
    resolve(p.d, yi, yj, yk, lp, fl);
    return ((int64_t)yi * p.d.ny + yj) * p.rs + yk + (int64_t)(18 - lp) * p.ps;
}

// Ladd's moving-wall term for bounce-back link l (isothermal.pyx:1112-1248):
// f_l = f_opp +/- cff*(u_a +/- u_b), u the velocity of the solid site
// x - c_l (mirrored through symmetric faces, zero for WALL/FIXED halos), a and
// b the non-zero components of c_l, cff = 2 rho_prev w_l 3. The velocity sum
// (u_a, or u_a +/- u_b for diagonal links) depends only on the walls, so it is
// tabulated per moving site (k_ladd_prep, 18 links) whenever the walls change;
// the step kernel adds cff * sum with the sign of c_l's first component.
template <int L>
LBD double ladd_sum(const Dims& d, const double* vel, int64_t nq, int i, int j, int k) {
    int yi = i - CI[L], yj = j - CJ[L], yk = k - CK[L], lp = L, fl = 0;
    double u[3] = {0.0, 0.0, 0.0};
    if (resolve(d, yi, yj, yk, lp, fl)) {
        const int64_t y = ((int64_t)yi * d.ny + yj) * d.nzp + yk;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            double v = vel[(int64_t)a * nq + y];
            u[a] = (fl >> a & 1) ? -v : v;
        }
    }
    constexpr int c[3] = {CI[L], CJ[L], CK[L]};
    constexpr int a = c[0] != 0 ? 0 : (c[1] != 0 ? 1 : 2);
    constexpr int b = nnz(L) == 1 ? -1 : (a == 0 ? (c[1] != 0 ? 1 : 2) : 2);
    if constexpr (b < 0) return u[a];
    else if constexpr (c[a] == c[b]) return u[a] + u[b];
    else return u[a] - u[b];
}

// table slot of link l (the rest population has none)
LBD constexpr int link_slot(int l) { return l < REST ? l : l - 1; }
// sign of the first non-zero component of c_l
LBD constexpr bool ladd_plus(int l) { return (CI[l] != 0 ? CI[l] : CJ[l] != 0 ? CJ[l] : CK[l]) > 0; }

template <class T>
LBD void ladd_apply(const StepParams& p, uint32_t c, uint32_t m, T* f) {
    const T den = (T)p.den[c];
    const T cff1 = T(2.0) * den * T(2.0 / 36.0) * T(3.0);
    const T cff2 = T(2.0) * den * T(1.0 / 36.0) * T(3.0);
    const double* U = p.ladd_u + (int64_t)p.mov_idx[c] * (Q - 1);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        if (l == REST || !(m & (1u << l))) continue;
        const T term = (nnz(l) == 1 ? cff1 : cff2) * (T)U[link_slot(l)];
        f[l] = ladd_plus(l) ? f[l] + term : f[l] - term;
    }
}

struct AdjParams {
    void* lam;          // cotangent array, population layout of f, compute type
    void* lam2;         // pull form: the array the step writes (null: in place)
    const void* rec;    // (rho, u*) recorded by the forward recompute of this step
    void* gacc;         // accumulated acceleration cotangent, compute type, 3 rows indexed like rec
    double* gface;      // accumulated face-velocity cotangent (face of gface_id), face layout x 3
    int gface_id;
    int seed;           // 1 on the objective's step
    int obj_face, obj_axis;  // obj_face LB_OBJ_DISSIPATION: the dissipation objective
    double obj_scale;   // 1 / (fluid sites on the objective face)
    const void* dseed;  // dissipation objective: the seeds of diss_site
};

// Launchers of the per-precision kernels, defined in lb_kernels.h and
// instantiated once per precision P in its own translation unit (lb_inst.hip,
// built with LB_INST_P = P). The adjoint launchers exist for LB_FP64 and
// LB_FP32_SHIFTED only.
template <int P> void launch_step_rlx(const StepParams& p, int rlx, bool odd, hipStream_t st);
template <int P> void launch_adj_rlx(const StepParams& p, const AdjParams& x, int rlx, bool odd, hipStream_t st);
template <int P> void launch_diss_rlx(const StepParams& p, int rlx, bool odd, hipStream_t st);
template <int P> void launch_init(const StepParams& p, hipStream_t st);
template <int P> void launch_post_f(const StepParams& p, int next_odd, double* out, hipStream_t st);

}  // namespace lb
