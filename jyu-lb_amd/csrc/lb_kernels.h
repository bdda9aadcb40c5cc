// Per-precision kernels of the D3Q19 engine: the fused AA stream+collide step
// (dense and sparse layouts, bulk and special sites), the dissipation-objective
// gather, the adjoint AA step, the initial populations and the f1 export,
// with their launchers. Included by lb_inst.hip only, which instantiates the
// launchers for one storage variant per translation unit, so the variants
// compile in parallel.
#pragma once
#include "lb_common.h"
#include "lb_adjoint.h"

namespace lb {

// Build switches of the bulk kernels. LB_PLAIN_PATH=1: waves whose 64 sites
// are all plain fluid (wave class 0) take the path that loads no link mask;
// 0 sends every wave through the mask path. LB_WAVES_PER_EU > 0 asks the
// compiler for that many waves per SIMD in k_step (0: its own choice; 5 or 6
// for the FP64 kernels spill registers and measured about half the speed).
#ifndef LB_PLAIN_PATH
#define LB_PLAIN_PATH 1
#endif
#ifndef LB_WAVES_PER_EU
#define LB_WAVES_PER_EU 0
#endif
#if LB_WAVES_PER_EU > 0
#define LB_OCC __attribute__((amdgpu_waves_per_eu(LB_WAVES_PER_EU, 8)))
#else
#define LB_OCC
#endif

// One site of the fused step, in one of three modes:
//   MODE_PLAIN    the site and its whole wave are interior fluid sites with no
//                 bounce-back link: no link mask, loads issue immediately
//   MODE_BULK     ordinary fluid site: bounce-back links from the link mask,
//                 including Ladd's moving-wall term
//   MODE_SPECIAL  sites with halo-crossing links (periodic / symmetric faces)
//                 or on a Zou-He face; processed by a separate
//                 kernel over a compact list so that these rarely taken paths
//                 do not raise the register budget of the bulk kernel
constexpr int MODE_PLAIN = 0, MODE_BULK = 1, MODE_SPECIAL = 2;

// Zou-He faces of site (i,j,k), in the order 1X, X1, 1Y, Y1, 1Z, Z1
// (isothermal.pyx:1280-1291), on the streamed populations f
template <class TC, bool SH>
LBD void zouhe_site(const StepParams& p, uint32_t m, int i, int j, int k, TC rho0, TC* f) {
#pragma unroll
    for (int fc = 0; fc < 6; ++fc) {
        if (!(m & (M_ZH0 << fc))) continue;
        int64_t fi;
        if (fc < 2) fi = (int64_t)(j + 1) * p.NK + (k + 1);
        else if (fc < 4) fi = (int64_t)(i + 1) * p.NK + (k + 1);
        else fi = (int64_t)(i + 1) * p.NJ + (j + 1);
        TC vb[3] = {(TC)p.fvel[fc][3 * fi], (TC)p.fvel[fc][3 * fi + 1], (TC)p.fvel[fc][3 * fi + 2]};
        if (p.d.bc[fc] == FIXED_DEN) zh_fix_den<TC, SH>(fc, (TC)p.fden[fc][fi], (TC)(p.fden[fc][fi] - p.rho0), vb, f);
        else zh_fix_vel<TC, SH>(fc, vb, rho0, f);
    }
}

// Dissipation objective J = (1/N) sum over the fluid sites of Pi^neq:Pi^neq,
// the non-equilibrium momentum flux of the final step's collision input f
// (streamed, bounce-back and Zou-He applied): Pi^neq = sum_l c_l c_l f_l -
// rho u* u* - rho/3 I with u* = j/rho + a/2. Writes the site's term and the
// seeds dJ/df_l = (2/N) [c_l.Pi.c_l - 2 (c_l + a/2).Pi.u* + u*.Pi.u* -
// tr(Pi)/3] and dJ/da = -(2/N) rho Pi.u* (Pi = Pi^neq, symmetric).
template <class TC, bool SH>
LBD void diss_site(const StepParams& p, const uint32_t c, const TC* f, const TC* acc) {
    TC s0 = TC(0), j[3] = {TC(0), TC(0), TC(0)}, Pi[6] = {TC(0), TC(0), TC(0), TC(0), TC(0), TC(0)};
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const TC cx = TC(CI[l]), cy = TC(CJ[l]), cz = TC(CK[l]);
        s0 += f[l];
        j[0] += cx * f[l];
        j[1] += cy * f[l];
        j[2] += cz * f[l];
        Pi[0] += cx * cx * f[l];
        Pi[1] += cx * cy * f[l];
        Pi[2] += cx * cz * f[l];
        Pi[3] += cy * cy * f[l];
        Pi[4] += cy * cz * f[l];
        Pi[5] += cz * cz * f[l];
    }
    const TC rho0 = (TC)p.rho0, third = TC(1.0) / TC(3.0);
    const TC rho = SH ? rho0 + s0 : s0;
    if (SH) {  // shifted populations: sum_l w_l c_l c_l rho0 = rho0/3 I
        Pi[0] += rho0 * third;
        Pi[3] += rho0 * third;
        Pi[5] += rho0 * third;
    }
    TC u[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) u[a] = j[a] / rho + TC(0.5) * acc[a];
    const TC N[6] = {Pi[0] - rho * u[0] * u[0] - rho * third, Pi[1] - rho * u[0] * u[1], Pi[2] - rho * u[0] * u[2],
                     Pi[3] - rho * u[1] * u[1] - rho * third, Pi[4] - rho * u[1] * u[2],
                     Pi[5] - rho * u[2] * u[2] - rho * third};
    const TC Nu[3] = {N[0] * u[0] + N[1] * u[1] + N[2] * u[2], N[1] * u[0] + N[3] * u[1] + N[4] * u[2],
                      N[2] * u[0] + N[4] * u[1] + N[5] * u[2]};
    const TC uNu = u[0] * Nu[0] + u[1] * Nu[1] + u[2] * Nu[2];
    const TC tr = N[0] + N[3] + N[5];
    const TC sc = TC(2.0) * (TC)p.dscale;
    TC* S = (TC*)p.dseed;
    const int64_t nq = p.nq;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const TC cx = TC(CI[l]), cy = TC(CJ[l]), cz = TC(CK[l]);
        const TC cNc = cx * (N[0] * cx + N[1] * cy + N[2] * cz) + cy * (N[1] * cx + N[3] * cy + N[4] * cz) +
                       cz * (N[2] * cx + N[4] * cy + N[5] * cz);
        const TC wNu =
            (cx + TC(0.5) * acc[0]) * Nu[0] + (cy + TC(0.5) * acc[1]) * Nu[1] + (cz + TC(0.5) * acc[2]) * Nu[2];
        S[(int64_t)l * nq + c] = sc * (cNc - TC(2.0) * wNu + uNu - tr * third);
    }
#pragma unroll
    for (int a = 0; a < 3; ++a) S[(int64_t)(Q + a) * nq + c] = -sc * rho * Nu[a];
    S[(int64_t)(Q + 3) * nq + c] =
        N[0] * N[0] + N[3] * N[3] + N[5] * N[5] + TC(2.0) * (N[1] * N[1] + N[2] * N[2] + N[4] * N[4]);
}

// PULL (with ODD): the odd step's loads, and the even step's stores into p.f2
// (the dense form of the sparse layout's pull step, k_step_sp_pull).
template <int P, int RLX, bool ODD, int MODE, bool DISS = false, bool PULL = false>
LBD void step_site(const StepParams& p, const uint32_t c, const uint32_t m) {
    constexpr bool PLAIN = MODE != MODE_SPECIAL;  // no special handling needed
    using TS = typename Var<P>::TS;
    using TC = typename Var<P>::TC;
    constexpr bool SH = Var<P>::SH;
    TS* F = (TS*)p.f;
    const int64_t nq = p.nq;
    const int64_t e = eidx(p, c);
    int i = 0, j = 0, k = 0;
    if (MODE == MODE_SPECIAL) decompose(c, p.d.ny, p.d.nzp, i, j, k);

    // Bulk sites address populations as a wave-uniform row base plus the
    // unsigned lane offset kk (a wave never straddles rows: nzp % 64 == 0), so
    // every load and store uses a scalar base and one shared 32-bit offset.
    // Bounce-back links are exec-masked loads from the site's own slot.
    const uint32_t nzp = (uint32_t)p.d.nzp;
    const uint32_t row = PLAIN ? __builtin_amdgcn_readfirstlane(c / nzp) : 0;
    const uint32_t kk = c - row * nzp;
    TS* const Fr = F + (int64_t)row * p.rs;
    auto own = [&](int l) { return uniform_ptr(Fr + (int64_t)l * p.ps); };
    auto up = [&](int l) {
        return uniform_ptr(Fr - ((int64_t)CI[l] * p.d.ny + CJ[l]) * p.rs + (int64_t)opp(l) * p.ps - CK[l]);
    };
    // z-periodic wrap of link l for this lane: the upstream element is the
    // other end of the same upstream row (a uniform address)
    const uint32_t nzm = (uint32_t)p.d.nz - 1u;
    auto wraps = [&](int l) {
        return MODE == MODE_BULK && (m & M_KWRAP) && ((CK[l] > 0 && kk == 0) || (CK[l] < 0 && kk == nzm));
    };
    auto upw = [&](int l) { return up(l) + CK[l] + (CK[l] > 0 ? nzm : 0u); };
    // wave-uniform: does any lane wrap? Other waves run the pull and store
    // loops without the per-lane wrap tests.
    const bool anyw = MODE == MODE_BULK && ODD && __ballot((m & M_KWRAP) != 0) != 0;
    constexpr bool NTLE = (LB_NT_MASK & 1) != 0, NTSE = (LB_NT_MASK & 2) != 0;
    constexpr bool NTLO = (LB_NT_MASK & 4) != 0, NTSO = (LB_NT_MASK & 8) != 0;

    // ---- stream (pull) ----
    TC f[Q];
    if constexpr (PLAIN && !ODD) {
#pragma unroll
        for (int l = 0; l < Q; ++l) f[l] = ldf<P, NTLE>(own(l) + kk, p.inv_scale);
    } else if constexpr (PLAIN) {
        auto pull = [&](auto wrap) {
            constexpr bool WRAP = decltype(wrap)::value;
#pragma unroll
            for (int l = 0; l < Q; ++l) {
                // k-shifted pulls straddle cache lines shared with the next wave:
                // LB_NT_MASK bit 4 keeps those cached and streams the rest
                constexpr bool NTK = (LB_NT_MASK & 16) ? true : NTLO;
                if (l == REST) f[l] = ldf<P, NTK>(own(l) + kk, p.inv_scale);
                else if (MODE == MODE_PLAIN || !(m & (1u << l))) {
                    if (CK[l] == 0) f[l] = ldf<P, NTK>(up(l) + kk, p.inv_scale);
                    else if (WRAP && wraps(l)) f[l] = ldf<P, NTLO>(upw(l), p.inv_scale);
                    else f[l] = ldf<P, NTLO>(up(l) + kk, p.inv_scale);
                } else f[l] = ldf<P, NTK>(own(l) + kk, p.inv_scale);
            }
        };
        if (anyw) pull(std::true_type{});
        else pull(std::false_type{});
    } else if constexpr (!ODD) {
#pragma unroll
        for (int l = 0; l < Q; ++l) f[l] = ldf<P, NTLE>(F + e + (int64_t)l * p.ps, p.inv_scale);
    } else {
#pragma unroll
        for (int l = 0; l < Q; ++l)
            f[l] = ldf<P, NTLO>(F + odd_loc<true>(p, m, e, l, i, j, k), p.inv_scale);
    }
    const TC rho0 = (TC)p.rho0;

    // ---- Ladd moving walls on bounce-back links (previous-step density) ----
    if (MODE != MODE_PLAIN && p.moving && (m & M_MOVING)) ladd_apply<TC>(p, c, m, f);

    // ---- Zou-He faces ----
    if (!PLAIN && (m & M_ZH_ANY)) zouhe_site<TC, SH>(p, m, i, j, k, rho0, f);

    // ---- acceleration and collision ----
    TC acc[3];
    const TC* A = (const TC*)p.acc;
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a] = (TC)p.g[a] + (A ? A[(int64_t)a * nq + c] : TC(0));
    if constexpr (DISS) {  // objective gather: nothing is stored to f
        diss_site<TC, SH>(p, c, f, acc);
        return;
    }
    TC prm[2] = {(TC)p.prm[0], (TC)p.prm[1]};
    TC vel[3];
    TC ds;
    const TC den = collide<RLX, TC, SH, P == LB_FP32_SHIFTED>(f, acc, prm, rho0, vel, ds);

    if (p.store_dv) {
        // shifted variants: rho0 + (den - rho0) in FP64, so the stored density
        // keeps the deviation's precision (it feeds FIXED_DEN face data back)
        p.den[c] = SH ? p.rho0 + (double)ds : (double)den;
        p.vel[c] = (double)vel[0];
        p.vel[nq + c] = (double)vel[1];
        p.vel[2 * nq + c] = (double)vel[2];
    } else if (MODE != MODE_PLAIN && p.moving && (m & M_MOVING)) {
        p.den[c] = SH ? p.rho0 + (double)ds : (double)den;
    }
    if (p.rec) {
        TC* R = (TC*)p.rec;
        R[c] = den;
        R[nq + c] = vel[0];
        R[2 * nq + c] = vel[1];
        R[3 * nq + c] = vel[2];
    }

    // ---- store ----
    if constexpr (PLAIN && !ODD) {
#pragma unroll
        for (int l = 0; l < Q; ++l) stf<P, NTSE>(own(opp(l)) + kk, f[l], p.scale);
    } else if constexpr (PLAIN && PULL) {
        TS* const Gr = (TS*)p.f2 + (int64_t)row * p.rs;
#pragma unroll
        for (int l = 0; l < Q; ++l) stf<P, NTSE>(uniform_ptr(Gr + (int64_t)opp(l) * p.ps) + kk, f[l], p.scale);
    } else if constexpr (PLAIN) {
        auto push = [&](auto wrap) {
            constexpr bool WRAP = decltype(wrap)::value;
#pragma unroll
            for (int l = 0; l < Q; ++l) {
                constexpr bool NTK = (LB_NT_MASK & 16) ? true : NTSO;
                if (l == REST) stf<P, NTK>(own(l) + kk, f[l], p.scale);
                else if (MODE == MODE_PLAIN || !(m & (1u << l))) {
                    if (CK[l] == 0) stf<P, NTK>(up(l) + kk, f[opp(l)], p.scale);
                    else if (WRAP && wraps(l)) stf<P, NTSO>(upw(l), f[opp(l)], p.scale);
                    else stf<P, NTSO>(up(l) + kk, f[opp(l)], p.scale);
                } else stf<P, NTK>(own(l) + kk, f[opp(l)], p.scale);
            }
        };
        if (anyw) push(std::true_type{});
        else push(std::false_type{});
    } else if constexpr (!ODD) {
#pragma unroll
        for (int l = 0; l < Q; ++l) stf<P, NTSE>(F + e + (int64_t)opp(l) * p.ps, f[l], p.scale);
    } else if constexpr (PULL) {
        TS* const G = (TS*)p.f2;
#pragma unroll
        for (int l = 0; l < Q; ++l) stf<P, NTSE>(G + e + (int64_t)opp(l) * p.ps, f[l], p.scale);
    } else {
#pragma unroll
        for (int l = 0; l < Q; ++l) stf<P, NTSO>(F + odd_loc<true>(p, m, e, l, i, j, k), f[opp(l)], p.scale);
    }
}

// Tile of this workgroup. With p.xcd the tiles are dealt so that each of the
// 8 XCDs (which receive workgroups round-robin) works through one contiguous
// range of rows; the grid is then padded to a multiple of 8. It is set for
// lattices of at least LB_XCD_REMAP_CELLS cells (an environment variable of
// that name, read at lb_create, overrides it): the 768x384x384 porous
// medium ran 5-7% faster with it, while the 256^3 cavity, the 512x128x128
// duct and the 128x64x64 porous channel ran 1.5-3.8% faster without in the
// dense layout (the sparse layout's own defaults: layout_defaults).
// With p.xcd = C other than XCD_CONTIGUOUS, the XCDs take chunks of C
// consecutive tiles in turn (chunk m to XCD m % 8) and the grid is padded to
// a multiple of 8C.
LBD uint32_t tile_index(const StepParams& p) {
    if (!p.xcd) return blockIdx.x;
    const uint32_t x = blockIdx.x & 7u, i = blockIdx.x >> 3;
    const uint32_t c = p.xcd == XCD_CONTIGUOUS ? gridDim.x >> 3 : p.xcd;
    return ((i / c) * 8 + x) * c + i % c;
}

// grid of a launch that reads tile_index
inline unsigned tile_grid(const StepParams& p) {
    if (!p.xcd) return p.ntiles;
    const unsigned q = p.xcd == XCD_CONTIGUOUS ? 8u : 8u * p.xcd;
    return (p.ntiles + q - 1) / q * q;
}

// Mask bits that send a site to the special-site kernel. An even step reads
// and writes only the site's own slots, so sites whose links cross a periodic
// or symmetric face (M_BND) need no address resolution then and stay in the
// bulk kernel; Zou-He sites need the face data on both parities.
template <bool ODD>
LBD constexpr uint32_t special_bits() { return ODD ? M_SPECIAL : M_ZH_ANY; }

// Bulk step: every fluid site that is not special for this parity.
template <int P, int RLX, bool ODD, bool DISS = false, bool PULL = false>
LBD void step_bulk(const StepParams& p) {
    const uint32_t tile = tile_index(p);
    if (tile >= p.ntiles) return;
    const uint32_t c = tile * p.tile + threadIdx.x;
    // nq is a multiple of 64, so a wave is entirely inside or outside [0, nq)
    const uint32_t w = __builtin_amdgcn_readfirstlane(c >> 6);
    if (w >= p.nwaves) return;
    if (LB_PLAIN_PATH && p.wclass[w] == 0) {
        step_site<P, RLX, ODD, MODE_PLAIN, DISS, PULL>(p, c, M_FLUID);
        return;
    }
    const uint32_t m = p.mask[c];
    if ((m & (M_FLUID | special_bits<ODD>())) != M_FLUID) return;
    step_site<P, RLX, ODD, MODE_BULK, DISS, PULL>(p, c, m);
}

template <int P, int RLX, bool ODD, bool PULL = false>
__global__ __launch_bounds__(STEP_TILE_MAX) LB_OCC void k_step(StepParams p) {
    step_bulk<P, RLX, ODD, false, PULL>(p);
}

// FP16 storage with TRT: the same kernel held to 64 VGPRs (8 waves per SIMD),
// which it reaches without spills (69 VGPRs otherwise, 7 waves; ~4% slower on
// the TRT cavity). BGK and the regularized scheme measured 9-10% slower with
// the hint (128x64x64 porous channel, 512x128x128 duct) and keep k_step.
template <int RLX, bool ODD, bool PULL = false>
__global__ __launch_bounds__(STEP_TILE_MAX) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_step_h(StepParams p) {
    step_bulk<LB_FP16_STORAGE, RLX, ODD, false, PULL>(p);
}

// Special sites, one thread per entry of the compact list.
template <int P, int RLX, bool ODD, bool PULL = false>
__global__ __launch_bounds__(TILE) void k_step_special(StepParams p) {
    const uint32_t t = blockIdx.x * TILE + threadIdx.x;
    if (t >= (ODD ? p.nspecial : p.nzh)) return;
    const uint32_t c = p.special[t];
    step_site<P, RLX, ODD, MODE_SPECIAL, false, PULL>(p, c, p.mask[c]);
}

// Dissipation objective of the step about to run (diss_site): the step's
// pre-collision part at every fluid site, nothing stored to f.
template <int P, int RLX, bool ODD>
__global__ __launch_bounds__(STEP_TILE_MAX) void k_diss(StepParams p) {
    step_bulk<P, RLX, ODD, true>(p);
}

template <int P, int RLX, bool ODD>
__global__ __launch_bounds__(TILE) void k_diss_special(StepParams p) {
    const uint32_t t = blockIdx.x * TILE + threadIdx.x;
    if (t >= (ODD ? p.nspecial : p.nzh)) return;
    const uint32_t c = p.special[t];
    step_site<P, RLX, ODD, MODE_SPECIAL, true>(p, c, p.mask[c]);
}

// ---------------------------------------------------------------------------
// sparse layout: fluid sites only
// ---------------------------------------------------------------------------
// Compact site n (fluid sites in ascending dense order, so consecutive threads
// mostly hold consecutive k of one lattice row) keeps population l at
// f[l*ps + n]. An even step reads and writes only these slots. An odd step
// reads link l from tab[link_slot(l)*ps + n], the element that the dense
// resolver (odd_loc) names for it, mapped to compact indices at
// lb_finalize: upstream slot, own slot for bounce-back, or the periodic /
// mirrored source; it writes f*_opp(l) back there. FULL sites may carry
// moving-wall or Zou-He bits, read from the dense mask; lattices with
// neither run the FULL = false kernel, which loads no mask.
//
// The step reads the table in a compact form (k_sparse_pack). Per wave of 64
// compact sites and link: a base and an alternate location (32 bits each).
// Per site and link a 4-bit code: below SP_ALT, the location is base + lane
// + code (in a run of fluid sites the upstream locations advance with the
// lane, so the codes stay small); SP_ALT, the wave's alternate location (it
// takes the one outlier of most waves: the lane whose link wraps around a
// periodic face at the end of a row); SP_BOUNCE, the own slot l; SP_FAR, a
// location read from the 32-bit table (0.2% of the links on a random-sphere
// medium). A site's 18 codes are two 32-bit words of eight links and a byte
// of two, in two rows of ps words and a row of ps bytes: three loads per
// wave, 9 bytes per site (5-bit codes in three words, 12 bytes, measured
// 0.7% slower in FP64 and 2% in FP32; 8-bit codes, 18 bytes, slower still).

// Odd-step locations of the 18 links of compact site n (loc[REST]: own slot).
// The codes are decoded with selects, so that all 18 population loads after
// this can be in flight together; the rare SP_FAR links are fixed up in one
// wave-uniform branch.
LBD void sparse_locs(const StepParams& p, const uint32_t n, uint32_t* loc) {
    const uint32_t ps = (uint32_t)p.ps;
    const uint32_t* B = p.tabw + (size_t)__builtin_amdgcn_readfirstlane(n >> 6) * SP_WAVE_WORDS;
    uint32_t base[Q - 1], alt[Q - 1], w[SP_WORDS + 1];
#pragma unroll
    for (int sl = 0; sl < Q - 1; ++sl) {
        base[sl] = B[sl];
        alt[sl] = B[Q - 1 + sl];
    }
    // used here: the loads are not sunk into a branch per link
#pragma unroll
    for (int sl = 0; sl < Q - 1; ++sl) asm volatile("" ::"s"(base[sl]), "s"(alt[sl]));
    // read once per step: streamed past L2, where the gathers' partly used
    // lines wait for the neighbouring waves (FP16 0.6% and FP32 about 1%
    // faster on the 768x384x384 medium, FP64 unchanged)
#pragma unroll
    for (int q = 0; q < SP_WORDS; ++q) w[q] = __builtin_nontemporal_load(p.tabc + (size_t)q * ps + n);
    w[SP_WORDS] = __builtin_nontemporal_load(p.tabb + n);
    const uint32_t lane = n & 63u;  // waves start at multiples of 64
    uint32_t far = 0;
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        if (l == REST) {
            loc[l] = REST * ps + n;
            continue;
        }
        const int sl = link_slot(l);
        const uint32_t code = sp_code(w, sl);
        const uint32_t near = base[sl] + lane + code;  // modulo 2^32: the base may be below 0
        loc[l] = code < SP_ALT ? near : code == SP_ALT ? alt[sl] : (uint32_t)l * ps + n;
        far |= (code == SP_FAR ? 1u : 0u) << sl;
    }
    if (__ballot(far != 0)) {
#pragma unroll
        for (int l = 0; l < Q; ++l)
            if (l != REST && ((far >> link_slot(l)) & 1u)) loc[l] = p.tab[link_slot(l) * ps + n];
    }
}

// Step of compact site n; loc: the odd step's locations (sparse_locs). PULL
// (with ODD): the odd step's loads, and the even step's stores into p.f2.
template <int P, int RLX, bool ODD, bool FULL, bool DISS = false, bool PULL = false>
LBD void step_sparse_at(const StepParams& p, const uint32_t n, const uint32_t* loc) {
    using TS = typename Var<P>::TS;
    using TC = typename Var<P>::TC;
    constexpr bool SH = Var<P>::SH;
    TS* F = (TS*)p.f;
    const uint32_t ps = (uint32_t)p.ps;  // 19 * ps < 2^32 (lb_finalize)
    const TC* A = (const TC*)p.acc;
    const bool need_c = FULL || DISS || A || p.store_dv;  // uniform
    const uint32_t c = need_c ? p.site[n] : 0u;
    const uint32_t m = FULL ? p.mask[c] : M_FLUID;
    constexpr bool NTLE = (LB_NT_MASK & 1) != 0, NTSE = (LB_NT_MASK & 2) != 0;
    constexpr bool NTLO = (LB_NT_MASK & 4) != 0, NTSO = (LB_NT_MASK & 8) != 0;

    // population l of this site: a wave-uniform base plus the 32-bit lane
    // offset n (n * sizeof(TS) < 2^32)
    auto own = [&](int l) { return uniform_ptr(F + (int64_t)l * ps) + n; };
    TC f[Q];
    if constexpr (!ODD) {
#pragma unroll
        for (int l = 0; l < Q; ++l) f[l] = ldf<P, NTLE>(own(l), p.inv_scale);
    } else {
#pragma unroll
        for (int l = 0; l < Q; ++l) f[l] = ldf<P, NTLO>(l == REST ? own(l) : F + loc[l], p.inv_scale);
    }
    const TC rho0 = (TC)p.rho0;
    if (FULL && p.moving && (m & M_MOVING)) ladd_apply<TC>(p, c, m, f);
    if (FULL && (m & M_ZH_ANY)) {
        int i, j, k;
        decompose(c, p.d.ny, p.d.nzp, i, j, k);
        zouhe_site<TC, SH>(p, m, i, j, k, rho0, f);
    }
    const int64_t nq = p.nq;
    TC acc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a] = (TC)p.g[a] + (A ? A[(int64_t)a * nq + c] : TC(0));
    if constexpr (DISS) {  // objective gather: nothing is stored to f
        diss_site<TC, SH>(p, c, f, acc);
        return;
    }
    TC prm[2] = {(TC)p.prm[0], (TC)p.prm[1]};
    TC vel[3];
    TC ds;
    const TC den = collide<RLX, TC, SH, P == LB_FP32_SHIFTED>(f, acc, prm, rho0, vel, ds);
    if (p.store_dv) {
        p.den[c] = SH ? p.rho0 + (double)ds : (double)den;
        p.vel[c] = (double)vel[0];
        p.vel[nq + c] = (double)vel[1];
        p.vel[2 * nq + c] = (double)vel[2];
    } else if (FULL && p.moving && (m & M_MOVING)) {
        p.den[c] = SH ? p.rho0 + (double)ds : (double)den;
    }
    if (p.rec) {  // by compact site: rows of ps values
        TC* R = (TC*)p.rec;
        R[n] = den;
        R[ps + n] = vel[0];
        R[2 * ps + n] = vel[1];
        R[3 * ps + n] = vel[2];
    }
    if constexpr (!ODD) {
#pragma unroll
        for (int l = 0; l < Q; ++l) stf<P, NTSE>(own(opp(l)), f[l], p.scale);
    } else if constexpr (PULL) {
        TS* G = (TS*)p.f2;
#pragma unroll
        for (int l = 0; l < Q; ++l) stf<P, NTSE>(uniform_ptr(G + (int64_t)opp(l) * ps) + n, f[l], p.scale);
    } else {
#pragma unroll
        for (int l = 0; l < Q; ++l) stf<P, NTSO>(l == REST ? own(l) : F + loc[l], f[opp(l)], p.scale);
    }
}

template <int P, int RLX, bool ODD, bool FULL, bool DISS = false>
LBD void step_sparse(const StepParams& p, const uint32_t n) {
    uint32_t loc[Q];
    if constexpr (ODD) sparse_locs(p, n, loc);
    step_sparse_at<P, RLX, ODD, FULL, DISS>(p, n, loc);
}

template <int P, int RLX, bool ODD, bool FULL>
__global__ __launch_bounds__(STEP_TILE_MAX) void k_step_sp(StepParams p) {
    const uint32_t tile = tile_index(p);
    if (tile >= p.ntiles) return;
    const uint32_t n = tile * p.tile + threadIdx.x;
    if (n >= p.nf) return;
    step_sparse<P, RLX, ODD, FULL>(p, n);
}

// The sparse layout's pull step: from the state an even step leaves (site y's
// post-collision population l in its own slot opp(l)) it gathers like the
// odd step, and stores like the even step, into the other population array
// p.f2, which then holds the next state in the same form. The odd step's
// stores go back to the gathered locations, which the upstream rows'
// compact indices place anywhere within a 128-byte line; partial-line stores
// measured a quarter slower than aligned ones, while misaligned loads cost
// nothing. The pull step's stores are all aligned, at the price of a second
// population array.
template <int P, int RLX, bool FULL>
__global__ __launch_bounds__(STEP_TILE_MAX) void k_step_sp_pull(StepParams p) {
    const uint32_t tile = tile_index(p);
    if (tile >= p.ntiles) return;
    const uint32_t n = tile * p.tile + threadIdx.x;
    if (n >= p.nf) return;
    uint32_t loc[Q];
    sparse_locs(p, n, loc);
    step_sparse_at<P, RLX, true, FULL, false, true>(p, n, loc);
}

// the dissipation objective's gather in the sparse layout (see k_diss)
template <int P, int RLX, bool ODD, bool FULL>
__global__ __launch_bounds__(STEP_TILE_MAX) void k_diss_sp(StepParams p) {
    const uint32_t tile = tile_index(p);
    if (tile >= p.ntiles) return;
    const uint32_t n = tile * p.tile + threadIdx.x;
    if (n >= p.nf) return;
    step_sparse<P, RLX, ODD, FULL, true>(p, n);
}

// ---------------------------------------------------------------------------
// adjoint step
// ---------------------------------------------------------------------------
// The adjoint of an AA step is an AA step of the same parity on the cotangent
// array: site x reads the cotangents of its post-collision populations from
// the locations its forward step wrote them to, pulls them back through the
// collision and the Zou-He faces (reverse order) and writes the cotangents of
// its pre-collision populations to the locations its forward step read them
// from. Those are the same 19 locations, so the step runs in place.

// Per-site core of the adjoint step: the collision's vector-Jacobian product
// from the recorded (rho, u*), the objective's seed, the acceleration
// cotangent and, on Zou-He sites (zh), the faces' VJP in reverse order.
// lam: cotangents of the post-collision populations; mu: of the streamed ones.
// c: the site's cell (read only for a per-site acceleration field and the
// dissipation seeds); r: its index in the record and in x.gacc, rows of rs
// values (the cell and nq in the dense layouts, the compact site and ps in
// the sparse layout).
template <int P, int RLX>
LBD void adj_core(const StepParams& p, const AdjParams& x, const uint32_t c, const uint32_t r, const int64_t rs,
                  const uint32_t m, const bool zh, int i, int j, int k, const typename Var<P>::TC* lam,
                  typename Var<P>::TC* mu) {
    using TC = typename Var<P>::TC;
    const int64_t nq = p.nq;
    const TC* R = (const TC*)x.rec;
    const TC rho = R[r];
    const TC us[3] = {R[rs + r], R[2 * rs + r], R[3 * rs + r]};
    TC acc[3];
    const TC* A = (const TC*)p.acc;
#pragma unroll
    for (int a = 0; a < 3; ++a) acc[a] = (TC)p.g[a] + (A ? A[(int64_t)a * nq + c] : TC(0));
    TC prm[2] = {(TC)p.prm[0], (TC)p.prm[1]};
    TC ga[3] = {TC(0), TC(0), TC(0)};
    collide_vjp<RLX, TC>(lam, rho, us, acc, prm, mu, ga);
    if (x.seed && x.obj_face == LB_OBJ_DISSIPATION) {
        // dissipation objective: seeds written by diss_site in the forward pass
        const TC* S = (const TC*)x.dseed;
#pragma unroll
        for (int l = 0; l < Q; ++l) mu[l] += S[(int64_t)l * nq + c];
#pragma unroll
        for (int a = 0; a < 3; ++a) ga[a] += S[(int64_t)(Q + a) * nq + c];
    } else if (x.seed) {
        // objective: mean over the face layer of rho u*_axis = j_axis + rho a_axis / 2
        const int ax = x.obj_face >> 1;
        const int pos = ax == 0 ? i : ax == 1 ? j : k;
        const int n = ax == 0 ? p.d.nx : ax == 1 ? p.d.ny : p.d.nz;
        if (pos == ((x.obj_face & 1) ? n - 1 : 0)) {
            const TC sc = (TC)x.obj_scale;
#pragma unroll
            for (int l = 0; l < Q; ++l) mu[l] += sc * (TC(ccomp(l, x.obj_axis)) + TC(0.5) * acc[x.obj_axis]);
            ga[x.obj_axis] += sc * TC(0.5) * rho;
        }
    }
    if (x.gacc) {
        TC* G = (TC*)x.gacc;
#pragma unroll
        for (int a = 0; a < 3; ++a) G[(int64_t)a * rs + r] += ga[a];
    }
    if (zh) {
        for (int fc = 5; fc >= 0; --fc) {
            if (!(m & (M_ZH0 << fc))) continue;
            int64_t fi;
            if (fc < 2) fi = (int64_t)(j + 1) * p.NK + (k + 1);
            else if (fc < 4) fi = (int64_t)(i + 1) * p.NK + (k + 1);
            else fi = (int64_t)(i + 1) * p.NJ + (j + 1);
            const TC vb[3] = {(TC)p.fvel[fc][3 * fi], (TC)p.fvel[fc][3 * fi + 1], (TC)p.fvel[fc][3 * fi + 2]};
            const int kind = p.d.bc[fc];
            const TC dbc = kind == FIXED_DEN ? (TC)p.fden[fc][fi] : rho;
            TC nu[Q], gv[3] = {TC(0), TC(0), TC(0)};
            zh_vjp<TC>(fc, kind, dbc, vb, mu, nu, gv);
#pragma unroll
            for (int l = 0; l < Q; ++l) mu[l] = nu[l];
            if (x.gface && fc == x.gface_id)
                for (int a = 0; a < 3; ++a) x.gface[3 * fi + a] += (double)gv[a];
        }
    }
}

// One site of the adjoint step. MODE_PLAIN / MODE_BULK address the
// cotangents like the bulk step kernel (wave-uniform row bases, bounce-back
// links from the link mask); MODE_SPECIAL resolves every link. PULL (with
// ODD): the pull form of the sparse layout's k_adj_sp_pull, the odd step's
// gathers from x.lam and the even step's stores into x.lam2.
template <int P, int RLX, bool ODD, int MODE, bool PULL = false>
LBD void adj_site(const StepParams& p, const AdjParams& x, const uint32_t c, const uint32_t m) {
    using TC = typename Var<P>::TC;
    constexpr bool SPECIAL = MODE == MODE_SPECIAL;
    const int64_t nq = p.nq;
    TC* L = (TC*)x.lam;
    int i = 0, j = 0, k = 0;
    if (SPECIAL || x.seed) decompose(c, p.d.ny, p.d.nzp, i, j, k);
    const int64_t e = eidx(p, c);
    const uint32_t nzp = (uint32_t)p.d.nzp;
    const uint32_t row = SPECIAL ? 0u : __builtin_amdgcn_readfirstlane(c / nzp);
    const uint32_t kk = c - row * nzp;
    TC* const Lr = L + (int64_t)row * p.rs;
    auto own = [&](int l) { return uniform_ptr(Lr + (int64_t)l * p.ps); };
    auto up = [&](int l) {
        return uniform_ptr(Lr - ((int64_t)CI[l] * p.d.ny + CJ[l]) * p.rs + (int64_t)opp(l) * p.ps - CK[l]);
    };
    // location of the cotangent of f_l's forward read (= of f*_opp(l)'s write)
    const uint32_t nzm = (uint32_t)p.d.nz - 1u;
    auto loc = [&](int l) -> TC* {
        if constexpr (SPECIAL) return L + odd_loc<true>(p, m, e, l, i, j, k);
        else if (l == REST || (MODE == MODE_BULK && (m & (1u << l)))) return own(l) + kk;
        else if (MODE == MODE_BULK && (m & M_KWRAP) && ((CK[l] > 0 && kk == 0) || (CK[l] < 0 && kk == nzm)))
            return up(l) + CK[l] + (CK[l] > 0 ? nzm : 0u);
        else return up(l) + kk;
    };
    constexpr bool NTE = (LB_NT_MASK & 1) != 0;
    TC lam[Q], mu[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        if constexpr (ODD) lam[opp(l)] = *loc(l);
        else if constexpr (SPECIAL) lam[l] = L[e + (int64_t)opp(l) * p.ps];
        else lam[l] = ld_stream<NTE>(own(opp(l)) + kk);
    }
    adj_core<P, RLX>(p, x, c, c, nq, m, SPECIAL && (m & M_ZH_ANY), i, j, k, lam, mu);
    if constexpr (PULL) {
        TC* L2 = (TC*)x.lam2;
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            if constexpr (SPECIAL) L2[e + (int64_t)l * p.ps] = mu[l];
            else st_stream<NTE>(uniform_ptr(L2 + (int64_t)row * p.rs + (int64_t)l * p.ps) + kk, mu[l]);
        }
        return;
    }
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        if constexpr (ODD) *loc(l) = mu[l];
        else if constexpr (SPECIAL) L[e + (int64_t)l * p.ps] = mu[l];
        else st_stream<NTE>(own(l) + kk, mu[l]);
    }
}

// Adjoint step over the non-special fluid sites (same tiling as k_step).
// PULL (with ODD): the pull form, at either parity; the bulk and special-site
// launches both read x.lam and write x.lam2.
template <int P, int RLX, bool ODD, bool PULL = false>
__global__ __launch_bounds__(TILE) void k_adj(StepParams p, AdjParams x) {
    const uint32_t tile = tile_index(p);
    if (tile >= p.ntiles) return;
    const uint32_t c = tile * p.tile + threadIdx.x;
    const uint32_t w = __builtin_amdgcn_readfirstlane(c >> 6);
    if (w >= p.nwaves) return;
    if (LB_PLAIN_PATH && p.wclass[w] == 0) {
        adj_site<P, RLX, ODD, MODE_PLAIN, PULL>(p, x, c, M_FLUID);
        return;
    }
    const uint32_t m = p.mask[c];
    if ((m & (M_FLUID | special_bits<ODD>())) != M_FLUID) return;
    adj_site<P, RLX, ODD, MODE_BULK, PULL>(p, x, c, m);
}

// Adjoint step over the special sites (compact list).
template <int P, int RLX, bool ODD, bool PULL = false>
__global__ __launch_bounds__(TILE) void k_adj_special(StepParams p, AdjParams x) {
    const uint32_t t = blockIdx.x * TILE + threadIdx.x;
    if (t >= (ODD ? p.nspecial : p.nzh)) return;
    const uint32_t c = p.special[t];
    adj_site<P, RLX, ODD, MODE_SPECIAL, PULL>(p, x, c, p.mask[c]);
}

// Adjoint step of the sparse layout, one thread per compact fluid site: the
// cotangents sit where the forward step's populations do (own slots on even
// steps, the decoded odd-step locations on odd steps). ZH: the lattice has
// Zou-He faces (their VJP is compiled in).
template <int P, int RLX, bool ODD, bool ZH>
__global__ __launch_bounds__(TILE) void k_adj_sp(StepParams p, AdjParams x) {
    using TC = typename Var<P>::TC;
    const uint32_t tile = tile_index(p);
    if (tile >= p.ntiles) return;
    const uint32_t n = tile * p.tile + threadIdx.x;
    if (n >= p.nf) return;
    TC* L = (TC*)x.lam;
    const uint32_t ps = (uint32_t)p.ps;
    const uint32_t c = ZH || x.seed || p.acc ? p.site[n] : 0u;  // uniform choice
    const uint32_t m = ZH ? p.mask[c] : M_FLUID;
    int i = 0, j = 0, k = 0;
    if ((m & M_ZH_ANY) || x.seed) decompose(c, p.d.ny, p.d.nzp, i, j, k);
    constexpr bool NTE = (LB_NT_MASK & 1) != 0;
    uint32_t loc[Q];
    TC lam[Q], mu[Q];
    if constexpr (ODD) {
        sparse_locs(p, n, loc);
#pragma unroll
        for (int l = 0; l < Q; ++l) lam[opp(l)] = L[loc[l]];
    } else {
#pragma unroll
        for (int l = 0; l < Q; ++l) lam[l] = ld_stream<NTE>(uniform_ptr(L + (int64_t)opp(l) * ps) + n);
    }
    adj_core<P, RLX>(p, x, c, n, ps, m, ZH && (m & M_ZH_ANY) != 0, i, j, k, lam, mu);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        if constexpr (ODD) L[loc[l]] = mu[l];
        else st_stream<NTE>(uniform_ptr(L + (int64_t)l * ps) + n, mu[l]);
    }
}

// Pull form of the sparse adjoint step, for a step of either parity. The
// cotangents of the streamed populations are kept in the form an even
// step's reads use: that of population l at site y in y's own slot l. The
// post-collision population l of site x streams to site x + c_l, which is
// the upstream site of link opp(l), so its cotangent is found where the
// forward odd step reads link opp(l) (loc[opp(l)]; on a bounce-back link
// that is x's own slot opp(l), where the bounced population's cotangent is).
// The step gathers those like the odd adjoint step and stores like the even
// one, into x.lam2, which then holds the previous step's cotangents in the
// same form. Arithmetic and results are those of k_adj_sp; all stores are
// aligned, at the price of a second cotangent array.
template <int P, int RLX, bool ZH>
__global__ __launch_bounds__(TILE) void k_adj_sp_pull(StepParams p, AdjParams x) {
    using TC = typename Var<P>::TC;
    const uint32_t tile = tile_index(p);
    if (tile >= p.ntiles) return;
    const uint32_t n = tile * p.tile + threadIdx.x;
    if (n >= p.nf) return;
    const TC* L = (const TC*)x.lam;
    TC* L2 = (TC*)x.lam2;
    const uint32_t ps = (uint32_t)p.ps;
    const uint32_t c = ZH || x.seed || p.acc ? p.site[n] : 0u;  // uniform choice
    const uint32_t m = ZH ? p.mask[c] : M_FLUID;
    int i = 0, j = 0, k = 0;
    if ((m & M_ZH_ANY) || x.seed) decompose(c, p.d.ny, p.d.nzp, i, j, k);
    constexpr bool NTE = (LB_NT_MASK & 1) != 0;
    uint32_t loc[Q];
    TC lam[Q], mu[Q];
    sparse_locs(p, n, loc);
#pragma unroll
    for (int l = 0; l < Q; ++l) lam[opp(l)] = L[loc[l]];
    adj_core<P, RLX>(p, x, c, n, ps, m, ZH && (m & M_ZH_ANY) != 0, i, j, k, lam, mu);
#pragma unroll
    for (int l = 0; l < Q; ++l) st_stream<NTE>(uniform_ptr(L2 + (int64_t)l * ps) + n, mu[l]);
}


// velocity of the site reached from x along +/-c_l for _comp_der_vel:
// resolved through periodic / symmetric halos, zero for WALL/FIXED halos
LBD void nb_vel(const Dims& d, const double* vel, uint32_t nq, int i, int j, int k, double* u) {
    int l = 0, fl = 0;
    u[0] = u[1] = u[2] = 0.0;
    if (!resolve(d, i, j, k, l, fl)) return;
    const int64_t y = ((int64_t)i * d.ny + j) * d.nzp + k;
    for (int a = 0; a < 3; ++a) {
        double v = vel[(int64_t)a * nq + y];
        u[a] = (fl >> a & 1) ? -v : v;
    }
}

// Solver._comp_der_vel, isothermal.pyx:958-1023
LBD void comp_der_vel(const Dims& d, const double* vel, uint32_t nq, uint32_t m, int i, int j, int k,
                             int fdir, const double* u0, double* der) {
    constexpr int FD[3][5] = {{14, 15, 16, 17, 18}, {4, 11, 12, 13, 18}, {3, 7, 10, 13, 17}};
    der[0] = der[1] = der[2] = 0.0;
    for (int fl = 0; fl < 5; ++fl) {
        const int l = FD[fdir][fl];
        double un[3], uo[3];
        nb_vel(d, vel, nq, i + CI[l], j + CJ[l], k + CK[l], un);
        nb_vel(d, vel, nq, i - CI[l], j - CJ[l], k - CK[l], uo);
        const bool sn = m & (1u << opp(l));  // x + c_l is SOLID
        const bool so = m & (1u << l);       // x - c_l is SOLID
        double c1 = 0.5, c2 = 0.0, c3 = -0.5;
        if (sn) {
            if (so) {
                c1 = 1.0;
                c3 = -1.0;
            } else {
                c1 = 4.0 / 3.0;
                c2 = -1.0;
                c3 = -1.0 / 3.0;
            }
        } else if (so) {
            c1 = 1.0 / 3.0;
            c2 = 1.0;
            c3 = -4.0 / 3.0;
        }
        const double w = nnz(l) == 1 ? 2.0 / 36.0 : 1.0 / 36.0;
        const double c0 = 2.0 * w * 3.0;
        der[0] += c0 * (c1 * un[0] + c2 * u0[0] + c3 * uo[0]);
        der[1] += c0 * (c1 * un[1] + c2 * u0[1] + c3 * uo[1]);
        der[2] += c0 * (c1 * un[2] + c2 * u0[2] + c3 * uo[2]);
    }
}

// Solver._init_f, isothermal.pyx:1028-1105: f1 = P(heq + (1-omega)(-2 tau/3) rho S),
// written as the state after an even AA step (f*_l in slot opp(l)).
template <int P>
__global__ void k_init(StepParams p) {
    using TS = typename Var<P>::TS;
    constexpr bool SH = Var<P>::SH;
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.nq) return;
    const uint32_t m = p.mask[c];
    if (!(m & M_FLUID)) return;
    int i, j, k;
    decompose(c, p.d.ny, p.d.nzp, i, j, k);
    const int64_t nq = p.nq;
    const double inv_tau = p.prm[0], tau = 1.0 / inv_tau, rlx_cff = 1.0 - inv_tau;
    const double den = p.den[c];
    double u[3] = {p.vel[c], p.vel[nq + c], p.vel[2 * nq + c]};
    double h[10], dx[3], dy[3], dz[3], f[Q];
    heq<double>(den, u, h);
    comp_der_vel(p.d, p.vel, p.nq, m, i, j, k, 0, u, dx);
    comp_der_vel(p.d, p.vel, p.nq, m, i, j, k, 1, u, dy);
    comp_der_vel(p.d, p.vel, p.nq, m, i, j, k, 2, u, dz);
    const double m2 = -2.0 * tau * (1.0 / 3.0) * den;
    h[4] += rlx_cff * m2 * dx[0];
    h[5] += rlx_cff * m2 * 0.5 * (dx[1] + dy[0]);
    h[6] += rlx_cff * m2 * 0.5 * (dx[2] + dz[0]);
    h[7] += rlx_cff * m2 * dy[1];
    h[8] += rlx_cff * m2 * 0.5 * (dy[2] + dz[1]);
    h[9] += rlx_cff * m2 * dz[2];
    if (SH) h[0] = den - p.rho0;
    project<double, false>(h, f);
    TS* F = (TS*)p.f;
    const int64_t e = eidx(p, c);
    for (int l = 0; l < Q; ++l) stf<P, false>(F + e + (int64_t)opp(l) * p.ps, (typename Var<P>::TC)f[l], p.scale);
}

// Post-collision populations of the newest step, unpacked into the reference's
// f1 layout restricted to the interior: out[l][i][j][k] (f64).
template <int P>
__global__ void k_post_f(StepParams p, int next_odd, double* out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= p.nq) return;
    int i, j, k;
    decompose(c, p.d.ny, p.d.nzp, i, j, k);
    if (k >= p.d.nz) return;
    const int64_t nc = (int64_t)p.d.nx * p.d.ny * p.d.nz;
    const int64_t o = ((int64_t)i * p.d.ny + j) * p.d.nz + k;
    const uint32_t m = p.mask[c];
    const typename Var<P>::TS* F = (const typename Var<P>::TS*)p.f;
    for (int l = 0; l < Q; ++l) {
        double v = 0.0;
        if (m & M_FLUID) {
            const int64_t e = eidx(p, c);
            const int lo = opp(l);
            int64_t loc;
            if (next_odd) loc = e + (int64_t)lo * p.ps;
            else if (!p.tab) loc = odd_loc(p, m, e, lo, i, j, k);
            else loc = lo == REST ? e + (int64_t)REST * p.ps : p.tab[(int64_t)link_slot(lo) * p.ps + e];
            v = (double)ldf<P, false>(F + loc, p.inv_scale);
            if (Var<P>::SH) v += (l == REST ? 12.0 / 36.0 : nnz(l) == 1 ? 2.0 / 36.0 : 1.0 / 36.0) * p.rho0;
        }
        out[(int64_t)l * nc + o] = v;
    }
}

// ---------------------------------------------------------------------------
// launchers (declared in lb_common.h)
// ---------------------------------------------------------------------------
template <int P, int RLX>
static void launch_step(const StepParams& p, bool odd, hipStream_t st) {
    const unsigned grid = tile_grid(p);
    if (p.tab) {  // sparse layout: one kernel, with the mask only where needed
        if (odd && p.f2) {
            if (p.moving || p.nzh) k_step_sp_pull<P, RLX, true><<<grid, p.tile, p.lds[1], st>>>(p);
            else k_step_sp_pull<P, RLX, false><<<grid, p.tile, p.lds[1], st>>>(p);
        } else if (p.moving || p.nzh) {
            if (odd) k_step_sp<P, RLX, true, true><<<grid, p.tile, p.lds[1], st>>>(p);
            else k_step_sp<P, RLX, false, true><<<grid, p.tile, p.lds[0], st>>>(p);
        } else {
            if (odd) k_step_sp<P, RLX, true, false><<<grid, p.tile, p.lds[1], st>>>(p);
            else k_step_sp<P, RLX, false, false><<<grid, p.tile, p.lds[0], st>>>(p);
        }
        return;
    }
    // a pull step (odd, into p.f2): the bulk and the special sites both read
    // p.f and write p.f2, so the two launches touch no common location
    const bool pull = odd && p.f2;
    if constexpr (P == LB_FP16_STORAGE && RLX == TRT) {
        if (pull) k_step_h<RLX, true, true><<<grid, p.tile, p.lds[1], st>>>(p);
        else if (odd) k_step_h<RLX, true><<<grid, p.tile, p.lds[1], st>>>(p);
        else k_step_h<RLX, false><<<grid, p.tile, p.lds[0], st>>>(p);
    } else {
        if (pull) k_step<P, RLX, true, true><<<grid, p.tile, p.lds[1], st>>>(p);
        else if (odd) k_step<P, RLX, true><<<grid, p.tile, p.lds[1], st>>>(p);
        else k_step<P, RLX, false><<<grid, p.tile, p.lds[0], st>>>(p);
    }
    const uint32_t ns = odd ? p.nspecial : p.nzh;
    if (ns) {
        const unsigned g2 = (ns + TILE - 1) / TILE;
        if (pull) k_step_special<P, RLX, true, true><<<g2, TILE, 0, st>>>(p);
        else if (odd) k_step_special<P, RLX, true><<<g2, TILE, 0, st>>>(p);
        else k_step_special<P, RLX, false><<<g2, TILE, 0, st>>>(p);
    }
}

template <int P>
void launch_step_rlx(const StepParams& p, int rlx, bool odd, hipStream_t st) {
    if (rlx == BGK) launch_step<P, BGK>(p, odd, st);
    else if (rlx == TRT) launch_step<P, TRT>(p, odd, st);
    else launch_step<P, REGUL>(p, odd, st);
}

template <int P, int RLX>
static void launch_adj(const StepParams& ps, const AdjParams& x, bool odd, hipStream_t st) {
    StepParams p = ps;  // the adjoint kernels run TILE threads per workgroup
    p.tile = TILE;
    p.ntiles = ((p.tab ? (uint32_t)p.ps : p.nq) + TILE - 1) / TILE;
    const unsigned grid = tile_grid(p);
    if (p.tab) {  // sparse layout
        if (x.lam2) {  // pull form, the same kernel at either parity
            if (p.nzh) k_adj_sp_pull<P, RLX, true><<<grid, TILE, 0, st>>>(p, x);
            else k_adj_sp_pull<P, RLX, false><<<grid, TILE, 0, st>>>(p, x);
        } else if (p.nzh) {
            if (odd) k_adj_sp<P, RLX, true, true><<<grid, TILE, 0, st>>>(p, x);
            else k_adj_sp<P, RLX, false, true><<<grid, TILE, 0, st>>>(p, x);
        } else {
            if (odd) k_adj_sp<P, RLX, true, false><<<grid, TILE, 0, st>>>(p, x);
            else k_adj_sp<P, RLX, false, false><<<grid, TILE, 0, st>>>(p, x);
        }
        return;
    }
    if (x.lam2) {  // pull form: the odd step's site classes at either parity
        k_adj<P, RLX, true, true><<<grid, TILE, 0, st>>>(p, x);
        if (p.nspecial) k_adj_special<P, RLX, true, true><<<(p.nspecial + TILE - 1) / TILE, TILE, 0, st>>>(p, x);
        return;
    }
    if (odd) k_adj<P, RLX, true><<<grid, TILE, 0, st>>>(p, x);
    else k_adj<P, RLX, false><<<grid, TILE, 0, st>>>(p, x);
    const uint32_t ns = odd ? p.nspecial : p.nzh;
    if (ns) {
        const unsigned g2 = (ns + TILE - 1) / TILE;
        if (odd) k_adj_special<P, RLX, true><<<g2, TILE, 0, st>>>(p, x);
        else k_adj_special<P, RLX, false><<<g2, TILE, 0, st>>>(p, x);
    }
}

template <int P>
void launch_adj_rlx(const StepParams& p, const AdjParams& x, int rlx, bool odd, hipStream_t st) {
    if (rlx == BGK) launch_adj<P, BGK>(p, x, odd, st);
    else if (rlx == TRT) launch_adj<P, TRT>(p, x, odd, st);
    else launch_adj<P, REGUL>(p, x, odd, st);
}

// the dissipation objective's gather before the step of parity `odd`
template <int P, int RLX>
static void launch_diss(const StepParams& p, bool odd, hipStream_t st) {
    const unsigned grid = tile_grid(p);
    if (p.tab) {  // sparse layout
        if (p.moving || p.nzh) {
            if (odd) k_diss_sp<P, RLX, true, true><<<grid, p.tile, 0, st>>>(p);
            else k_diss_sp<P, RLX, false, true><<<grid, p.tile, 0, st>>>(p);
        } else {
            if (odd) k_diss_sp<P, RLX, true, false><<<grid, p.tile, 0, st>>>(p);
            else k_diss_sp<P, RLX, false, false><<<grid, p.tile, 0, st>>>(p);
        }
        return;
    }
    if (odd) k_diss<P, RLX, true><<<grid, p.tile, 0, st>>>(p);
    else k_diss<P, RLX, false><<<grid, p.tile, 0, st>>>(p);
    const uint32_t ns = odd ? p.nspecial : p.nzh;
    if (ns) {
        const unsigned g2 = (ns + TILE - 1) / TILE;
        if (odd) k_diss_special<P, RLX, true><<<g2, TILE, 0, st>>>(p);
        else k_diss_special<P, RLX, false><<<g2, TILE, 0, st>>>(p);
    }
}

template <int P>
void launch_diss_rlx(const StepParams& p, int rlx, bool odd, hipStream_t st) {
    if (rlx == BGK) launch_diss<P, BGK>(p, odd, st);
    else if (rlx == TRT) launch_diss<P, TRT>(p, odd, st);
    else launch_diss<P, REGUL>(p, odd, st);
}

template <int P>
void launch_init(const StepParams& p, hipStream_t st) {
    k_init<P><<<(p.nq + 255) / 256, 256, 0, st>>>(p);
}

template <int P>
void launch_post_f(const StepParams& p, int next_odd, double* out, hipStream_t st) {
    k_post_f<P><<<(p.nq + 255) / 256, 256, 0, st>>>(p, next_odd, out);
}

}  // namespace lb
