// MI355X (gfx950) D3Q19 evolve engine: fused AA-pattern stream+collide
// kernel, its set-up kernels and the C ABI declared in include/jyulb_gpu.h.
//
// Storage. One population array (AA pattern, in place), structure of arrays:
// population l of interior site (i,j,k) lives at f[l*nq + (i*ny + j)*nzp + k]
// with the unit-stride axis k (the reference's fastest C-order axis,
// isothermal.pyx:1532) padded to nzp = round_up(nz, 64) so every row starts on
// a 128-byte line for all three storage types. No halo layers are stored:
// the reference's per-step halo passes for periodic and symmetric faces
// (_enforce_distr_per_sym_bc, isothermal.pyx:691-837) are folded into
// address resolution, and WALL / FIXED faces are SOLID halos, i.e. bounce-back
// links recorded in a per-site link mask.
//
// AA pattern. After an "even" step site x holds its post-collision f*_l in
// its own slot opp(l). An odd step pulls f_l(x) = f*_{l'}(y') from slot
// opp(l') of the resolved upstream site (y',l') = resolve(x - c_l, l), or from
// its own slot l for a bounce-back link, collides, and writes f*_{opp(l)}(x)
// back into the very location it read f_l from. An even step reads its own
// slots l and writes f*_l into its own slot opp(l). Every location is read
// and written by one thread only, so one array suffices (the C++ exe of the
// reference uses the same scheme with a tabulated index map,
// include/kernel/kernel.cpp:94-139, 469-552; here addresses are computed).
//
// This translation unit holds the host side (the C ABI, solver state, launch
// tuning) and the set-up, import/export and reduction kernels; the fused step,
// adjoint and initialisation kernels are built per storage variant from
// lb_kernels.h (lb_inst.hip).
#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <math.h>
#include <stdarg.h>
#include <algorithm>
#include <utility>
#include <vector>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <mutex>
#include <string>

#include "lb_common.h"

using namespace lb;

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess)                                                            \
            return fail(LB_EHIP, "%s failed: %s (%s:%d)", #call, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                             \
    } while (0)

// objective: sum over fluid sites of the face layer of rho u*_axis (FP64)
__global__ __launch_bounds__(RED_BLOCK_ADJ) void k_obj(const uint32_t* mask, const double* den, const double* vel,
                                                   Dims d, uint32_t nq, int face, int axis, double* part) {
    double a = 0.0, n = 0.0;
    const int ax = face >> 1;
    for (uint32_t c = blockIdx.x * RED_BLOCK_ADJ + threadIdx.x; c < nq; c += gridDim.x * RED_BLOCK_ADJ) {
        if (!(mask[c] & M_FLUID)) continue;
        int i, j, k;
        decompose(c, d.ny, d.nzp, i, j, k);
        const int pos = ax == 0 ? i : ax == 1 ? j : k;
        const int nn = ax == 0 ? d.nx : ax == 1 ? d.ny : d.nz;
        if (pos != ((face & 1) ? nn - 1 : 0)) continue;
        a += den[c] * vel[(int64_t)axis * nq + c];
        n += 1.0;
    }
    for (int o = 32; o > 0; o >>= 1) {
        a += __shfl_xor(a, o, 64);
        n += __shfl_xor(n, o, 64);
    }
    __shared__ double sh[RED_BLOCK_ADJ / 64][2];
    if ((threadIdx.x & 63) == 0) {
        sh[threadIdx.x >> 6][0] = a;
        sh[threadIdx.x >> 6][1] = n;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double s0 = 0.0, s1 = 0.0;
        for (int w = 0; w < RED_BLOCK_ADJ / 64; ++w) {
            s0 += sh[w][0];
            s1 += sh[w][1];
        }
        part[2 * blockIdx.x] = s0;
        part[2 * blockIdx.x + 1] = s1;
    }
}

template <class TC>
__global__ void k_export_vec3(const TC* src, double* out, Dims d, uint32_t nq) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq) return;
    int i, j, k;
    decompose(c, d.ny, d.nzp, i, j, k);
    if (k >= d.nz) return;
    const int64_t o = ((int64_t)i * d.ny + j) * d.nz + k;
    for (int a = 0; a < 3; ++a) out[3 * o + a] = (double)src[(int64_t)a * nq + c];
}

// The same from rows indexed by compact site (nf sites, rows of nfp values;
// site[n] is the cell of compact site n) into an output zeroed beforehand.
template <class TC>
__global__ void k_export_vec3_sp(const TC* src, const uint32_t* site, uint32_t nf, uint32_t nfp, double* out, Dims d) {
    const uint32_t n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= nf) return;
    int i, j, k;
    decompose(site[n], d.ny, d.nzp, i, j, k);
    const int64_t o = ((int64_t)i * d.ny + j) * d.nz + k;
    for (int a = 0; a < 3; ++a) out[3 * o + a] = (double)src[(int64_t)a * nfp + n];
}

// Device-to-device copy, one 8-byte element per thread over a grid that
// covers the array once: on MI355X 5.8 TB/s (read + write) on 2 GiB of
// nonzero data, where hipMemcpy's device copy reaches 4.9 TB/s and grid-
// stride loops of 8- or 16-byte elements 4.4-5.7 TB/s (all of them 0.3-0.5
// TB/s faster on zero-filled buffers).
__global__ __launch_bounds__(256) void k_stream_copy(const uint64_t* __restrict__ src, uint64_t* __restrict__ dst,
                                                     size_t n) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

static hipError_t stream_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
    if (((uintptr_t)dst | (uintptr_t)src | bytes) & 7)
        return hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, st);
    const size_t n = bytes / 8, chunk = (size_t)1 << 38;  // 2^30 workgroups per launch at most
    for (size_t o = 0; o < n; o += chunk) {
        const size_t m = std::min(chunk, n - o);
        k_stream_copy<<<(unsigned)((m + 255) / 256), 256, 0, st>>>((const uint64_t*)src + o, (uint64_t*)dst + o, m);
    }
    return hipGetLastError();
}

template <class T>
__global__ void k_fill(T* p, size_t n, T v) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}


// ---------------------------------------------------------------------------
// set-up kernels
// ---------------------------------------------------------------------------

// Link mask of every site (see lb_lattice.h) and the class of every wave of
// 64 sites: 0 when all its sites are fluid with a bare M_FLUID mask (no
// bounce-back, halo-crossing, Zou-He or moving-wall link), 1 otherwise.
// `vel` (wall velocities) may be null, then no site is flagged MOVING.
__global__ void k_mask(const uint8_t* phase, const double* vel, uint32_t* mask, uint8_t* wclass, Dims d,
                       uint32_t nq) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq) return;  // whole waves only: nq % 64 == 0
    int i, j, k;
    decompose(c, d.ny, d.nzp, i, j, k);
    uint32_t m = 0;
    if (k < d.nz && phase[c] != SOLID) {
        m = M_FLUID;
        for (int l = 0; l < Q; ++l) {
            if (l == REST) continue;
            int yi = i - CI[l], yj = j - CJ[l], yk = k - CK[l], lp = l, fl = 0;
            const bool crossed_ij = yi < 0 || yi >= d.nx || yj < 0 || yj >= d.ny;
            const bool crossed_k = yk < 0 || yk >= d.nz;
            if (!resolve(d, yi, yj, yk, lp, fl)) {
                m |= 1u << l;
                continue;
            }
            const int64_t y = ((int64_t)yi * d.ny + yj) * d.nzp + yk;
            if (phase[y] == SOLID) {
                m |= 1u << l;
                if (vel && (vel[y] != 0.0 || vel[nq + y] != 0.0 || vel[2 * (int64_t)nq + y] != 0.0))
                    m |= M_MOVING;
            } else if (crossed_ij || (crossed_k && d.bc[4] != PERIODIC)) {
                m |= M_BND;
            } else if (crossed_k) {
                m |= M_KWRAP;
            }
        }
        for (int fc = 0; fc < 6; ++fc) {
            const bool fx = d.bc[fc] == FIXED_DEN || d.bc[fc] == FIXED_VEL;
            const int a = fc >> 1, x = a == 0 ? i : a == 1 ? j : k, n = a == 0 ? d.nx : a == 1 ? d.ny : d.nz;
            if (fx && x == ((fc & 1) ? n - 1 : 0)) m |= M_ZH0 << fc;
        }
    }
    if (m & M_BND) m &= ~M_KWRAP;  // resolved in full by the special-site kernel
    mask[c] = m;
    const unsigned long long notplain = __ballot(m != M_FLUID);
    if ((threadIdx.x & 63) == 0) wclass[c >> 6] = notplain ? 1 : 0;
}


// interior contiguous (den[nc], vel[nc][3]) <-> padded SoA device fields
__global__ void k_import_dv(const double* st, double* den, double* vel, Dims d, uint32_t nq, int with_den) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq) return;
    int i, j, k;
    decompose(c, d.ny, d.nzp, i, j, k);
    if (k >= d.nz) return;
    const int64_t nc = (int64_t)d.nx * d.ny * d.nz;
    const int64_t o = ((int64_t)i * d.ny + j) * d.nz + k;
    if (with_den) den[c] = st[o];
    const double* v = st + (with_den ? nc : 0) + 3 * o;
    vel[c] = v[0];
    vel[nq + c] = v[1];
    vel[2 * (int64_t)nq + c] = v[2];
}

__global__ void k_export_dv(double* st, const double* den, const double* vel, Dims d, uint32_t nq) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq) return;
    int i, j, k;
    decompose(c, d.ny, d.nzp, i, j, k);
    if (k >= d.nz) return;
    const int64_t nc = (int64_t)d.nx * d.ny * d.nz;
    const int64_t o = ((int64_t)i * d.ny + j) * d.nz + k;
    st[o] = den[c];
    double* v = st + nc + 3 * o;
    v[0] = vel[c];
    v[1] = vel[nq + c];
    v[2] = vel[2 * (int64_t)nq + c];
}

template <class TC>
__global__ void k_import_acc(const double* st, TC* acc, Dims d, uint32_t nq) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq) return;
    int i, j, k;
    decompose(c, d.ny, d.nzp, i, j, k);
    if (k >= d.nz) {
        acc[c] = acc[nq + c] = acc[2 * (int64_t)nq + c] = TC(0);
        return;
    }
    const int64_t o = ((int64_t)i * d.ny + j) * d.nz + k;
    acc[c] = (TC)st[3 * o];
    acc[nq + c] = (TC)st[3 * o + 1];
    acc[2 * (int64_t)nq + c] = (TC)st[3 * o + 2];
}

__global__ void k_import_phase(const uint8_t* st, uint8_t* phase, Dims d, uint32_t nq) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq) return;
    int i, j, k;
    decompose(c, d.ny, d.nzp, i, j, k);
    phase[c] = k < d.nz ? st[((int64_t)i * d.ny + j) * d.nz + k] : (uint8_t)SOLID;
}

// ---------------------------------------------------------------------------
// FP64 reductions: wave shuffles -> LDS -> per-block partials -> one block
// ---------------------------------------------------------------------------
constexpr int RED_N = 8;
constexpr int RED_BLOCK = 256;

LBD double wave_sum(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
LBD double wave_max(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

__global__ __launch_bounds__(RED_BLOCK) void k_reduce1(const uint32_t* mask, const double* den, const double* vel,
                                                      uint32_t nq, double* part) {
    double a[RED_N] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t c = blockIdx.x * RED_BLOCK + threadIdx.x; c < nq; c += gridDim.x * RED_BLOCK) {
        if (!(mask[c] & M_FLUID)) continue;
        const double r = den[c], u = vel[c], v = vel[nq + c], w = vel[2 * (int64_t)nq + c];
        a[0] += r;
        a[1] += u;
        a[2] += v;
        a[3] += w;
        a[4] += r * u;
        a[5] += r * v;
        a[6] += r * w;
        a[7] = fmax(a[7], sqrt(u * u + v * v + w * w));
    }
    __shared__ double sh[RED_BLOCK / 64][RED_N];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < RED_N; ++q) {
        double s = q == 7 ? wave_max(a[q]) : wave_sum(a[q]);
        if (lane == 0) sh[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < RED_N) {
        const int q = threadIdx.x;
        double s = sh[0][q];
        for (int w2 = 1; w2 < RED_BLOCK / 64; ++w2) s = q == 7 ? fmax(s, sh[w2][q]) : s + sh[w2][q];
        part[(int64_t)blockIdx.x * RED_N + q] = s;
    }
}

__global__ __launch_bounds__(RED_BLOCK) void k_reduce2(const double* part, int nparts, double* out) {
    __shared__ double sh[RED_BLOCK / 64][RED_N];
    double a[RED_N] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < nparts; b += RED_BLOCK)
        for (int q = 0; q < RED_N; ++q) a[q] = q == 7 ? fmax(a[q], part[(int64_t)b * RED_N + q]) : a[q] + part[(int64_t)b * RED_N + q];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < RED_N; ++q) {
        double s = q == 7 ? wave_max(a[q]) : wave_sum(a[q]);
        if (lane == 0) sh[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < RED_N) {
        const int q = threadIdx.x;
        double s = sh[0][q];
        for (int w2 = 1; w2 < RED_BLOCK / 64; ++w2) s = q == 7 ? fmax(s, sh[w2][q]) : s + sh[w2][q];
        out[q] = s;
    }
}

// sum |u - u_prev|^2 and sum |u|^2 over fluid sites, then u_prev := u
__global__ __launch_bounds__(RED_BLOCK) void k_conv1(const uint32_t* mask, const double* vel, double* uprev,
                                                    uint32_t nq, double* part) {
    double a[RED_N] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (uint32_t c = blockIdx.x * RED_BLOCK + threadIdx.x; c < nq; c += gridDim.x * RED_BLOCK) {
        if (!(mask[c] & M_FLUID)) continue;
        for (int q = 0; q < 3; ++q) {
            const double u = vel[(int64_t)q * nq + c], p = uprev[(int64_t)q * nq + c];
            a[0] += (u - p) * (u - p);
            a[1] += u * u;
            uprev[(int64_t)q * nq + c] = u;
        }
    }
    __shared__ double sh[RED_BLOCK / 64][RED_N];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    for (int q = 0; q < RED_N; ++q) {
        double s = wave_sum(a[q]);
        if (lane == 0) sh[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < RED_N) {
        const int q = threadIdx.x;
        double s = sh[0][q];
        for (int w2 = 1; w2 < RED_BLOCK / 64; ++w2) s += sh[w2][q];
        part[(int64_t)blockIdx.x * RED_N + q] = s;
    }
}

// ---------------------------------------------------------------------------
// calc_flow_info (include/solver/solver.cpp:64-129): FP64 sums, extrema and
// the fluid-site count in one pass. Slots FI_DEN..FI_CNT are sums, FI_MAXDEN
// and FI_MAXU maxima, FI_MINDEN a minimum; the extrema of the density start
// from the reference density and the speed maximum from 0, as in the
// reference. The sites come from an accessor: a solver's padded SoA fields
// (SolverSites) or caller-provided interior C-order fields (PlainSites).
// ---------------------------------------------------------------------------
enum {
    FI_DEN, FI_UX, FI_UY, FI_UZ, FI_U2, FI_JX, FI_JY, FI_JZ, FI_J2, FI_CNT,  // sums
    FI_MAXDEN, FI_MINDEN, FI_MAXU, FLOW_N
};
constexpr int FLOW_SUMS = FI_CNT + 1;

LBD double flow_op(int q, double a, double b) {
    return q < FLOW_SUMS ? a + b : q == FI_MINDEN ? fmin(a, b) : fmax(a, b);
}

struct SolverSites {
    const uint32_t* mask;
    const double* den;
    const double* vel;  // 3 * nq, SoA
    uint32_t n;
    LBD bool get(uint32_t c, double& r, double& u, double& v, double& w) const {
        if (!(mask[c] & M_FLUID)) return false;
        r = den[c];
        u = vel[c];
        v = vel[n + c];
        w = vel[2 * (int64_t)n + c];
        return true;
    }
};

struct PlainSites {
    const uint8_t* phase;  // [n], SOLID = 0
    const double* den;     // [n]
    const double* vel;     // [n][3]
    uint32_t n;
    LBD bool get(uint32_t c, double& r, double& u, double& v, double& w) const {
        if (phase[c] == SOLID) return false;
        r = den[c];
        u = vel[3 * (int64_t)c];
        v = vel[3 * (int64_t)c + 1];
        w = vel[3 * (int64_t)c + 2];
        return true;
    }
};

template <class S>
__global__ __launch_bounds__(RED_BLOCK) void k_flow1(S sites, double rho0, double* part) {
    double a[FLOW_N];
#pragma unroll
    for (int q = 0; q < FLOW_N; ++q) a[q] = q < FLOW_SUMS ? 0.0 : q == FI_MAXU ? 0.0 : rho0;
    for (uint32_t c = blockIdx.x * RED_BLOCK + threadIdx.x; c < sites.n; c += gridDim.x * RED_BLOCK) {
        double r, u, v, w;
        if (!sites.get(c, r, u, v, w)) continue;
        const double u2 = sqrt(u * u + v * v + w * w);
        a[FI_DEN] += r;
        a[FI_UX] += u;
        a[FI_UY] += v;
        a[FI_UZ] += w;
        a[FI_U2] += u2;
        a[FI_JX] += r * u;
        a[FI_JY] += r * v;
        a[FI_JZ] += r * w;
        a[FI_J2] += r * u2;
        a[FI_CNT] += 1.0;
        a[FI_MAXDEN] = fmax(a[FI_MAXDEN], r);
        a[FI_MINDEN] = fmin(a[FI_MINDEN], r);
        a[FI_MAXU] = fmax(a[FI_MAXU], u2);
    }
    __shared__ double sh[RED_BLOCK / 64][FLOW_N];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < FLOW_N; ++q) {
        double s = a[q];
        for (int o = 32; o > 0; o >>= 1) s = flow_op(q, s, __shfl_xor(s, o, 64));
        if (lane == 0) sh[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < FLOW_N) {
        const int q = threadIdx.x;
        double s = sh[0][q];
        for (int w2 = 1; w2 < RED_BLOCK / 64; ++w2) s = flow_op(q, s, sh[w2][q]);
        part[(int64_t)blockIdx.x * FLOW_N + q] = s;
    }
}

// Final reduction of the per-block partials, written as lb_flow_info_t
// (IsothermalFlowInfo, solver.cpp:112-128, plus the fluid-site count).
__global__ __launch_bounds__(RED_BLOCK) void k_flow2(const double* part, int nparts, double rho0, double* out) {
    double a[FLOW_N];
#pragma unroll
    for (int q = 0; q < FLOW_N; ++q) a[q] = q < FLOW_SUMS ? 0.0 : q == FI_MAXU ? 0.0 : rho0;
    for (int b = threadIdx.x; b < nparts; b += RED_BLOCK)
#pragma unroll
        for (int q = 0; q < FLOW_N; ++q) a[q] = flow_op(q, a[q], part[(int64_t)b * FLOW_N + q]);
    __shared__ double sh[RED_BLOCK / 64][FLOW_N];
    __shared__ double r[FLOW_N];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < FLOW_N; ++q) {
        double s = a[q];
        for (int o = 32; o > 0; o >>= 1) s = flow_op(q, s, __shfl_xor(s, o, 64));
        if (lane == 0) sh[wv][q] = s;
    }
    __syncthreads();
    if (threadIdx.x < FLOW_N) {
        const int q = threadIdx.x;
        double s = sh[0][q];
        for (int w2 = 1; w2 < RED_BLOCK / 64; ++w2) s = flow_op(q, s, sh[w2][q]);
        r[q] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const double n = r[FI_CNT];
        const double ave_cff = n > 0 ? 1.0 / n : 0.0;
        const double v[FLOW_N] = {r[FI_JX], r[FI_JY], r[FI_JZ], r[FI_J2], r[FI_DEN], ave_cff * r[FI_UX],
                                  ave_cff * r[FI_UY], ave_cff * r[FI_UZ], ave_cff * r[FI_U2], r[FI_MAXU],
                                  r[FI_MAXDEN], r[FI_MINDEN], n};
        for (int q = 0; q < FLOW_N; ++q) out[q] = v[q];
    }
}
static_assert(sizeof(lb_flow_info_t) == FLOW_N * sizeof(double), "lb_flow_info_t is FLOW_N doubles");

// ---------------------------------------------------------------------------
// scratch of the allocation-free entry point lb_flow_info_fields_noalloc: the
// per-block partials of k_flow1 (RED_PARTS x FLOW_N doubles, 104 KiB per
// device). In host code the name is only a stand-in for the symbol; its
// address on each device comes from hipGetSymbolAddress (scratch_ptr).
// ---------------------------------------------------------------------------
static __device__ __align__(256) double scratch[1024 * FLOW_N];

// ---------------------------------------------------------------------------
// the Solver's halo-padded host fields (lb_bind_host_fields / lb_evolve_host),
// staged on the device in the host layout: phase u8 [NI][NJ][NK], den f64
// [NI][NJ][NK], vel and acc f64 [NI][NJ][NK][3] (isothermal.pyx:243-248)
// ---------------------------------------------------------------------------
LBD int64_t host_idx(const Dims& d, int i, int j, int k) {
    return ((int64_t)(i + 1) * (d.ny + 2) + (j + 1)) * (d.nz + 2) + (k + 1);
}

constexpr unsigned SYNC_PHASE = 1u, SYNC_WALLS = 2u, SYNC_ACC = 4u;

// Inputs of an evolve call from the staged host fields: den/vel of every
// interior site (the wall velocities and the densities that Ladd's term reads
// in the first step, isothermal.pyx:1117-1119), the acceleration field in the
// compute type (acc may be null: only the flag is raised). Flags, OR-ed per
// wave into *flags: SYNC_PHASE if a site's fluid/solid state differs from the
// finalized phase, SYNC_WALLS if a SOLID site's velocity changed, SYNC_ACC if
// the acceleration field is nonzero anywhere.
template <class TC>
__global__ void k_sync_in(const uint8_t* hphase, const double* hden, const double* hvel, const double* hacc,
                          const uint8_t* phase, double* den, double* vel, TC* acc, Dims d, uint32_t nq,
                          unsigned int* flags) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned fl = 0;
    if (c < nq) {
        int i, j, k;
        decompose(c, d.ny, d.nzp, i, j, k);
        if (k < d.nz) {
            const int64_t o = host_idx(d, i, j, k);
            const bool fluid = phase[c] != SOLID;
            if ((hphase[o] != SOLID) != fluid) fl |= SYNC_PHASE;
            const double v0 = hvel[3 * o], v1 = hvel[3 * o + 1], v2 = hvel[3 * o + 2];
            if (!fluid && (v0 != vel[c] || v1 != vel[nq + c] || v2 != vel[2 * (int64_t)nq + c])) fl |= SYNC_WALLS;
            den[c] = hden[o];
            vel[c] = v0;
            vel[nq + c] = v1;
            vel[2 * (int64_t)nq + c] = v2;
            const double a0 = hacc[3 * o], a1 = hacc[3 * o + 1], a2 = hacc[3 * o + 2];
            if (a0 != 0.0 || a1 != 0.0 || a2 != 0.0) fl |= SYNC_ACC;
            if (acc) {
                acc[c] = (TC)a0;
                acc[nq + c] = (TC)a1;
                acc[2 * (int64_t)nq + c] = (TC)a2;
            }
        } else if (acc) {
            acc[c] = acc[nq + c] = acc[2 * (int64_t)nq + c] = TC(0);
        }
    }
    const unsigned w = (__ballot((fl & SYNC_PHASE) != 0) ? SYNC_PHASE : 0u) |
                       (__ballot((fl & SYNC_WALLS) != 0) ? SYNC_WALLS : 0u) |
                       (__ballot((fl & SYNC_ACC) != 0) ? SYNC_ACC : 0u);
    if ((threadIdx.x & 63) == 0 && w) atomicOr(flags, w);
}

// den/vel of the fluid sites into the staged host fields (the SOLID sites
// keep the host's values: the reference never writes them)
__global__ void k_scatter_fluid(const uint32_t* mask, const double* den, const double* vel, double* hden,
                                double* hvel, Dims d, uint32_t nq) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq || !(mask[c] & M_FLUID)) return;
    int i, j, k;
    decompose(c, d.ny, d.nzp, i, j, k);
    const int64_t o = host_idx(d, i, j, k);
    hden[o] = den[c];
    hvel[3 * o] = vel[c];
    hvel[3 * o + 1] = vel[nq + c];
    hvel[3 * o + 2] = vel[2 * (int64_t)nq + c];
}

// _enforce_den_vel_per_sym_bc (isothermal.pyx:582-686) on the staged host
// fields, one axis per launch in the reference's order x, y, z: the x pass
// covers interior j,k, the y pass all i and interior k, the z pass all i,j.
// A periodic face copies the opposite interior layer into its halo, a
// symmetric face mirrors the adjacent layer with the normal velocity negated.
__global__ void k_halo_dv(double* hden, double* hvel, Dims d, int axis) {
    const int NI = d.nx + 2, NJ = d.ny + 2, NK = d.nz + 2;
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int a, b;  // the two coordinates across the face
    int n1, n2;
    if (axis == 0) { n1 = NJ - 2; n2 = NK - 2; }
    else if (axis == 1) { n1 = NI; n2 = NK - 2; }
    else { n1 = NI; n2 = NJ; }
    if (t >= (int64_t)n1 * n2) return;
    a = (int)(t / n2);
    b = (int)(t - (int64_t)a * n2);
    const int N = axis == 0 ? NI : axis == 1 ? NJ : NK;
    auto at = [&](int pos) -> int64_t {
        if (axis == 0) return ((int64_t)pos * NJ + (a + 1)) * NK + (b + 1);
        if (axis == 1) return ((int64_t)a * NJ + pos) * NK + (b + 1);
        return ((int64_t)a * NJ + b) * NK + pos;
    };
    auto copy = [&](int dst, int src, bool neg) {
        const int64_t o = at(dst), s = at(src);
        hden[o] = hden[s];
        for (int q = 0; q < 3; ++q) {
            const double v = hvel[3 * s + q];
            hvel[3 * o + q] = (neg && q == axis) ? -v : v;
        }
    };
    const int lo = d.bc[2 * axis], hi = d.bc[2 * axis + 1];
    if (lo == PERIODIC) copy(N - 1, 1, false);
    else if (lo == SYMMETRIC) copy(0, 1, true);
    if (hi == PERIODIC) copy(0, N - 2, false);
    else if (hi == SYMMETRIC) copy(N - 1, N - 2, true);
}

// ---------------------------------------------------------------------------
// FP16 storage guard: the largest stored magnitude (|bits| & 0x7fff, so that
// inf and NaN, exponent all ones, compare above every finite value) and an
// exact rescale by a power of two
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_f16_amax(const uint4* f, size_t n8, unsigned int* out) {
    uint32_t m = 0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n8; i += (size_t)gridDim.x * 256) {
        const uint4 v = f[i];
        const uint32_t w[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int q = 0; q < 4; ++q) m = max(m, max(w[q] & 0x7fffu, (w[q] >> 16) & 0x7fffu));
    }
    for (int o = 32; o > 0; o >>= 1) m = max(m, (uint32_t)__shfl_xor((int)m, o, 64));
    if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

__global__ __launch_bounds__(256) void k_f16_scale(_Float16* f, size_t n, float factor) {
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        f[i] = (_Float16)((float)f[i] * factor);
}

// bound of |f - w rho0| implied by the Zou-He data of one face:
// |den - rho0| + 3 |u|_1 + 4.5 |u|_1^2 (FIXED_DEN faces use den, FIXED_VEL
// faces the velocity; both are bounded)
__global__ void k_face_bound(const double* fden, const double* fvel, size_t n, double rho0, double* part) {
    double a = 0.0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const double u = fabs(fvel[3 * i]) + fabs(fvel[3 * i + 1]) + fabs(fvel[3 * i + 2]);
        a = fmax(a, fabs(fden[i] - rho0) + 3.0 * u + 4.5 * u * u);
    }
    a = wave_max(a);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
struct lb_solver {
    lb_config cfg;
    Dims d;
    int NI, NJ, NK;
    uint32_t nq;
    int64_t ncell;
    int64_t rs, ps, fsize;  // population-array strides and element count
    int prec;
    double prm[2];
    double rho0;
    double scale;
    hipStream_t stream = nullptr;
    uint64_t xcd_cells = LB_XCD_REMAP_CELLS;  // tile order threshold (tile_index)
    uint32_t xcd_chunk = 0;                   // tiles per XCD chunk in that order (0: contiguous ranges)
    bool even_rr = false;                     // even steps in round-robin order whatever the above
    uint32_t step_lds[2] = {0, 0};            // see StepParams::lds
    uint32_t step_tile = TILE;                // see StepParams::tile
    bool tune = false;                        // choose the step launch on the next long evolve (tune_step)
    bool tune_xcd = true;                     // ... including the tile order (not pinned by the user)
    bool tune_verbose = false;
    uint32_t* mov_idx = nullptr;  // moving-site index per site (k_ladd_prep)
    double* ladd_u = nullptr;     // 18 Ladd velocity sums per moving site
    size_t ladd_cap = 0;          // moving sites ladd_u has room for
    void* f = nullptr;
    uint32_t* mask = nullptr;
    uint8_t* wclass = nullptr;
    uint32_t* special = nullptr;  // compact list of special sites
    uint32_t nspecial = 0, special_cap = 0;
    uint32_t nzh = 0;  // Zou-He sites at the head of the special list
    // sparse layout: compact index per site, site per compact index, odd-step
    // table; nfp = fluid sites rounded up to 64 (the population stride)
    bool sparse = false;
    int layout = LB_LAYOUT_AUTO;  // resolved at lb_create, or for LB_LAYOUT_AUTO at the first lb_finalize
    uint32_t* cidx = nullptr;
    uint32_t* site = nullptr;
    uint32_t* tab = nullptr;
    uint32_t* tabc = nullptr;  // SP_CODE_BYTES per compact site: SP_WORDS rows of words, a row of bytes
    uint32_t* tabw = nullptr;
    void* f2 = nullptr;   // sparse layout: the second population array of the pull steps (run_steps), or null
    int sp_pull = -1;     // pull steps: 1 on, 0 off, -1 by sp_pull_on / the launch tuner (LB_SP_PULL sets 1 or 0)
    bool pull_active = false;  // run_steps takes pull steps (f2 allocated); a dense f2 may be a tuning candidate
    // the state before the last step of the last evolve call, for the
    // reference's f2 (lb_keep_previous): copy, its parity and FP16 scale
    void* prev = nullptr;
    size_t prev_bytes = 0;
    int prev_odd = 1;
    double prev_scale = 1.0;
    bool keep_prev = false, prev_valid = false;
    uint64_t sp_kinds[4] = {0, 0, 0, 0};  // link codes by kind (lb_sparse_codes)
    uint32_t nfp = 0;
    uint8_t* phase = nullptr;
    double* den = nullptr;
    double* vel = nullptr;
    void* acc = nullptr;
    bool acc_on = false;  // the step kernels read acc (it is nonzero somewhere)
    double* fden[6] = {};
    double* fvel[6] = {};
    double* stage = nullptr;  // 4 * ncell doubles (import / export)
    double* part = nullptr;   // reduction partials (RED_PARTS x FLOW_N)
    double* red = nullptr;    // FLOW_N doubles
    // the Solver's halo-padded host fields (lb_bind_host_fields) and their
    // staged copies on the device, ncnt = NI * NJ * NK sites
    uint8_t* h_phase = nullptr;
    double *h_den = nullptr, *h_vel = nullptr, *h_acc = nullptr;
    uint8_t* st_phase = nullptr;
    double *st_den = nullptr, *st_vel = nullptr, *st_acc = nullptr;
    int64_t ncnt = 0;
    // FP16 storage guard: steps between checks of the stored magnitudes, steps
    // since the last check, and whether the scale was chosen automatically
    uint32_t f16_check = 64, f16_since = 0;
    bool f16_auto = true;
    bool f16_dirty = true;  // inputs changed since the last guard check
    double* uprev = nullptr;
    bool have_uprev = false;
    bool phase_set = false;
    bool finalized = false;
    bool moving = false;
    int64_t nfluid = 0;
    uint64_t steps = 0;
    int next_odd = 1;
    uint64_t bytes = 0;
    // adjoint (checkpointed reverse pass)
    void* ckpt = nullptr;  // (nseg + 1) states: segment starts and the final state
    size_t ckpt_bytes = 0;
    void* lam = nullptr;
    void* lam2 = nullptr;  // the second cotangent array of the pull-form adjoint steps, or null
    int adj_pull = -1;     // pull-form adjoint steps: 1 on, 0 off, -1 from 1 GiB of cotangents (LB_ADJ_PULL sets 1 or 0)
    std::vector<uint8_t> ckpt_odd;  // next_odd of each checkpoint
    void* rec = nullptr;
    size_t rec_bytes = 0;
    void* gacc = nullptr;
    double* gface = nullptr;
    uint32_t adj_steps = 0, adj_interval = 0;
    int adj_po = 1;
    uint64_t adj_step_end = 0;
    int adj_face = FX1, adj_axis = 0;
    double adj_scale = 0.0;
    bool adj_ready = false;
    bool rec_fwd = false;  // rec holds the last segment, recorded by lb_adj_forward
    void* dseed = nullptr;  // dissipation objective: (Q + 4) rows of nq, compute type
    // one snapshot slot (lb_snapshot / lb_restore)
    void* snap = nullptr;
    double* snap_dv = nullptr;
    uint64_t snap_steps = 0;
    int snap_odd = 1;
};

static constexpr int RED_PARTS = 1024;
static_assert(sizeof(scratch) == sizeof(double) * RED_PARTS * FLOW_N, "scratch holds RED_PARTS partials");

static size_t ts_size(int prec) { return prec == LB_FP64 ? 8 : prec == LB_FP16_STORAGE ? 2 : 4; }
static size_t tc_size(int prec) { return prec == LB_FP64 || prec == LB_FP32_STORAGE ? 8 : 4; }

// Values per row of the adjoint's (rho, u*) record and acceleration
// cotangent: one per cell in the dense layouts, one per compact site in the
// sparse layout (whose kernels then read no cell index for them).
static size_t rec_row(const lb_solver* s) { return s->sparse ? (size_t)s->nfp : (size_t)s->nq; }

static int dalloc(lb_solver* s, void** p, size_t n) {
    hipError_t e = hipMalloc(p, n);
    if (e != hipSuccess) {
        *p = nullptr;
        return fail(LB_ENOMEM, "hipMalloc(%zu bytes) failed: %s", n, hipGetErrorString(e));
    }
    s->bytes += n;
    return LB_OK;
}

// An optional array (the second population or cotangent array of the pull
// steps): null, and the steps run in place, if the device cannot hold it.
static void dalloc_optional(lb_solver* s, void** p, size_t n) {
    if (hipMalloc(p, n) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return;
    }
    s->bytes += n;
}

static void dfree(lb_solver* s, void** p, size_t n) {
    if (!*p) return;
    (void)hipFree(*p);
    *p = nullptr;
    s->bytes -= n;
}

static StepParams params(const lb_solver* s) {
    StepParams p;
    memset(&p, 0, sizeof(p));
    p.f = s->f;
    p.mask = s->mask;
    p.den = s->den;
    p.vel = s->vel;
    p.acc = s->acc_on ? s->acc : nullptr;
    for (int fc = 0; fc < 6; ++fc) {
        p.fden[fc] = s->fden[fc];
        p.fvel[fc] = s->fvel[fc];
    }
    for (int a = 0; a < 3; ++a) p.g[a] = s->cfg.g[a];
    p.prm[0] = s->prm[0];
    p.prm[1] = s->prm[1];
    p.rho0 = s->rho0;
    p.scale = (float)s->scale;
    p.inv_scale = (float)(1.0 / s->scale);
    p.d = s->d;
    p.NJ = s->NJ;
    p.NK = s->NK;
    p.nq = s->nq;
    p.tile = s->step_tile;
    p.ntiles = ((s->sparse ? s->nfp : s->nq) + p.tile - 1) / p.tile;
    // the tuned launch may set xcd_cells to 0 or UINT64_MAX
    p.xcd = s->nq >= s->xcd_cells ? (s->xcd_chunk ? s->xcd_chunk : XCD_CONTIGUOUS) : 0u;
    p.lds[0] = s->step_lds[0];
    p.lds[1] = s->step_lds[1];
    p.rs = s->rs;
    p.ps = s->ps;
    p.wclass = s->wclass;
    p.nwaves = s->nq / 64;
    p.special = s->special;
    p.nspecial = s->nspecial;
    p.nzh = s->nzh;

    p.moving = s->moving ? 1 : 0;
    p.mov_idx = s->mov_idx;
    p.ladd_u = s->ladd_u;
    if (s->sparse) {
        p.cidx = s->cidx;
        p.site = s->site;
        p.tab = s->tab;
        p.tabc = s->tabc;
        p.tabb = (const uint8_t*)(s->tabc + (size_t)SP_WORDS * s->nfp);
        p.tabw = s->tabw;
        p.nf = (uint32_t)s->nfluid;
    }
    return p;
}

// An input of the step changed (or the state moved on): the forward pass's
// record no longer matches, and an adjoint pass recorded before refers to
// other inputs, so the reverse pass is refused until the next lb_adj_forward.
static void inputs_changed(lb_solver* s) {
    s->rec_fwd = false;
    s->adj_ready = false;
    s->f16_dirty = true;
}

// The state moves on (evolve steps): the same for the adjoint, while the
// FP16 guard keeps its own step count.
static void state_advanced(lb_solver* s) {
    s->rec_fwd = false;
    s->adj_ready = false;
}

static int face_bound(lb_solver* s, double* out);

// LB_LAYOUT_AUTO picks the sparse layout from this fraction of solid cells on
// (LB_AUTO_SPARSE_SOLID in the environment overrides it)
#ifndef LB_AUTO_SPARSE_SOLID
#define LB_AUTO_SPARSE_SOLID 0.05
#endif

static unsigned blocks_for(uint32_t n, unsigned b) { return (n + b - 1) / b; }

// _check_bc, isothermal.pyx:378-426
static int check_config(const lb_config* c) {
    if (c->lat[0] < 1 || c->lat[1] < 1 || c->lat[2] < 1) return fail(LB_EINVAL, "Invalid lattice size input.");
    // sites are indexed with 32-bit unsigned integers, and the site counts
    // of the launch arithmetic are ints: nx * ny * round_up(nz, 64) < 2^31
    const uint64_t padded = (uint64_t)c->lat[0] * (uint64_t)c->lat[1] * (((uint64_t)c->lat[2] + 63) / 64 * 64);
    if (padded >= (1ull << 31))
        return fail(LB_EINVAL, "lattice of %llu cells (z padded to a multiple of 64) exceeds the engine's limit "
                    "of 2^31 padded cells", (unsigned long long)padded);
    if (c->rlx_scheme < 0 || c->rlx_scheme > 2) return fail(LB_EINVAL, "Invalid relaxation scheme input.");
    if (!(c->kvisc > 0.0)) return fail(LB_EINVAL, "Invalid kinematic viscosity input.");
    for (int f = 0; f < 6; ++f)
        if (c->bc[f] < WALL || c->bc[f] > SYMMETRIC) return fail(LB_EINVAL, "Invalid boundary condition input.");
    const char* ax = "xyz";
    for (int a = 0; a < 3; ++a) {
        const int lo = c->bc[2 * a], hi = c->bc[2 * a + 1];
        if ((lo == PERIODIC || hi == PERIODIC) && lo != hi)
            return fail(LB_EINVAL, "Invalid periodic bc setup (in %c-dir.).", ax[a]);
    }
    bool fx[3];
    for (int a = 0; a < 3; ++a) {
        const int lo = c->bc[2 * a], hi = c->bc[2 * a + 1];
        fx[a] = lo == FIXED_DEN || lo == FIXED_VEL || hi == FIXED_DEN || hi == FIXED_VEL;
    }
    if ((fx[0] && (fx[1] || fx[2])) || (fx[1] && (fx[0] || fx[2])) || (fx[2] && (fx[0] || fx[1])))
        return fail(LB_EINVAL, "Invalid fixed variable bc setup.");
    if (c->precision < LB_FP64 || c->precision > LB_FP32_STORAGE) return fail(LB_EINVAL, "Invalid precision.");
    if (c->layout < LB_LAYOUT_ROWS || c->layout > LB_LAYOUT_AUTO) return fail(LB_EINVAL, "Invalid layout.");
    return LB_OK;
}

static int ensure_stage(lb_solver* s) {
    if (s->stage) return LB_OK;
    return dalloc(s, (void**)&s->stage, (size_t)4 * s->ncell * 8);
}

static int import_acc_from_stage(lb_solver* s) {
    if (!s->acc) {
        int rc = dalloc(s, &s->acc, (size_t)3 * s->nq * tc_size(s->prec));
        if (rc) return rc;
    }
    if (tc_size(s->prec) == 8)
        k_import_acc<double><<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->stage, (double*)s->acc, s->d, s->nq);
    else
        k_import_acc<float><<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->stage, (float*)s->acc, s->d, s->nq);
    HIPCHK(hipGetLastError());
    s->acc_on = true;
    return LB_OK;
}

static int build_mask(lb_solver* s, bool with_walls) {
    k_mask<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->phase, with_walls ? s->vel : nullptr, s->mask,
                                                          s->wclass, s->d, s->nq);
    HIPCHK(hipGetLastError());
    return LB_OK;
}

// count fluid, moving-wall, special and Zou-He sites from the mask
__global__ void k_count(const uint32_t* mask, uint32_t nq, unsigned long long* out) {
    __shared__ unsigned long long sh[4];
    if (threadIdx.x < 4) sh[threadIdx.x] = 0;
    __syncthreads();
    unsigned long long a = 0, b = 0, c3 = 0, c4 = 0;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nq; c += gridDim.x * blockDim.x) {
        const uint32_t m = mask[c];
        a += (m & M_FLUID) ? 1 : 0;
        b += (m & M_MOVING) ? 1 : 0;
        c3 += ((m & M_FLUID) && (m & M_SPECIAL)) ? 1 : 0;
        c4 += ((m & M_FLUID) && (m & M_ZH_ANY)) ? 1 : 0;
    }
    atomicAdd(&sh[0], a);
    atomicAdd(&sh[1], b);
    atomicAdd(&sh[2], c3);
    atomicAdd(&sh[3], c4);
    __syncthreads();
    if (threadIdx.x < 4) atomicAdd(&out[threadIdx.x], sh[threadIdx.x]);
}

// compact list of the fluid sites with a bit of `need` and none of `skip`,
// appended at *counter: one atomic per wave, ascending order within a wave
__global__ void k_list_special(const uint32_t* mask, uint32_t nq, uint32_t need, uint32_t skip, uint32_t* list,
                               unsigned int* counter) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq) return;  // whole waves only: nq % 64 == 0
    const uint32_t m = mask[c];
    const bool sp = (m & M_FLUID) && (m & need) && !(m & skip);
    const unsigned long long b = __ballot(sp);
    if (!b) return;
    const int lane = threadIdx.x & 63;
    unsigned int base = 0;
    if (lane == 0) base = atomicAdd(counter, (unsigned int)__popcll(b));
    base = __shfl(base, 0, 64);
    if (sp) list[base + __popcll(b & ((1ull << lane) - 1ull))] = c;
}

// Moving sites (fluid sites with a bounce-back link to a moving solid site):
// an index per site and the 18 velocity sums of Ladd's term per site.
template <int... Ls>
LBD void ladd_sums(const Dims& d, const double* vel, int64_t nq, int i, int j, int k, uint32_t m, double* out,
                   std::integer_sequence<int, Ls...>) {
    ((out[Ls] = (Ls != REST && (m & (1u << Ls))) ? ladd_sum<Ls>(d, vel, nq, i, j, k) : 0.0), ...);
}

__global__ void k_ladd_prep(const uint32_t* mask, const double* vel, Dims d, uint32_t nq, uint32_t* mov_idx,
                            double* ladd_u, unsigned int* counter) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq) return;  // whole waves only: nq % 64 == 0
    const uint32_t m = mask[c];
    const bool mv = (m & M_FLUID) && (m & M_MOVING);
    const unsigned long long b = __ballot(mv);
    if (!b) return;
    const int lane = threadIdx.x & 63;
    unsigned int base = 0;
    if (lane == 0) base = atomicAdd(counter, (unsigned int)__popcll(b));
    base = __shfl(base, 0, 64);
    if (!mv) return;
    const uint32_t q = base + __popcll(b & ((1ull << lane) - 1ull));
    mov_idx[c] = q;
    int i, j, k;
    decompose(c, d.ny, d.nzp, i, j, k);
    double* out = ladd_u + (int64_t)q * (Q - 1);
    double tmp[Q];  // by link; the rest population has no table slot
    ladd_sums(d, vel, nq, i, j, k, m, tmp, std::make_integer_sequence<int, Q>{});
#pragma unroll
    for (int l = 0; l < Q; ++l)
        if (l != REST) out[link_slot(l)] = tmp[l];
}

// sparse layout set-up: fluid sites per wave of 64 sites
__global__ void k_wave_fluid(const uint32_t* mask, uint32_t nq, uint32_t* cnt) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq) return;  // whole waves only: nq % 64 == 0
    const unsigned long long b = __ballot((mask[c] & M_FLUID) != 0);
    if ((threadIdx.x & 63) == 0) cnt[c >> 6] = (uint32_t)__popcll(b);
}

// compact index of every fluid site from the per-wave offsets (ascending
// dense order), the site of every compact index; solid sites get ~0u
__global__ void k_compact(const uint32_t* mask, uint32_t nq, const uint32_t* woff, uint32_t* cidx, uint32_t* site) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nq) return;
    const bool fl = (mask[c] & M_FLUID) != 0;
    const unsigned long long b = __ballot(fl);
    const int lane = threadIdx.x & 63;
    if (!fl) {
        cidx[c] = ~0u;
        return;
    }
    const uint32_t n = woff[c >> 6] + (uint32_t)__popcll(b & ((1ull << lane) - 1ull));
    cidx[c] = n;
    site[n] = c;
}

// odd-step locations of every fluid site in compact indices. `q` addresses a
// dense SoA view (element l*nq + c), in which odd_loc resolves bounce-back,
// periodic, symmetric and z-wrapped links; a location on a non-fluid site
// (which the link mask rules out) raises *err instead of being stored.
__global__ void k_sparse_tab(StepParams q, const uint32_t* cidx, uint32_t nfp, uint32_t* tab, unsigned int* err) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= q.nq) return;
    const uint32_t m = q.mask[c];
    if (!(m & M_FLUID)) return;
    int i, j, k;
    decompose(c, q.d.ny, q.d.nzp, i, j, k);
    const uint32_t n = cidx[c];
    for (int l = 0; l < Q; ++l) {
        if (l == REST) continue;
        const int64_t e = odd_loc<true>(q, m, (int64_t)c, l, i, j, k);
        const int64_t ls = e / q.nq, cs = e - ls * q.nq;
        const uint32_t ns = (ls >= 0 && ls < Q && cs >= 0 && cs < q.nq) ? cidx[cs] : ~0u;
        if (ns == ~0u) {
            atomicOr(err, 1u);
            tab[(int64_t)link_slot(l) * nfp + n] = (uint32_t)REST * nfp + n;  // a valid location
        } else {
            tab[(int64_t)link_slot(l) * nfp + n] = (uint32_t)ls * nfp + ns;
        }
    }
}

// Whether the sparse layout steps by pull steps (run_steps), given the bytes
// of one population array. Pull steps against odd AA steps in place (ms per
// step, sparse porous media of 20% solid): FP64 0.0295 / 0.0254 at
// 128x64x64, 0.197 / 0.186 at 256x128x128, 0.650 / 0.686 at 384x192x192,
// 1.545 / 1.587 at 512x256x256, 5.28 / 5.46 at 768x384x384; FP32 2.76 /
// 2.83 and FP32 storage 2.77 / 2.80 at 768x384x384, within 1% either way
// from 384x192x192 to 512x256x256; FP16 1.59 / 1.57 at 768x384x384. Below
// about 1 GiB of populations the lattice sits largely in the 256 MiB last-level
// cache and the even AA step's purely local access wins; FP16 keeps the AA
// steps (and half the population memory).
static bool sp_pull_on(const lb_solver* s, size_t array_bytes) {
    if (s->sp_pull >= 0) return s->sp_pull != 0;
    return s->prec != LB_FP16_STORAGE && array_bytes >= ((size_t)1 << 30);
}

// Compact form of the odd-step table, one wave per 64 compact sites (see
// sparse_locs). Per link, over the wave's lanes that do not bounce back, the
// base is the value of loc - lane (modulo 2^32) whose window [base, base +
// SP_ALT) holds the most of them (the lowest such lane on ties); the
// alternate is the location of the lowest lane outside it, and every lane
// with that location takes SP_ALT.
// Counts of the link codes go to kinds[0..3]: near, SP_ALT, SP_BOUNCE, SP_FAR.
__global__ __launch_bounds__(64) void k_sparse_pack(const uint32_t* tab, uint32_t nfp, uint32_t nf, uint32_t* tabc,
                                                    uint32_t* tabw, unsigned long long* kinds) {
    const uint32_t w = blockIdx.x, lane = threadIdx.x;
    const uint32_t n = w * 64 + lane;  // < nfp: the grid is nfp / 64 waves
    const bool live = n < nf;
    uint32_t word[SP_WORDS + 1] = {};
    for (int l = 0; l < Q; ++l) {
        if (l == REST) continue;
        const uint32_t sl = link_slot(l);
        const uint32_t loc = tab[(size_t)sl * nfp + n];
        const bool bounce = live && loc == (uint32_t)l * nfp + n;
        const bool cand = live && !bounce;
        const uint32_t d = loc - lane;
        const uint64_t cands = __ballot(cand);
        int best = -1;
        uint32_t base = 0;
        for (int c = 0; c < 64; ++c) {
            const uint32_t dc = (uint32_t)__shfl((int)d, c, 64);
            if (!((cands >> c) & 1u)) continue;
            const int cnt = __popcll(__ballot(cand && d - dc < SP_ALT));
            if (cnt > best) {
                best = cnt;
                base = dc;
            }
        }
        const bool in = cand && d - base < SP_ALT;
        const uint64_t out = __ballot(cand && !in);
        const uint32_t alt = out ? (uint32_t)__shfl((int)loc, __ffsll((unsigned long long)out) - 1, 64) : 0u;
        uint32_t code = SP_FAR;  // also for lanes past the last site, which no step reads
        if (bounce) code = SP_BOUNCE;
        else if (in) code = d - base;
        else if (cand && loc == alt) code = SP_ALT;
        word[sl / SP_PER_WORD] |= code << (SP_BITS * (sl % SP_PER_WORD));
        const unsigned kn[4] = {(unsigned)__popcll(__ballot(in)),
                                (unsigned)__popcll(__ballot(cand && !in && code == SP_ALT)),
                                (unsigned)__popcll(__ballot(bounce)),
                                (unsigned)__popcll(__ballot(cand && code == SP_FAR))};
        if (lane == 0) {
            tabw[(size_t)w * SP_WAVE_WORDS + sl] = base;
            tabw[(size_t)w * SP_WAVE_WORDS + Q - 1 + sl] = alt;
            for (int k = 0; k < 4; ++k)
                if (kn[k]) atomicAdd(kinds + k, (unsigned long long)kn[k]);
        }
    }
    for (int q = 0; q < SP_WORDS; ++q) tabc[(size_t)q * nfp + n] = word[q];
    ((uint8_t*)(tabc + (size_t)SP_WORDS * nfp))[n] = (uint8_t)word[SP_WORDS];
}

static int count_mask(lb_solver* s, int64_t* nfluid, int64_t* nmoving) {
    unsigned long long* d = (unsigned long long*)s->red;
    HIPCHK(hipMemsetAsync(d, 0, 32, s->stream));
    k_count<<<256, 256, 0, s->stream>>>(s->mask, s->nq, d);
    HIPCHK(hipGetLastError());
    unsigned long long h[4];
    HIPCHK(hipMemcpyAsync(h, d, 32, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    *nfluid = (int64_t)h[0];
    *nmoving = (int64_t)h[1];
    const uint32_t nsp = (uint32_t)h[2];
    if (nsp > s->special_cap) {
        if (s->special) {
            HIPCHK(hipFree(s->special));
            s->bytes -= (size_t)s->special_cap * 4;
            s->special = nullptr;
        }
        int rc = dalloc(s, (void**)&s->special, (size_t)nsp * 4);
        if (rc) return rc;
        s->special_cap = nsp;
    }
    s->nspecial = nsp;
    s->nzh = (uint32_t)h[3];
    if (nsp > 0) {
        // Zou-He sites first (both parities), then the rest (odd steps only)
        unsigned int* cnt = (unsigned int*)s->red;
        HIPCHK(hipMemsetAsync(cnt, 0, 4, s->stream));
        k_list_special<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->mask, s->nq, M_ZH_ANY, 0u, s->special,
                                                                     cnt);
        k_list_special<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->mask, s->nq, M_SPECIAL, M_ZH_ANY,
                                                                     s->special, cnt);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    const size_t nmov = (size_t)*nmoving;
    if (nmov > 0) {
        int rc;
        if (!s->mov_idx && (rc = dalloc(s, (void**)&s->mov_idx, (size_t)s->nq * 4))) return rc;
        if (nmov > s->ladd_cap) {
            if (s->ladd_u) {
                HIPCHK(hipFree(s->ladd_u));
                s->bytes -= s->ladd_cap * (Q - 1) * 8;
                s->ladd_u = nullptr;
            }
            if ((rc = dalloc(s, (void**)&s->ladd_u, nmov * (Q - 1) * 8))) return rc;
            s->ladd_cap = nmov;
        }
        unsigned int* cnt = (unsigned int*)s->red;
        HIPCHK(hipMemsetAsync(cnt, 0, 4, s->stream));
        k_ladd_prep<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->mask, s->vel, s->d, s->nq, s->mov_idx,
                                                                  s->ladd_u, cnt);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s->stream));
    }
    return LB_OK;
}

// max |f - w rho0| of the initial state, for the FP16 storage scale
__global__ void k_maxdev(const uint32_t* mask, const double* den, const double* vel, uint32_t nq, double rho0,
                         double* part) {
    double a = 0.0;
    for (uint32_t c = blockIdx.x * blockDim.x + threadIdx.x; c < nq; c += gridDim.x * blockDim.x) {
        const double u = fabs(vel[c]) + fabs(vel[nq + c]) + fabs(vel[2 * (int64_t)nq + c]);
        if (mask[c] & M_FLUID) a = fmax(a, fabs(den[c] - rho0) + 3.0 * u + 4.5 * u * u);
        else a = fmax(a, 3.0 * u + 4.5 * u * u);  // wall velocities
    }
    a = wave_max(a);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = fmax(fmax(sh[0], sh[1]), fmax(sh[2], sh[3]));
}


static void launch_any_step(const lb_solver* s, const StepParams& ps, bool odd) {
    StepParams p = ps;
    if (!odd && s->even_rr) p.xcd = 0;
    switch (s->prec) {
    case LB_FP64: launch_step_rlx<LB_FP64>(p, s->cfg.rlx_scheme, odd, s->stream); break;
    case LB_FP32_SHIFTED: launch_step_rlx<LB_FP32_SHIFTED>(p, s->cfg.rlx_scheme, odd, s->stream); break;
    case LB_FP32_STORAGE: launch_step_rlx<LB_FP32_STORAGE>(p, s->cfg.rlx_scheme, odd, s->stream); break;
    default: launch_step_rlx<LB_FP16_STORAGE>(p, s->cfg.rlx_scheme, odd, s->stream); break;
    }
}

// sum of one nq-long row of the compute type, 256 partial sums into part
template <class TC>
__global__ __launch_bounds__(256) void k_sum_row(const TC* v, uint32_t n, double* part) {
    double a = 0.0;
    for (uint32_t c = blockIdx.x * 256 + threadIdx.x; c < n; c += gridDim.x * 256) a += (double)v[c];
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    __shared__ double sh[4];
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = a;
    __syncthreads();
    if (threadIdx.x == 0) part[blockIdx.x] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

const char* lb_last_error(void) { return g_err.c_str(); }

int lb_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

// Lattices below this many cells are not launch-tuned by default.
constexpr uint64_t TUNE_MIN_CELLS = 1ull << 22;

// Launch defaults of a population layout, then the environment overrides.
static void layout_defaults(lb_solver* s, int layout) {
    const bool sparse = layout == LB_LAYOUT_SPARSE;
    // The sparse step on the 768x384x384 medium, by tile order (ms per step,
    // FP64 / FP32 / FP16 / FP32 storage): round-robin 6.29 / 2.86 / 1.66 /
    // 2.81, contiguous ranges per XCD 5.90 / 3.03 / 1.66 / 3.05, and chunks
    // of 32 / 4 / 8 / 2 tiles dealt to the XCDs in turn 5.49 / 2.82 / 1.57 /
    // 2.79 (chunks of 3, 6, 12 or 24 tiles ran 1-4% slower than the powers
    // of two beside them). The dense layouts keep contiguous ranges: chunks
    // were within +-1.5% on the 256^3 cavity and the 512x128x128 duct and
    // 2-7% slower on the medium.
    // The sparse FP64 even step, which reads and writes only its own slots,
    // streamed 2.8% faster in round-robin order (4.97 against 5.11 ms) while
    // its odd step ran 27% slower there; so the even steps of that variant
    // alone run round-robin.
    const uint32_t chunk[4] = {32, 4, 8, 2};  // by LB_FP64, LB_FP32_SHIFTED, LB_FP16_STORAGE, LB_FP32_STORAGE
    s->xcd_chunk = sparse ? chunk[s->prec] : 0u;
    // Small sparse lattices, which the launch tuner leaves alone: with 4- and
    // 2-byte populations contiguous ranges per XCD, whose L2 then holds most
    // of the upstream rows of its range. Sparse media of 20% solid, ms per
    // step in contiguous ranges / round-robin: 128x64x64 FP32 0.0135 /
    // 0.0156, FP16 0.0120 / 0.0135, FP32 storage 0.0149 / 0.0165, FP64
    // 0.0273 / 0.0265; 256x128x128 FP32 0.0944 / 0.0984, FP16 0.0627 /
    // 0.0650 (tuned there, with the XCD orders among the candidates);
    // 384x192x192 FP32 0.359 / 0.345.
    if (sparse && s->prec != LB_FP64 && s->nq < TUNE_MIN_CELLS) {
        s->xcd_cells = 0;
        s->xcd_chunk = 0;
    }
    s->even_rr = sparse && s->prec == LB_FP64;
    if (const char* env = getenv("LB_EVEN_RR")) s->even_rr = strtol(env, nullptr, 10) != 0;
    s->sp_pull = -1;  // decided by sp_pull_on at lb_finalize
    if (const char* env = getenv("LB_SP_PULL")) s->sp_pull = strtol(env, nullptr, 10) != 0 ? 1 : 0;
    s->adj_pull = -1;
    if (const char* env = getenv("LB_ADJ_PULL")) s->adj_pull = strtol(env, nullptr, 10) != 0 ? 1 : 0;
    if (const char* env = getenv("LB_XCD_CHUNK")) {
        s->xcd_chunk = (uint32_t)strtoul(env, nullptr, 10);
        s->tune_xcd = false;
    }
    if (const char* env = getenv("LB_XCD_REMAP_CELLS")) {
        s->xcd_cells = strtoull(env, nullptr, 10);
        s->tune_xcd = false;
    }
    // Unused dynamic LDS caps the bulk step's occupancy. The FP32 kernel needs
    // few registers (8 waves per SIMD); its even step streamed 3-4% faster at
    // 6 waves per SIMD (26 KB per workgroup of 256 threads), while capping the
    // other variants or the odd step measured slower. In the sparse layout the
    // same cap speeds up the FP32 pull step (60 VGPRs): 2.684-2.706 against
    // 2.754 ms per step on the 768x384x384 medium, 0.808 against 0.826 at
    // 512x256x256 (FP32 storage, FP64 and FP16 unchanged or slower); the odd
    // AA step, at 5 waves per SIMD, is not capped by it.
    if (s->prec == LB_FP32_SHIFTED && !sparse) s->step_lds[0] = 26000;
    if (s->prec == LB_FP32_SHIFTED && sparse) s->step_lds[1] = 27000;
    if (const char* env = getenv("LB_STEP_LDS_EVEN")) s->step_lds[0] = (uint32_t)strtoul(env, nullptr, 10);
    if (const char* env = getenv("LB_STEP_LDS_ODD")) s->step_lds[1] = (uint32_t)strtoul(env, nullptr, 10);
    for (uint32_t& b : s->step_lds) b = b > 65536u ? 65536u : b;  // 64 KiB keeps one workgroup per CU
    // Bulk step workgroups: the dense FP64 kernel (4 waves per SIMD) measured
    // 4% faster with 512 threads on the 256^3 cavity (1.3% slower on the
    // 512x128x128 duct, even on porous media); FP32 storage measured 8% slower
    // with 512 on the cavity, and the sparse FP64 odd step holds 3 waves per
    // SIMD, of which a 512-thread workgroup would leave 2: those and the
    // rest keep 256. LB_STEP_TILE (a multiple of 64 up to 512) overrides it.
    if (s->prec == LB_FP64 && !sparse) s->step_tile = 512;
    if (const char* env = getenv("LB_STEP_TILE")) {
        const unsigned long t = strtoul(env, nullptr, 10);
        if (t >= 64 && t <= (unsigned long)STEP_TILE_MAX && t % 64 == 0) s->step_tile = (uint32_t)t;
    }
}

// Strides and population array of a layout: row-interleaved f[i][j][l][k]
// (ROWS), plain SoA f[l][i][j][k], or SoA with the population stride padded
// by 2304 bytes so that equal offsets of different populations do not share
// a channel. The sparse layout sizes its arrays from the fluid sites at
// lb_finalize (build_sparse).
static int layout_arrays(lb_solver* s, int layout) {
    s->layout = layout;
    s->sparse = layout == LB_LAYOUT_SPARSE;
    if (s->sparse) {
        s->fsize = 0;
        return LB_OK;
    }
    const int64_t rows = (int64_t)s->d.nx * s->d.ny;
    if (layout == LB_LAYOUT_SOA || layout == LB_LAYOUT_SOA_PADDED) {
        s->rs = s->d.nzp;
        s->ps = (int64_t)s->nq + (layout == LB_LAYOUT_SOA_PADDED ? (int64_t)(2304 / ts_size(s->prec)) : 0);
        s->fsize = (int64_t)Q * s->ps;
    } else {
        s->rs = (int64_t)Q * s->d.nzp;
        s->ps = s->d.nzp;
        s->fsize = rows * s->rs;
    }
    const size_t fb = (size_t)s->fsize * ts_size(s->prec);
    int rc = dalloc(s, &s->f, fb);
    if (rc) return rc;
    HIPCHK(hipMemset(s->f, 0, fb));
    // Dense layouts: pull steps when asked for (LB_SP_PULL=1), or as the
    // launch tuner's last candidate (tune_step), which frees the second
    // array again if the in-place steps stay faster.
    if (s->sp_pull == 1 || (s->sp_pull < 0 && s->tune)) {
        dalloc_optional(s, &s->f2, fb);
        if (s->f2) HIPCHK(hipMemset(s->f2, 0, fb));
        s->pull_active = s->f2 && s->sp_pull == 1;
    }
    return LB_OK;
}

// The second array of a dense pull-step candidate that was not chosen.
static void drop_pull_candidate(lb_solver* s) {
    if (s->f2 && !s->pull_active) dfree(s, &s->f2, (size_t)s->fsize * ts_size(s->prec));
}

int lb_create(const lb_config* cfg, lb_solver** out) {
    if (!cfg || !out) return fail(LB_EINVAL, "null argument");
    *out = nullptr;
    int rc = check_config(cfg);
    if (rc) return rc;
    if (lb_device_count() < 1) return fail(LB_EHIP, "no HIP device available");
    lb_solver* s = new lb_solver();
    s->cfg = *cfg;
    s->d.nx = cfg->lat[0];
    s->d.ny = cfg->lat[1];
    s->d.nz = cfg->lat[2];
    s->d.nzp = (cfg->lat[2] + 63) / 64 * 64;
    for (int f = 0; f < 6; ++f) s->d.bc[f] = cfg->bc[f];
    s->NI = s->d.nx + 2;
    s->NJ = s->d.ny + 2;
    s->NK = s->d.nz + 2;
    s->nq = (uint32_t)((int64_t)s->d.nx * s->d.ny * s->d.nzp);
    s->ncell = (int64_t)s->d.nx * s->d.ny * s->d.nz;
    s->prec = cfg->precision;
    if (cfg->layout != LB_LAYOUT_AUTO) layout_defaults(s, cfg->layout);
    // Launch tuning on the first long evolve call, for lattices of at least
    // 2^22 cells (below that launch overhead dominates the timings), unless
    // the launch is pinned by LB_STEP_TILE / LB_STEP_LDS_* or LB_AUTOTUNE=0.
    // LB_AUTOTUNE_MIN_CELLS changes the size, LB_AUTOTUNE_VERBOSE=1 prints
    // the timings.
    {
        uint64_t min_cells = TUNE_MIN_CELLS;
        if (const char* env = getenv("LB_AUTOTUNE_MIN_CELLS")) min_cells = strtoull(env, nullptr, 10);
        const char* at = getenv("LB_AUTOTUNE");
        s->tune = (!at || strtol(at, nullptr, 10) != 0) && !getenv("LB_STEP_TILE") && !getenv("LB_STEP_LDS_EVEN") &&
                  !getenv("LB_STEP_LDS_ODD") && (uint64_t)s->nq >= min_cells;
        const char* vb = getenv("LB_AUTOTUNE_VERBOSE");
        s->tune_verbose = vb && strtol(vb, nullptr, 10) != 0;
    }
    s->rho0 = cfg->rho0 > 0 ? cfg->rho0 : 1.0;
    s->scale = cfg->fp16_scale > 0 ? cfg->fp16_scale : 1.0;
    s->f16_auto = !(cfg->fp16_scale > 0);
    if (const char* env = getenv("LB_FP16_CHECK")) {
        const unsigned long v = strtoul(env, nullptr, 10);
        if (v >= 1) s->f16_check = (uint32_t)v;
    }
    // relaxation parameters, isothermal.pyx:273-287
    const double tau_e = 0.5 + 3.0 * cfg->kvisc;
    const double inv_tau_e = 1.0 / tau_e;
    s->prm[0] = inv_tau_e;
    s->prm[1] = cfg->rlx_scheme == TRT ? 8.0 * (2.0 - inv_tau_e) / (8.0 - inv_tau_e) : 0.0;
    const size_t nq = s->nq;
    if ((cfg->layout != LB_LAYOUT_AUTO && (rc = layout_arrays(s, cfg->layout))) ||
        (rc = dalloc(s, (void**)&s->mask, nq * 4)) || (rc = dalloc(s, (void**)&s->phase, nq)) ||
        (rc = dalloc(s, (void**)&s->wclass, nq / 64)) ||
        (rc = dalloc(s, (void**)&s->den, nq * 8)) || (rc = dalloc(s, (void**)&s->vel, 3 * nq * 8)) ||
        (rc = dalloc(s, (void**)&s->part, (size_t)RED_PARTS * FLOW_N * 8)) ||
        (rc = dalloc(s, (void**)&s->red, FLOW_N * 8))) {
        std::string msg = g_err;
        lb_destroy(s);
        g_err = msg;
        return rc;
    }
    // den = 1, vel = 0, phase SOLID, mask 0 (isothermal.pyx:243-248)
    (void)hipMemset(s->mask, 0, nq * 4);
    (void)hipMemset(s->wclass, 1, nq / 64);
    (void)hipMemset(s->phase, 0, nq);
    (void)hipMemset(s->vel, 0, 3 * nq * 8);
    std::vector<double> ones;  // den = 1 via a host copy of one row block
    {
        const size_t chunk = 1 << 20;
        ones.assign(chunk, 1.0);
        for (size_t o = 0; o < nq; o += chunk)
            (void)hipMemcpy(s->den + o, ones.data(), std::min(chunk, nq - o) * 8, hipMemcpyHostToDevice);
    }
    for (int f = 0; f < 6; ++f) {
        if (cfg->bc[f] != FIXED_DEN && cfg->bc[f] != FIXED_VEL) continue;
        const size_t n = (size_t)(f < 2 ? s->NJ * s->NK : f < 4 ? s->NI * s->NK : s->NI * s->NJ);
        if ((rc = dalloc(s, (void**)&s->fden[f], n * 8)) || (rc = dalloc(s, (void**)&s->fvel[f], 3 * n * 8))) {
            std::string msg = g_err;
            lb_destroy(s);
            g_err = msg;
            return rc;
        }
        std::vector<double> o1(n, 1.0), z3(3 * n, 0.0);
        (void)hipMemcpy(s->fden[f], o1.data(), n * 8, hipMemcpyHostToDevice);
        (void)hipMemcpy(s->fvel[f], z3.data(), 3 * n * 8, hipMemcpyHostToDevice);
    }
    if (hipDeviceSynchronize() != hipSuccess) {
        lb_destroy(s);
        return fail(LB_EHIP, "device initialisation failed");
    }
    *out = s;
    return LB_OK;
}

int lb_destroy(lb_solver* s) {
    if (!s) return LB_OK;
    if (s->stream) (void)hipStreamSynchronize(s->stream);
    else (void)hipDeviceSynchronize();
    void* ptrs[] = {s->f,     s->mask, s->wclass, s->special, s->phase, s->den,  s->vel,  s->acc,
                    s->stage, s->part, s->red,    s->uprev,   s->ckpt,  s->lam,  s->lam2, s->rec,  s->gacc, s->gface,
                    s->snap,  s->snap_dv, s->mov_idx, s->ladd_u, s->cidx, s->site, s->tab,
                    s->tabc, s->tabw, s->f2, s->prev, s->dseed, s->st_phase, s->st_den, s->st_vel, s->st_acc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int f = 0; f < 6; ++f) {
        if (s->fden[f]) (void)hipFree(s->fden[f]);
        if (s->fvel[f]) (void)hipFree(s->fvel[f]);
    }
    delete s;
    return LB_OK;
}

int lb_set_stream(lb_solver* s, void* st) {
    if (!s) return fail(LB_EINVAL, "null handle");
    s->stream = (hipStream_t)st;
    return LB_OK;
}

int lb_set_gravity(lb_solver* s, const double g[3]) {
    if (!s || !g) return fail(LB_EINVAL, "null argument");
    inputs_changed(s);
    for (int a = 0; a < 3; ++a) s->cfg.g[a] = g[a];
    return LB_OK;
}

int lb_upload_phase(lb_solver* s, const uint8_t* phase) {
    if (!s || !phase) return fail(LB_EINVAL, "null argument");
    inputs_changed(s);
    int rc = ensure_stage(s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s->stage, phase, (size_t)s->ncell, hipMemcpyHostToDevice, s->stream));
    k_import_phase<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>((const uint8_t*)s->stage, s->phase, s->d, s->nq);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    s->phase_set = true;
    return LB_OK;
}

int lb_upload_den_vel(lb_solver* s, const double* den, const double* vel) {
    if (!s || !den || !vel) return fail(LB_EINVAL, "null argument");
    inputs_changed(s);
    int rc = ensure_stage(s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s->stage, den, (size_t)s->ncell * 8, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->stage + s->ncell, vel, (size_t)s->ncell * 24, hipMemcpyHostToDevice, s->stream));
    k_import_dv<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->stage, s->den, s->vel, s->d, s->nq, 1);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

int lb_upload_acc(lb_solver* s, const double* acc) {
    if (!s) return fail(LB_EINVAL, "null handle");
    inputs_changed(s);
    if (!acc) {
        if (s->acc) {
            HIPCHK(hipStreamSynchronize(s->stream));
            (void)hipFree(s->acc);
            s->bytes -= (size_t)3 * s->nq * tc_size(s->prec);
            s->acc = nullptr;
        }
        s->acc_on = false;
        return LB_OK;
    }
    int rc = ensure_stage(s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s->stage, acc, (size_t)s->ncell * 24, hipMemcpyHostToDevice, s->stream));
    if ((rc = import_acc_from_stage(s))) return rc;
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

int lb_upload_acc_device(lb_solver* s, const double* acc_dev) {
    if (!s || !acc_dev) return fail(LB_EINVAL, "null argument");
    inputs_changed(s);
    int rc = ensure_stage(s);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s->stage, acc_dev, (size_t)s->ncell * 24, hipMemcpyDeviceToDevice, s->stream));
    return import_acc_from_stage(s);
}

int lb_set_face_data(lb_solver* s, int face, const double* den_face, const double* vel_face) {
    if (!s || face < 0 || face > 5 || !den_face || !vel_face) return fail(LB_EINVAL, "invalid argument");
    inputs_changed(s);
    if (!s->fden[face]) return fail(LB_EINVAL, "face %d has no FIXED_DEN/FIXED_VEL condition", face);
    const size_t n = (size_t)(face < 2 ? s->NJ * s->NK : face < 4 ? s->NI * s->NK : s->NI * s->NJ);
    HIPCHK(hipMemcpyAsync(s->fden[face], den_face, n * 8, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->fvel[face], vel_face, 3 * n * 8, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

// Sparse layout: compact the fluid sites in ascending dense order, size the
// population array to them (stride nfp) and tabulate every odd-step location.
static int build_sparse(lb_solver* s) {
    const uint64_t nf = (uint64_t)s->nfluid;
    const uint64_t nfp = std::max<uint64_t>(64, (nf + 63) / 64 * 64);
    if ((uint64_t)Q * nfp >= (1ull << 32))
        return fail(LB_EINVAL, "sparse layout: %llu fluid sites exceed its 32-bit location table",
                    (unsigned long long)nf);
    const size_t ts = ts_size(s->prec);
    int rc;
    if (s->nfp != nfp) {
        const size_t o = s->nfp;
        dfree(s, &s->f, Q * o * ts);
        dfree(s, &s->snap, Q * o * ts);
        dfree(s, (void**)&s->site, o * 4);
        dfree(s, (void**)&s->tab, (Q - 1) * o * 4);
        dfree(s, (void**)&s->tabc, SP_CODE_BYTES * o);
        dfree(s, (void**)&s->tabw, SP_WAVE_WORDS * (o / 64) * 4);
        dfree(s, &s->f2, Q * o * ts);
        s->nfp = 0;
        if ((rc = dalloc(s, &s->f, Q * nfp * ts)) || (rc = dalloc(s, (void**)&s->site, nfp * 4)) ||
            (rc = dalloc(s, (void**)&s->tab, (Q - 1) * nfp * 4)) ||
            (rc = dalloc(s, (void**)&s->tabc, SP_CODE_BYTES * nfp)) ||
            (rc = dalloc(s, (void**)&s->tabw, SP_WAVE_WORDS * (nfp / 64) * 4)))
            return rc;
        if (sp_pull_on(s, Q * nfp * ts)) dalloc_optional(s, &s->f2, Q * nfp * ts);
        s->pull_active = s->f2 != nullptr;
        s->nfp = (uint32_t)nfp;
    }
    if (!s->cidx && (rc = dalloc(s, (void**)&s->cidx, (size_t)s->nq * 4))) return rc;
    s->ps = (int64_t)nfp;
    s->rs = 0;
    s->fsize = (int64_t)Q * nfp;
    HIPCHK(hipMemsetAsync(s->f, 0, Q * nfp * ts, s->stream));  // padding sites stay zero
    if (s->f2) HIPCHK(hipMemsetAsync(s->f2, 0, Q * nfp * ts, s->stream));
    HIPCHK(hipMemsetAsync(s->site, 0, nfp * 4, s->stream));
    // fluid sites per wave -> exclusive scan on the host -> compact indices
    const uint32_t nw = s->nq / 64;
    uint32_t* woff = nullptr;
    if ((rc = dalloc(s, (void**)&woff, (size_t)nw * 4))) return rc;
    k_wave_fluid<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->mask, s->nq, woff);
    std::vector<uint32_t> h(nw);
    hipError_t e = hipMemcpyAsync(h.data(), woff, (size_t)nw * 4, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    uint64_t run = 0;
    for (uint32_t& v : h) {
        const uint32_t t = v;
        v = (uint32_t)run;
        run += t;
    }
    if (e == hipSuccess && run == nf)
        e = hipMemcpyAsync(woff, h.data(), (size_t)nw * 4, hipMemcpyHostToDevice, s->stream);
    if (e == hipSuccess && run == nf) {
        k_compact<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->mask, s->nq, woff, s->cidx, s->site);
        e = hipStreamSynchronize(s->stream);
    }
    dfree(s, (void**)&woff, (size_t)nw * 4);
    if (e != hipSuccess) return fail(LB_EHIP, "sparse layout set-up: %s", hipGetErrorString(e));
    if (run != nf) return fail(LB_EHIP, "sparse layout set-up: %llu fluid sites counted, %llu listed",
                               (unsigned long long)nf, (unsigned long long)run);
    // odd-step locations, resolved in a dense SoA view: element l*nq + c
    StepParams q = params(s);
    q.cidx = nullptr;
    q.tab = nullptr;
    q.rs = s->d.nzp;
    q.ps = s->nq;
    unsigned int* err = (unsigned int*)s->red;
    HIPCHK(hipMemsetAsync(err, 0, 4, s->stream));
    k_sparse_tab<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(q, s->cidx, (uint32_t)nfp, s->tab, err);
    HIPCHK(hipGetLastError());
    unsigned int herr = 0;
    HIPCHK(hipMemcpyAsync(&herr, err, 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (herr) return fail(LB_EHIP, "sparse layout set-up: a link resolves to a non-fluid site");
    unsigned long long* kinds = (unsigned long long*)s->red;
    HIPCHK(hipMemsetAsync(kinds, 0, sizeof(s->sp_kinds), s->stream));
    k_sparse_pack<<<(unsigned)(nfp / 64), 64, 0, s->stream>>>(s->tab, (uint32_t)nfp, (uint32_t)nf, s->tabc,
                                                             s->tabw, kinds);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(s->sp_kinds, kinds, sizeof(s->sp_kinds), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

int lb_finalize(lb_solver* s, int64_t* nfluid) {
    if (!s) return fail(LB_EINVAL, "null handle");
    inputs_changed(s);
    if (!s->phase_set) return fail(LB_ESTATE, "phase field not uploaded");
    s->finalized = false;  // until the populations below are initialised
    s->prev_valid = false;
    int rc = build_mask(s, true);
    if (rc) return rc;
    int64_t nf, nm;
    if ((rc = count_mask(s, &nf, &nm))) return rc;
    s->nfluid = nf;
    s->moving = nm > 0;
    if (s->layout == LB_LAYOUT_AUTO) {
        // the sparse layout moves the fluid sites' populations only, but reads
        // 18 B of location codes per odd step: it wins from a few per cent of
        // solid cells on
        double thr = LB_AUTO_SPARSE_SOLID;
        if (const char* env = getenv("LB_AUTO_SPARSE_SOLID")) thr = strtod(env, nullptr);
        const double solid = 1.0 - (double)nf / (double)s->ncell;
        const int layout = solid >= thr ? LB_LAYOUT_SPARSE : LB_LAYOUT_ROWS;
        layout_defaults(s, layout);
        if ((rc = layout_arrays(s, layout))) return rc;
    }
    if (s->sparse && (rc = build_sparse(s))) return rc;
    if (s->prec == LB_FP16_STORAGE && s->f16_auto) {
        k_maxdev<<<256, 256, 0, s->stream>>>(s->mask, s->den, s->vel, s->nq, s->rho0, s->part);
        HIPCHK(hipGetLastError());
        std::vector<double> h(256);
        HIPCHK(hipMemcpyAsync(h.data(), s->part, 256 * 8, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        double mx = 0.0;
        for (double v : h) mx = fmax(mx, v);
        double fb = 0.0;
        if ((rc = face_bound(s, &fb))) return rc;
        // populations are O(w * mx) with w <= 1/3; keep them near 2^10 so that the
        // f16 range (6.1e-5 .. 65504) leaves ~64x headroom for growth, which the
        // guard (fp16_guard) watches from then on
        const double est = fmax(fmax(mx, fb) / 3.0, 1e-4);
        s->scale = ldexp(1.0, (int)floor(log2(1024.0 / est)));
        s->scale = fmin(fmax(s->scale, 1.0), ldexp(1.0, 24));
    }
    s->f16_since = 0;
    switch (s->prec) {
    case LB_FP64: launch_init<LB_FP64>(params(s), s->stream); break;
    case LB_FP32_SHIFTED: launch_init<LB_FP32_SHIFTED>(params(s), s->stream); break;
    case LB_FP32_STORAGE: launch_init<LB_FP32_STORAGE>(params(s), s->stream); break;
    default: launch_init<LB_FP16_STORAGE>(params(s), s->stream); break;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(s->stream));
    s->finalized = true;
    s->steps = 0;
    s->next_odd = 1;
    s->have_uprev = false;
    if (nfluid) *nfluid = nf;
    return LB_OK;
}

int lb_refresh_walls(lb_solver* s) {
    if (!s) return fail(LB_EINVAL, "null handle");
    inputs_changed(s);
    if (!s->phase_set) return fail(LB_ESTATE, "phase field not uploaded");
    int rc = build_mask(s, true);
    if (rc) return rc;
    int64_t nf, nm;
    if ((rc = count_mask(s, &nf, &nm))) return rc;
    s->moving = nm > 0;
    return LB_OK;
}

// Zou-He data bound of |f - w rho0| over the FIXED faces (k_face_bound), 0
// without such faces
static int face_bound(lb_solver* s, double* out) {
    *out = 0.0;
    int nf = 0;
    for (int f = 0; f < 6; ++f) {
        if (!s->fden[f]) continue;
        const size_t n = (size_t)(f < 2 ? s->NJ * s->NK : f < 4 ? s->NI * s->NK : s->NI * s->NJ);
        k_face_bound<<<64, 256, 0, s->stream>>>(s->fden[f], s->fvel[f], n, s->rho0, s->part + 64 * nf);
        ++nf;
    }
    if (!nf) {
        HIPCHK(hipStreamSynchronize(s->stream));
        return LB_OK;
    }
    HIPCHK(hipGetLastError());
    std::vector<double> h(64 * nf);
    HIPCHK(hipMemcpyAsync(h.data(), s->part, h.size() * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    for (double v : h) *out = fmax(*out, v);
    return LB_OK;
}

static double half_bits_value(uint32_t b) {
    const int e = (int)((b >> 10) & 0x1f), m = (int)(b & 0x3ff);
    return e ? ldexp(1024.0 + m, e - 25) : ldexp((double)m, -24);
}

// FP16 storage guard. The stored values are s (f - w rho0). Their largest
// magnitude, and the bound the Zou-He face data imply, are kept at or below
// 2^13 (and, with an automatic scale, at or above 2^6) by exact power-of-two
// rescales of the whole array, so that a flow which accelerates (gravity, an
// inlet switched on) does not overflow the f16 range (65504) and one that
// stays slow keeps its precision. Runs at the start of every evolve call and
// every f16_check steps within one. A non-finite stored value (an overflow
// between two checks) is an error, never a silent inf.
static int fp16_guard(lb_solver* s) {
    if (s->prec != LB_FP16_STORAGE || !s->f) return LB_OK;
    unsigned int* d = (unsigned int*)s->red;
    HIPCHK(hipMemsetAsync(d, 0, 4, s->stream));
    const size_t n8 = (size_t)s->fsize / 8;  // fsize is a multiple of 64
    k_f16_amax<<<1024, 256, 0, s->stream>>>((const uint4*)s->f, n8, d);
    HIPCHK(hipGetLastError());
    unsigned int bits = 0;
    HIPCHK(hipMemcpyAsync(&bits, d, 4, hipMemcpyDeviceToHost, s->stream));
    double fb = 0.0;
    int rc = face_bound(s, &fb);  // synchronises the stream
    if (rc) return rc;
    s->f16_since = 0;
    s->f16_dirty = false;
    if (bits >= 0x7c00u)
        return fail(LB_ESTATE, "FP16 storage overflowed: non-finite populations after step %llu (scale %g); "
                    "use a smaller fp16_scale or FP32", (unsigned long long)s->steps, s->scale);
    const double m = fmax(half_bits_value(bits), s->scale * fb / 3.0);
    int e = 0;
    if (m > 8192.0) e = 10 - (int)ceil(log2(m));
    else if (s->f16_auto && m > 0.0 && m < 64.0) e = std::min(10 - (int)ceil(log2(m)), 24 - (int)lround(log2(s->scale)));
    if (e == 0) return LB_OK;
    k_f16_scale<<<1024, 256, 0, s->stream>>>((_Float16*)s->f, (size_t)s->fsize, ldexpf(1.0f, e));
    HIPCHK(hipGetLastError());
    s->scale = ldexp(s->scale, e);
    return LB_OK;
}

// `n` steps with the solver's current launch configuration; the last one
// stores den/vel if `store_last`. FP16 storage runs its guard every
// f16_check steps.
// One step from the state in s->f, whose parity is `odd` (updated). Where
// pull steps are active and the second array is there, an odd step is a pull
// step: it leaves the state an even step leaves, in s->f2, which then becomes
// s->f, and the parity stays odd. Otherwise an in-place AA step.
static void one_step(lb_solver* s, StepParams& p, int& odd) {
    if (odd && s->pull_active && s->f2) {
        p.f2 = s->f2;
        launch_any_step(s, p, true);
        std::swap(s->f, s->f2);
        p.f = s->f;
        p.f2 = nullptr;
    } else {
        launch_any_step(s, p, odd != 0);
        odd ^= 1;
    }
}

static int run_steps(lb_solver* s, uint32_t n, bool store_last) {
    StepParams p = params(s);
    for (uint32_t t = 0; t < n; ++t) {
        if (s->prec == LB_FP16_STORAGE && s->f16_since >= s->f16_check) {
            int rc = fp16_guard(s);
            if (rc) return rc;
            p = params(s);
        }
        p.store_dv = (store_last && t + 1 == n) ? 1 : 0;
        one_step(s, p, s->next_odd);
        s->steps++;
        s->f16_since++;
    }
    return LB_OK;
}

// Launch configurations of the bulk step that the first long evolve call
// chooses between (tune_step). The best one depends on the workload: for
// FP64, 4 waves per SIMD in 512-thread workgroups were best on the 256^3
// cavity and 3 waves (256 threads, LDS-capped) 5% faster on the 512x128x128
// duct. Every configuration gives bit-identical results, so the candidates
// are timed on the call's own steps.
struct StepCfg {
    uint32_t tile, lds_even, lds_odd;
    uint64_t xcd_cells;  // see tile_index: 0 = XCD order, UINT64_MAX = round-robin
    uint32_t chunk;      // XCD order: tiles per chunk (0: contiguous ranges)
};
constexpr int TUNE_WARM = 2, TUNE_STEPS = 4;
constexpr float TUNE_MARGIN = 0.02f;

// the XCD tile orders are candidates too in the sparse layout and on dense
// lattices of at least 2^24 cells
constexpr uint64_t TUNE_XCD_CELLS = 1ull << 24;
constexpr int TUNE_MAX = 10;

static int tune_candidates(const lb_solver* s, StepCfg* c) {
    int n = 0;
    const uint64_t xd = s->nq >= s->xcd_cells ? 0 : UINT64_MAX;  // the default order
    const uint32_t ch = s->xcd_chunk;
    c[n++] = {s->step_tile, s->step_lds[0], s->step_lds[1], xd, ch};
    auto add = [&](StepCfg x) {
        if (x.xcd_cells) x.chunk = ch;  // unused in round-robin order
        for (int i = 0; i < n; ++i)
            if (c[i].tile == x.tile && c[i].lds_even == x.lds_even && c[i].lds_odd == x.lds_odd &&
                c[i].xcd_cells == x.xcd_cells && c[i].chunk == x.chunk)
                return;
        if (n < TUNE_MAX) c[n++] = x;
    };
    if (!s->sparse) {
        if (s->prec == LB_FP64 || s->prec == LB_FP32_STORAGE) {
            add({512, 0, 0, xd, ch});
            add({256, 0, 0, xd, ch});
            add({256, 40000, 40000, xd, ch});  // 4 workgroups per CU: 4 waves per SIMD
            add({256, 50000, 50000, xd, ch});  // 3 waves per SIMD
        } else if (s->prec == LB_FP32_SHIFTED) {
            add({256, 0, 0, xd, ch});
            add({256, 26000, 0, xd, ch});
        }
    }
    if (s->tune_xcd && (s->sparse || (uint64_t)s->nq >= TUNE_XCD_CELLS)) {  // each of them in the other orders
        const int m = n;
        for (int i = 0; i < m; ++i) {
            add({c[i].tile, c[i].lds_even, c[i].lds_odd, xd ? 0 : UINT64_MAX, ch});
            if (ch) add({c[i].tile, c[i].lds_even, c[i].lds_odd, 0, 0});  // contiguous ranges
        }
    }
    return n;
}

// The dense layouts' pull steps as the tuner's last candidate, with the
// chosen launch: the in-place steps and then the pull steps, TUNE_STEPS each;
// the pull steps stay if more than TUNE_MARGIN faster. Measured with the
// tuning off (ms per step, pull / in place): 256^3 cavity FP64 0.929 / 0.956,
// FP16 0.284 / 0.298, FP32 0.515 / 0.493, FP32 storage 0.507 / 0.515;
// 512x128x128 duct FP64 0.465 / 0.452, FP16 0.152 / 0.159; the dense
// 768x384x384 medium FP64 7.06 / 7.48, FP16 2.22 / 2.48.
constexpr int TUNE_PULL_STEPS = 2 * TUNE_STEPS + 1;
static int tune_pull(lb_solver* s) {
    if (!s->f2 || s->pull_active) return LB_OK;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool ok = true;
    for (hipEvent_t& e : ev) ok = ok && hipEventCreate(&e) == hipSuccess;
    int rc = LB_OK;
    float ms_aa = 0.0f, ms_pull = 0.0f;
    if (ok) {
        HIPCHK(hipEventRecord(ev[0], s->stream));
        rc = run_steps(s, TUNE_STEPS, false);
        if (!rc) {
            HIPCHK(hipEventRecord(ev[1], s->stream));
            s->pull_active = true;
            rc = run_steps(s, 1, false);  // reaches the state pull steps start from
        }
        if (!rc) {
            HIPCHK(hipEventRecord(ev[2], s->stream));
            rc = run_steps(s, TUNE_STEPS, false);
        }
        if (!rc) {
            HIPCHK(hipEventRecord(ev[3], s->stream));
            HIPCHK(hipEventSynchronize(ev[3]));
            HIPCHK(hipEventElapsedTime(&ms_aa, ev[0], ev[1]));
            HIPCHK(hipEventElapsedTime(&ms_pull, ev[2], ev[3]));
            if (s->tune_verbose)
                fprintf(stderr, "[jyulb] in-place steps %.4f ms, pull steps %.4f ms per step\n", ms_aa / TUNE_STEPS,
                        ms_pull / TUNE_STEPS);
        }
    }
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    s->pull_active = ok && !rc && ms_pull < (1.0f - TUNE_MARGIN) * ms_aa;
    if (!s->sparse) drop_pull_candidate(s);  // the state is in s->f either way
    return rc;
}

static int tune_step(lb_solver* s) {
    StepCfg c[TUNE_MAX];
    const int n = tune_candidates(s, c);
    s->tune = false;
    if (n < 2) return tune_pull(s);
    std::vector<hipEvent_t> ev(2 * n, nullptr);
    for (hipEvent_t& e : ev)
        if (hipEventCreate(&e) != hipSuccess) e = nullptr;
    for (hipEvent_t e : ev)
        if (!e) {
            for (hipEvent_t x : ev)
                if (x) (void)hipEventDestroy(x);
            return tune_pull(s);  // no timing: keep the default (and the in-place steps)
        }
    int rc = run_steps(s, TUNE_WARM, false);
    for (int i = 0; i < n && !rc; ++i) {
        s->step_tile = c[i].tile;
        s->step_lds[0] = c[i].lds_even;
        s->step_lds[1] = c[i].lds_odd;
        s->xcd_cells = c[i].xcd_cells;
        s->xcd_chunk = c[i].chunk;
        HIPCHK(hipEventRecord(ev[2 * i], s->stream));
        rc = run_steps(s, TUNE_STEPS, false);
        HIPCHK(hipEventRecord(ev[2 * i + 1], s->stream));
    }
    if (rc) {
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        return rc;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventSynchronize(ev[2 * n - 1]));
    int best = 0;
    float tbest = 0.0f, tdef = 0.0f;
    for (int i = 0; i < n; ++i) {
        float ms = 0.0f;
        HIPCHK(hipEventElapsedTime(&ms, ev[2 * i], ev[2 * i + 1]));
        if (i == 0) tdef = ms;
        if (i == 0 || ms < tbest) {
            best = i;
            tbest = ms;
        }
        if (s->tune_verbose)
            fprintf(stderr, "[jyulb] step launch tile %u, LDS %u/%u, %s order (chunk %u): %.4f ms per step\n",
                    c[i].tile, c[i].lds_even, c[i].lds_odd, c[i].xcd_cells ? "round-robin" : "XCD", c[i].chunk,
                    ms / TUNE_STEPS);
    }
    for (hipEvent_t e : ev) (void)hipEventDestroy(e);
    // a few timed steps cannot separate near-ties: the default stays unless
    // beaten by more than TUNE_MARGIN (on the 768x384x384 medium in FP16 the
    // two tile orders timed within 0.5% here but sustained 37.6k against 35.0k
    // MLUPS over 500 steps, in favour of the default)
    if (tbest > (1.0f - TUNE_MARGIN) * tdef) best = 0;
    s->step_tile = c[best].tile;
    s->step_lds[0] = c[best].lds_even;
    s->step_lds[1] = c[best].lds_odd;
    s->xcd_cells = c[best].xcd_cells;
    s->xcd_chunk = c[best].chunk;
    return tune_pull(s);
}

// The state before a call's last step, kept for the reference's f2: after
// evolve the reference's f2 holds the post-collision populations of the
// step before the last one (its two arrays are swapped every step,
// isothermal.pyx:1408-1415).
static int capture_prev(lb_solver* s) {
    const size_t fb = (size_t)s->fsize * ts_size(s->prec);
    if (s->prev && s->prev_bytes != fb) {
        dfree(s, &s->prev, s->prev_bytes);
        s->prev_bytes = 0;
    }
    int rc;
    if (!s->prev) {
        if ((rc = dalloc(s, &s->prev, fb))) return rc;
        s->prev_bytes = fb;
    }
    HIPCHK(stream_copy(s->prev, s->f, fb, s->stream));
    s->prev_odd = s->next_odd;
    s->prev_scale = s->scale;
    s->prev_valid = true;
    return LB_OK;
}

// The steps of one evolve call: the FP16 guard, the launch tuning of the
// first long call, then the steps (den/vel stored on the last one).
static int evolve_steps(lb_solver* s, uint32_t steps) {
    int rc = LB_OK;
    if (s->prec == LB_FP16_STORAGE && (s->f16_dirty || s->f16_since >= s->f16_check) && (rc = fp16_guard(s)))
        return rc;
    if (s->tune) {
        // tuning takes TUNE_WARM + TUNE_STEPS per candidate of this call's
        // steps; a shorter call leaves it to a later one
        StepCfg c[TUNE_MAX];
        const int n = tune_candidates(s, c);
        const int np = s->f2 && !s->pull_active ? TUNE_PULL_STEPS : 0;
        if (n < 2 && !np) {
            s->tune = false;
        } else if (steps > (uint32_t)(TUNE_WARM + n * TUNE_STEPS + np)) {
            const uint64_t before = s->steps;
            if ((rc = tune_step(s))) return rc;
            steps -= (uint32_t)(s->steps - before);
        }
    }
    if (s->keep_prev) {
        if ((rc = run_steps(s, steps - 1, false)) || (rc = capture_prev(s)) || (rc = run_steps(s, 1, true)))
            return rc;
    } else if ((rc = run_steps(s, steps, true))) {
        return rc;
    }
    HIPCHK(hipGetLastError());
    return LB_OK;
}

int lb_evolve(lb_solver* s, uint32_t steps) {
    if (!s) return fail(LB_EINVAL, "null handle");
    if (!s->finalized) return fail(LB_ESTATE, "Finalize solver initialization before calling evolve");
    state_advanced(s);  // an earlier adjoint forward pass no longer applies
    if (s->nfluid < 1 || steps == 0) return LB_OK;
    return evolve_steps(s, steps);
}

int lb_host_alloc(size_t bytes, void** ptr) {
    if (!ptr) return fail(LB_EINVAL, "null argument");
    *ptr = nullptr;
    if (bytes == 0) return LB_OK;
    hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
    if (e != hipSuccess) {
        *ptr = nullptr;
        return fail(LB_ENOMEM, "hipHostMalloc(%zu bytes) failed: %s", bytes, hipGetErrorString(e));
    }
    return LB_OK;
}

int lb_host_free(void* ptr) {
    if (ptr) HIPCHK(hipHostFree(ptr));
    return LB_OK;
}

int lb_bind_host_fields(lb_solver* s, uint8_t* phase, double* den, double* vel, double* acc) {
    if (!s) return fail(LB_EINVAL, "null handle");
    if (s->stream) HIPCHK(hipStreamSynchronize(s->stream));
    const int64_t ncnt = (int64_t)s->NI * s->NJ * s->NK;
    if (!phase && !den && !vel && !acc) {  // unbind
        s->h_phase = nullptr;
        s->h_den = s->h_vel = s->h_acc = nullptr;
        return LB_OK;
    }
    if (!phase || !den || !vel || !acc) return fail(LB_EINVAL, "lb_bind_host_fields: all four fields are needed");
    int rc;
    if (!s->st_phase) {
        if ((rc = dalloc(s, (void**)&s->st_phase, (size_t)ncnt)) ||
            (rc = dalloc(s, (void**)&s->st_den, (size_t)ncnt * 8)) ||
            (rc = dalloc(s, (void**)&s->st_vel, (size_t)ncnt * 24)) ||
            (rc = dalloc(s, (void**)&s->st_acc, (size_t)ncnt * 24)))
            return rc;
    }
    if (!s->acc && (rc = dalloc(s, &s->acc, (size_t)3 * s->nq * tc_size(s->prec)))) return rc;
    s->ncnt = ncnt;
    s->h_phase = phase;
    s->h_den = den;
    s->h_vel = vel;
    s->h_acc = acc;
    return LB_OK;
}

static bool has_halo_faces(const lb_solver* s) {
    for (int f = 0; f < 6; ++f)
        if (s->cfg.bc[f] == PERIODIC || s->cfg.bc[f] == SYMMETRIC) return true;
    return false;
}

// den/vel of the fluid sites into the staged host fields, optionally the
// den/vel halo pass there
static int scatter_staged(lb_solver* s, bool halo) {
    k_scatter_fluid<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->mask, s->den, s->vel, s->st_den, s->st_vel,
                                                                   s->d, s->nq);
    HIPCHK(hipGetLastError());
    if (halo) {
        const int64_t n[3] = {(int64_t)s->d.ny * s->d.nz, (int64_t)s->NI * s->d.nz, (int64_t)s->NI * s->NJ};
        for (int a = 0; a < 3; ++a) {
            if (s->cfg.bc[2 * a] != PERIODIC && s->cfg.bc[2 * a] != SYMMETRIC && s->cfg.bc[2 * a + 1] != PERIODIC &&
                s->cfg.bc[2 * a + 1] != SYMMETRIC)
                continue;
            k_halo_dv<<<(unsigned)((n[a] + 255) / 256), 256, 0, s->stream>>>(s->st_den, s->st_vel, s->d, a);
        }
        HIPCHK(hipGetLastError());
    }
    return LB_OK;
}

static int halo_only(lb_solver* s) {
    const int64_t n[3] = {(int64_t)s->d.ny * s->d.nz, (int64_t)s->NI * s->d.nz, (int64_t)s->NI * s->NJ};
    for (int a = 0; a < 3; ++a)
        k_halo_dv<<<(unsigned)((n[a] + 255) / 256), 256, 0, s->stream>>>(s->st_den, s->st_vel, s->d, a);
    HIPCHK(hipGetLastError());
    return LB_OK;
}

// The host fields as they are now (the user may have written any of them):
// staged on the device, checked against the finalized phase, and taken as
// the inputs of the next steps (k_sync_in).
static int load_host_fields(lb_solver* s) {
    const int64_t n = s->ncnt;
    HIPCHK(hipMemcpyAsync(s->st_phase, s->h_phase, (size_t)n, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->st_den, s->h_den, (size_t)n * 8, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->st_vel, s->h_vel, (size_t)n * 24, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->st_acc, s->h_acc, (size_t)n * 24, hipMemcpyHostToDevice, s->stream));
    unsigned int* flags = (unsigned int*)s->red;
    HIPCHK(hipMemsetAsync(flags, 0, 4, s->stream));
    if (tc_size(s->prec) == 8)
        k_sync_in<double><<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(
            s->st_phase, s->st_den, s->st_vel, s->st_acc, s->phase, s->den, s->vel, (double*)s->acc, s->d, s->nq,
            flags);
    else
        k_sync_in<float><<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(
            s->st_phase, s->st_den, s->st_vel, s->st_acc, s->phase, s->den, s->vel, (float*)s->acc, s->d, s->nq,
            flags);
    HIPCHK(hipGetLastError());
    unsigned int hf = 0;
    HIPCHK(hipMemcpyAsync(&hf, flags, 4, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (hf & SYNC_PHASE)
        return fail(LB_ESTATE, "the phase field was changed after finalize_init; the engine builds its links "
                    "from the phase at finalize_init, so phase edits after it are not supported");
    if (hf & SYNC_WALLS) {  // wall velocities changed: moving-wall flags and Ladd table
        int rc = build_mask(s, true);
        if (rc) return rc;
        int64_t nf, nm;
        if ((rc = count_mask(s, &nf, &nm))) return rc;
        s->moving = nm > 0;
    }
    s->acc_on = (hf & SYNC_ACC) != 0;
    return LB_OK;
}

int lb_load_host_fields(lb_solver* s) {
    if (!s) return fail(LB_EINVAL, "null handle");
    if (!s->finalized) return fail(LB_ESTATE, "solver not finalized");
    if (!s->h_phase) return fail(LB_ESTATE, "no host fields bound (lb_bind_host_fields)");
    inputs_changed(s);
    return load_host_fields(s);
}

int lb_evolve_host(lb_solver* s, uint32_t steps) {
    if (!s) return fail(LB_EINVAL, "null handle");
    if (!s->finalized) return fail(LB_ESTATE, "Finalize solver initialization before calling evolve");
    if (!s->h_phase) return fail(LB_ESTATE, "no host fields bound (lb_bind_host_fields)");
    inputs_changed(s);
    if (s->nfluid < 1 || steps == 0) return LB_OK;
    int rc = load_host_fields(s);
    if (rc) return rc;
    // Isothermal.evolve: the halo den/vel pass precedes every step, so after
    // the call the halos hold the state of step T-1 (isothermal.pyx:1408-1415)
    const bool halo = has_halo_faces(s);
    if (halo && steps >= 2) {
        if ((rc = evolve_steps(s, steps - 1))) return rc;
        if ((rc = scatter_staged(s, true))) return rc;
    } else if (halo && (rc = halo_only(s))) {
        return rc;
    }
    if ((rc = evolve_steps(s, halo ? 1 : steps))) return rc;
    if ((rc = scatter_staged(s, false))) return rc;
    if (s->prec == LB_FP16_STORAGE && (rc = fp16_guard(s))) return rc;  // no silent inf in the fields
    const int64_t n = s->ncnt;
    HIPCHK(hipMemcpyAsync(s->h_den, s->st_den, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(s->h_vel, s->st_vel, (size_t)n * 24, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

int lb_store_host_fields(lb_solver* s, int halo_pass) {
    if (!s) return fail(LB_EINVAL, "null handle");
    if (!s->h_phase) return fail(LB_ESTATE, "no host fields bound (lb_bind_host_fields)");
    const int64_t n = s->ncnt;
    // the SOLID sites and halos keep the host's values
    HIPCHK(hipMemcpyAsync(s->st_den, s->h_den, (size_t)n * 8, hipMemcpyHostToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->st_vel, s->h_vel, (size_t)n * 24, hipMemcpyHostToDevice, s->stream));
    int rc = scatter_staged(s, halo_pass != 0 && has_halo_faces(s));
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(s->h_den, s->st_den, (size_t)n * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(s->h_vel, s->st_vel, (size_t)n * 24, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

int lb_flow_info(lb_solver* s, lb_flow_info_t* info) {
    if (!s || !info) return fail(LB_EINVAL, "null argument");
    const int nb = (int)std::min<uint32_t>(RED_PARTS, blocks_for(s->nq, RED_BLOCK));
    SolverSites sites{s->mask, s->den, s->vel, s->nq};
    k_flow1<SolverSites><<<nb, RED_BLOCK, 0, s->stream>>>(sites, s->rho0, s->part);
    k_flow2<<<1, RED_BLOCK, 0, s->stream>>>(s->part, nb, s->rho0, s->red);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(info, s->red, sizeof(lb_flow_info_t), hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

// Address of `scratch` on the current device, resolved once per device with
// that device current (each device holds its own copy of the array).
static double* scratch_ptr(int* rc_out) {
    constexpr int MAX_DEV = 64;
    static double* ptrs[MAX_DEV] = {};
    static std::mutex mu;
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess || dev < 0 || dev >= MAX_DEV) {
        *rc_out = fail(LB_EHIP, "hipGetDevice failed or device %d beyond %d: %s", dev, MAX_DEV, hipGetErrorString(e));
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!ptrs[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *rc_out = fail(LB_EHIP, "hipGetSymbolAddress(scratch) failed on device %d: %s", dev, hipGetErrorString(e));
            return nullptr;
        }
        ptrs[dev] = (double*)p;
    }
    *rc_out = LB_OK;
    return ptrs[dev];
}

int lb_flow_info_fields_noalloc(const uint8_t* phase_dev, const double* den_dev, const double* vel_dev,
                                const int32_t lat[3], double rho0, double* info_dev, void* hip_stream) {
    if (!phase_dev || !den_dev || !vel_dev || !lat || !info_dev) return fail(LB_EINVAL, "null argument");
    if (lat[0] < 1 || lat[1] < 1 || lat[2] < 1) return fail(LB_EINVAL, "Invalid lattice size input.");
    const uint64_t n = (uint64_t)lat[0] * (uint64_t)lat[1] * (uint64_t)lat[2];
    if (n >= (1ull << 31))
        return fail(LB_EINVAL, "lb_flow_info_fields_noalloc: %llu sites exceed its limit of 2^31 - 1",
                    (unsigned long long)n);
    int rc;
    double* part = scratch_ptr(&rc);
    if (!part) return rc;
    const hipStream_t st = (hipStream_t)hip_stream;
    const int nb = (int)std::min<uint64_t>(RED_PARTS, (n + RED_BLOCK - 1) / RED_BLOCK);
    PlainSites sites{phase_dev, den_dev, vel_dev, (uint32_t)n};
    k_flow1<PlainSites><<<nb, RED_BLOCK, 0, st>>>(sites, rho0, part);
    k_flow2<<<1, RED_BLOCK, 0, st>>>(part, nb, rho0, info_dev);
    HIPCHK(hipGetLastError());
    return LB_OK;
}

int lb_download_den_vel(lb_solver* s, double* den, double* vel) {
    if (!s || !den || !vel) return fail(LB_EINVAL, "null argument");
    int rc = ensure_stage(s);
    if (rc) return rc;
    k_export_dv<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->stage, s->den, s->vel, s->d, s->nq);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(den, s->stage, (size_t)s->ncell * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipMemcpyAsync(vel, s->stage + s->ncell, (size_t)s->ncell * 24, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

int lb_download_den_vel_device(lb_solver* s, double* den_dev, double* vel_dev) {
    if (!s || !den_dev || !vel_dev) return fail(LB_EINVAL, "null argument");
    int rc = ensure_stage(s);
    if (rc) return rc;
    k_export_dv<<<blocks_for(s->nq, 256), 256, 0, s->stream>>>(s->stage, s->den, s->vel, s->d, s->nq);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(den_dev, s->stage, (size_t)s->ncell * 8, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(vel_dev, s->stage + s->ncell, (size_t)s->ncell * 24, hipMemcpyDeviceToDevice, s->stream));
    return LB_OK;
}

// Post-collision populations of the state in `fsrc` (parity next_odd, FP16
// scale `scale`) unpacked to f64 [19][nx][ny][nz] in host memory `f`.
static int download_post_f(lb_solver* s, void* fsrc, int next_odd, double scale, double* f) {
    double* buf = nullptr;
    int rc = dalloc(s, (void**)&buf, (size_t)Q * s->ncell * 8);
    if (rc) return rc;
    StepParams p = params(s);
    p.f = fsrc;
    p.scale = (float)scale;
    p.inv_scale = (float)(1.0 / scale);
    switch (s->prec) {
    case LB_FP64: launch_post_f<LB_FP64>(p, next_odd, buf, s->stream); break;
    case LB_FP32_SHIFTED: launch_post_f<LB_FP32_SHIFTED>(p, next_odd, buf, s->stream); break;
    case LB_FP32_STORAGE: launch_post_f<LB_FP32_STORAGE>(p, next_odd, buf, s->stream); break;
    default: launch_post_f<LB_FP16_STORAGE>(p, next_odd, buf, s->stream); break;
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(f, buf, (size_t)Q * s->ncell * 8, hipMemcpyDeviceToHost, s->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
    (void)hipFree(buf);
    s->bytes -= (size_t)Q * s->ncell * 8;
    if (e != hipSuccess) return fail(LB_EHIP, "lb_download_post_f: %s", hipGetErrorString(e));
    return LB_OK;
}

int lb_download_post_f(lb_solver* s, double* f) {
    if (!s || !f) return fail(LB_EINVAL, "null argument");
    if (!s->finalized) return fail(LB_ESTATE, "solver not finalized");
    return download_post_f(s, s->f, s->next_odd, s->scale, f);
}

int lb_keep_previous(lb_solver* s, int on) {
    if (!s) return fail(LB_EINVAL, "null handle");
    s->keep_prev = on != 0;
    if (!s->keep_prev) {
        dfree(s, &s->prev, s->prev_bytes);
        s->prev_bytes = 0;
        s->prev_valid = false;
    }
    return LB_OK;
}

int lb_download_previous_post_f(lb_solver* s, double* f) {
    if (!s || !f) return fail(LB_EINVAL, "null argument");
    if (!s->finalized) return fail(LB_ESTATE, "solver not finalized");
    if (!s->prev_valid)
        return fail(LB_ESTATE, "no previous step kept: lb_keep_previous before an evolve call");
    return download_post_f(s, s->prev, s->prev_odd, s->prev_scale, f);
}

int lb_reduce(lb_solver* s, double out[8]) {
    if (!s || !out) return fail(LB_EINVAL, "null argument");
    const int nb = (int)std::min<uint32_t>(RED_PARTS, blocks_for(s->nq, RED_BLOCK));
    k_reduce1<<<nb, RED_BLOCK, 0, s->stream>>>(s->mask, s->den, s->vel, s->nq, s->part);
    k_reduce2<<<1, RED_BLOCK, 0, s->stream>>>(s->part, nb, s->red);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, s->red, RED_N * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

int lb_convergence(lb_solver* s, double* rel) {
    if (!s || !rel) return fail(LB_EINVAL, "null argument");
    if (!s->uprev) {
        int rc = dalloc(s, (void**)&s->uprev, (size_t)3 * s->nq * 8);
        if (rc) return rc;
    }
    const int nb = (int)std::min<uint32_t>(RED_PARTS, blocks_for(s->nq, RED_BLOCK));
    k_conv1<<<nb, RED_BLOCK, 0, s->stream>>>(s->mask, s->vel, s->uprev, s->nq, s->part);
    k_reduce2<<<1, RED_BLOCK, 0, s->stream>>>(s->part, nb, s->red);
    HIPCHK(hipGetLastError());
    double h[RED_N];
    HIPCHK(hipMemcpyAsync(h, s->red, RED_N * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    if (!s->have_uprev) {
        s->have_uprev = true;
        *rel = 1.0;
    } else {
        *rel = h[1] > 0 ? sqrt(h[0] / h[1]) : (h[0] > 0 ? INFINITY : 0.0);
    }
    return LB_OK;
}

static int objective(lb_solver* s, int face, int axis, double* J, double* count) {
    const int nb = 256;
    k_obj<<<nb, RED_BLOCK_ADJ, 0, s->stream>>>(s->mask, s->den, s->vel, s->d, s->nq, face, axis, s->part);
    HIPCHK(hipGetLastError());
    std::vector<double> h(2 * nb);
    HIPCHK(hipMemcpyAsync(h.data(), s->part, 2 * nb * 8, hipMemcpyDeviceToHost, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    double a = 0.0, n = 0.0;
    for (int b = 0; b < nb; ++b) {
        a += h[2 * b];
        n += h[2 * b + 1];
    }
    if (count) *count = n;
    *J = n > 0 ? a / n : 0.0;
    return LB_OK;
}

int lb_objective(lb_solver* s, int face, int axis, double* J) {
    if (!s || !J || face < 0 || face > 5 || axis < 0 || axis > 2) return fail(LB_EINVAL, "invalid argument");
    return objective(s, face, axis, J, nullptr);
}

// The (rho, u*) record of one checkpoint segment: `len` steps of 4 rows.
static int ensure_rec(lb_solver* s, uint32_t len) {
    const size_t recb = (size_t)len * 4 * rec_row(s) * tc_size(s->prec);
    if (s->rec_bytes >= recb) return LB_OK;
    s->rec_fwd = false;
    if (s->rec) {
        HIPCHK(hipFree(s->rec));
        s->bytes -= s->rec_bytes;
        s->rec = nullptr;
        s->rec_bytes = 0;
    }
    int rc = dalloc(s, &s->rec, recb);
    if (rc) return rc;
    s->rec_bytes = recb;
    return LB_OK;
}

int lb_adj_forward(lb_solver* s, uint32_t steps, uint32_t interval, int face, int axis, double* J) {
    const bool diss = face == LB_OBJ_DISSIPATION;
    if (!s || !J || face < 0 || face > LB_OBJ_DISSIPATION || (!diss && (axis < 0 || axis > 2)))
        return fail(LB_EINVAL, "invalid argument");
    if (!s->finalized) return fail(LB_ESTATE, "solver not finalized");
    if (steps < 1 || interval < 1) return fail(LB_EINVAL, "steps and checkpoint interval must be >= 1");
    if (s->prec != LB_FP64 && s->prec != LB_FP32_SHIFTED)
        return fail(LB_EINVAL, "the adjoint runs in FP64 or FP32-shifted precision");
    if (s->moving) return fail(LB_EINVAL, "the adjoint supports stationary walls only (no moving solid sites)");
    const uint32_t nseg = (steps + interval - 1) / interval;
    const size_t fbytes = (size_t)s->fsize * ts_size(s->prec);
    int rc;
    s->adj_ready = false;  // until this pass completes: the checkpoints are overwritten
    s->rec_fwd = false;
    if (s->ckpt_bytes < (size_t)(nseg + 1) * fbytes) {
        if (s->ckpt) {
            HIPCHK(hipFree(s->ckpt));
            s->bytes -= s->ckpt_bytes;
            s->ckpt = nullptr;
        }
        if ((rc = dalloc(s, &s->ckpt, (size_t)(nseg + 1) * fbytes))) return rc;
        s->ckpt_bytes = (size_t)(nseg + 1) * fbytes;
    }
    const size_t tcs = tc_size(s->prec);
    if ((rc = ensure_rec(s, std::min(interval, steps)))) return rc;
    const size_t drows = (size_t)(Q + 4) * s->nq;
    if (diss) {
        if (!s->dseed && (rc = dalloc(s, &s->dseed, drows * tcs))) return rc;
        const unsigned fb = 256;
        if (s->prec == LB_FP64)
            k_fill<double><<<(unsigned)((drows + fb - 1) / fb), fb, 0, s->stream>>>((double*)s->dseed, drows, 0.0);
        else k_fill<float><<<(unsigned)((drows + fb - 1) / fb), fb, 0, s->stream>>>((float*)s->dseed, drows, 0.0f);
    }
    s->adj_po = s->next_odd;
    s->ckpt_odd.assign(nseg + 1, 1);
    StepParams p = params(s);
    for (uint32_t seg = 0; seg < nseg; ++seg) {
        HIPCHK(stream_copy((char*)s->ckpt + (size_t)seg * fbytes, s->f, fbytes, s->stream));
        s->ckpt_odd[seg] = (uint8_t)s->next_odd;
        const uint32_t len = std::min(interval, steps - seg * interval);
        for (uint32_t t = 0; t < len; ++t) {
            p.store_dv = (seg * interval + t + 1 == steps) ? 1 : 0;
            // the last segment records (rho, u*) as the reverse pass's recompute would
            p.rec = seg + 1 == nseg ? (char*)s->rec + (size_t)t * 4 * rec_row(s) * tcs : nullptr;
            if (diss && p.store_dv) {  // the objective is on the final step's collision input
                StepParams q = p;
                q.rec = nullptr;
                q.dseed = s->dseed;
                q.dscale = 1.0 / (double)s->nfluid;
                if (s->prec == LB_FP64) launch_diss_rlx<LB_FP64>(q, s->cfg.rlx_scheme, s->next_odd != 0, s->stream);
                else launch_diss_rlx<LB_FP32_SHIFTED>(q, s->cfg.rlx_scheme, s->next_odd != 0, s->stream);
            }
            one_step(s, p, s->next_odd);
            s->steps++;
        }
    }
    HIPCHK(hipGetLastError());
    HIPCHK(stream_copy((char*)s->ckpt + (size_t)nseg * fbytes, s->f, fbytes, s->stream));
    s->ckpt_odd[nseg] = (uint8_t)s->next_odd;
    double n = 0.0;
    if (diss) {
        const int nb = 256;
        const size_t off = (size_t)(Q + 3) * s->nq;
        if (s->prec == LB_FP64)
            k_sum_row<double><<<nb, 256, 0, s->stream>>>((const double*)s->dseed + off, s->nq, s->part);
        else k_sum_row<float><<<nb, 256, 0, s->stream>>>((const float*)s->dseed + off, s->nq, s->part);
        HIPCHK(hipGetLastError());
        std::vector<double> h(nb);
        HIPCHK(hipMemcpyAsync(h.data(), s->part, nb * 8, hipMemcpyDeviceToHost, s->stream));
        HIPCHK(hipStreamSynchronize(s->stream));
        double a = 0.0;
        for (double v : h) a += v;
        n = (double)s->nfluid;
        *J = a / n;
    } else if ((rc = objective(s, face, axis, J, &n))) {
        return rc;
    }
    if (n <= 0) return fail(LB_EINVAL, "objective face %d has no fluid site", face);
    s->adj_steps = steps;
    s->adj_interval = interval;
    s->adj_step_end = s->steps;
    s->adj_face = face;
    s->adj_axis = axis;
    s->adj_scale = 1.0 / n;
    s->adj_ready = true;
    s->rec_fwd = true;
    return LB_OK;
}

static int adj_backward(lb_solver* s, int gface_id, double* grad_acc, double* grad_face, bool device) {
    if (!s) return fail(LB_EINVAL, "null handle");
    if (!s->adj_ready)
        return fail(LB_ESTATE, "lb_adj_backward needs the lb_adj_forward pass of the current state and inputs: "
                    "an evolve step, an input upload or a restore since the forward pass invalidates it");
    if (s->steps != s->adj_step_end)
        return fail(LB_ESTATE, "the state advanced since lb_adj_forward (step %llu, forward pass ended at %llu)",
                    (unsigned long long)s->steps, (unsigned long long)s->adj_step_end);
    if (gface_id >= 0 && (gface_id > 5 || !s->fden[gface_id]))
        return fail(LB_EINVAL, "face %d has no FIXED_DEN/FIXED_VEL condition", gface_id);
    const size_t tcs = tc_size(s->prec);
    const size_t fbytes = (size_t)s->fsize * ts_size(s->prec);
    int rc;
    if (!s->lam && (rc = dalloc(s, &s->lam, (size_t)s->fsize * tcs))) return rc;
    // Pull-form adjoint steps from 1 GiB of cotangents, like the sparse
    // layout's forward pull steps: on the 768x384x384 medium the reverse pass
    // took 5.26 ms per step in FP32 (10.39 in FP64) against 5.58 (10.99) in
    // place, both with the record still indexed by cell; on the 128x64x64
    // channel, whose arrays stay in the caches, the in-place steps are faster
    // (22.5 against 24.3 us per adjoint step).
    const size_t lbytes = (size_t)s->fsize * tcs;
    const bool want_lpull = s->adj_pull >= 0 ? s->adj_pull != 0 : lbytes >= ((size_t)1 << 30);
    if (want_lpull && !s->lam2) dalloc_optional(s, &s->lam2, lbytes);
    const bool lpull = want_lpull && s->lam2;
    const size_t rn = rec_row(s);
    if (!s->gacc && (rc = dalloc(s, &s->gacc, 3 * rn * tcs))) return rc;
    const size_t nface = (size_t)(s->NI > s->NJ ? s->NI : s->NJ) * (s->NJ > s->NK ? s->NJ : s->NK) * 3;
    if (!s->gface && (rc = dalloc(s, (void**)&s->gface, std::max<size_t>(nface, 3 * (size_t)s->NI * s->NJ) * 8)))
        return rc;
    if ((rc = ensure_rec(s, std::min(s->adj_interval, s->adj_steps)))) return rc;
    const size_t fface = gface_id < 0 ? 0
                         : (size_t)(gface_id < 2 ? s->NJ * s->NK : gface_id < 4 ? s->NI * s->NK : s->NI * s->NJ);
    const unsigned fb = 256;
    if (s->prec == LB_FP64) {
        k_fill<double><<<(unsigned)((s->fsize + fb - 1) / fb), fb, 0, s->stream>>>((double*)s->lam, s->fsize, 0.0);
        if (lpull)
            k_fill<double><<<(unsigned)((s->fsize + fb - 1) / fb), fb, 0, s->stream>>>((double*)s->lam2, s->fsize, 0.0);
        k_fill<double><<<(unsigned)((3 * rn + fb - 1) / fb), fb, 0, s->stream>>>((double*)s->gacc, 3 * rn, 0.0);
    } else {
        k_fill<float><<<(unsigned)((s->fsize + fb - 1) / fb), fb, 0, s->stream>>>((float*)s->lam, s->fsize, 0.0f);
        if (lpull)
            k_fill<float><<<(unsigned)((s->fsize + fb - 1) / fb), fb, 0, s->stream>>>((float*)s->lam2, s->fsize, 0.0f);
        k_fill<float><<<(unsigned)((3 * rn + fb - 1) / fb), fb, 0, s->stream>>>((float*)s->gacc, 3 * rn, 0.0f);
    }
    if (fface) k_fill<double><<<(unsigned)((3 * fface + fb - 1) / fb), fb, 0, s->stream>>>(s->gface, 3 * fface, 0.0);
    HIPCHK(hipGetLastError());
    const uint32_t steps = s->adj_steps, interval = s->adj_interval;
    const uint32_t nseg = (steps + interval - 1) / interval;
    StepParams p = params(s);
    p.store_dv = 0;
    AdjParams x;
    x.lam = s->lam;
    x.lam2 = nullptr;
    x.gacc = s->gacc;
    x.gface = fface ? s->gface : nullptr;
    x.gface_id = gface_id;
    x.obj_face = s->adj_face;
    x.obj_axis = s->adj_axis;
    x.obj_scale = s->adj_scale;
    x.dseed = s->dseed;
    for (int seg = (int)nseg - 1; seg >= 0; --seg) {
        const uint32_t g0 = (uint32_t)seg * interval;
        const uint32_t len = std::min(interval, steps - g0);
        if (seg + 1 == (int)nseg && s->rec_fwd) {
            // recorded by the forward pass; the populations are the segment's end state
            HIPCHK(stream_copy(s->f, (char*)s->ckpt + (size_t)nseg * fbytes, fbytes, s->stream));
        } else {
            s->rec_fwd = false;  // rec now holds an earlier segment
            HIPCHK(stream_copy(s->f, (char*)s->ckpt + (size_t)seg * fbytes, fbytes, s->stream));
            int odd = s->ckpt_odd[seg];  // as the forward pass stepped from this checkpoint
            for (uint32_t t = 0; t < len; ++t) {  // recompute, recording (rho, u*)
                p.rec = (char*)s->rec + (size_t)t * 4 * rec_row(s) * tcs;
                one_step(s, p, odd);
            }
            p.rec = nullptr;
        }
        // The adjoint steps need only the record, whichever steps (in place
        // or pull) the forward pass took: in place, the adjoint of step g is
        // the adjoint AA step of g's parity in the in-place sequence from
        // adj_po; in pull form, one kernel, alternating lam and lam2.
        for (int t = (int)len - 1; t >= 0; --t) {
            const uint32_t g = g0 + (uint32_t)t;
            x.rec = (const char*)s->rec + (size_t)t * 4 * rec_row(s) * tcs;
            x.seed = (g + 1 == steps) ? 1 : 0;
            const bool odd = (s->adj_po ^ (g & 1)) != 0;
            if (lpull) {
                x.lam = s->lam;
                x.lam2 = s->lam2;
            }
            if (s->prec == LB_FP64) launch_adj_rlx<LB_FP64>(p, x, s->cfg.rlx_scheme, odd, s->stream);
            else launch_adj_rlx<LB_FP32_SHIFTED>(p, x, s->cfg.rlx_scheme, odd, s->stream);
            if (lpull) std::swap(s->lam, s->lam2);
        }
        HIPCHK(hipGetLastError());
    }
    // restore the state after the forward pass
    HIPCHK(stream_copy(s->f, (char*)s->ckpt + (size_t)nseg * fbytes, fbytes, s->stream));
    if ((rc = ensure_stage(s))) return rc;
    if (grad_acc) {
        if (s->sparse) {  // fluid sites from the compact rows, zero elsewhere
            const size_t no = (size_t)s->ncell * 3;
            k_fill<double><<<(unsigned)((no + 255) / 256), 256, 0, s->stream>>>(s->stage, no, 0.0);
            const uint32_t nf = (uint32_t)s->nfluid;
            if (nf && s->prec == LB_FP64)
                k_export_vec3_sp<double><<<blocks_for(nf, 256), 256, 0, s->stream>>>((const double*)s->gacc, s->site, nf,
                                                                                  s->nfp, s->stage, s->d);
            else if (nf)
                k_export_vec3_sp<float><<<blocks_for(nf, 256), 256, 0, s->stream>>>((const float*)s->gacc, s->site, nf,
                                                                                 s->nfp, s->stage, s->d);
        } else if (s->prec == LB_FP64) {
            k_export_vec3<double><<<blocks_for(s->nq, 256), 256, 0, s->stream>>>((const double*)s->gacc, s->stage, s->d, s->nq);
        } else {
            k_export_vec3<float><<<blocks_for(s->nq, 256), 256, 0, s->stream>>>((const float*)s->gacc, s->stage, s->d, s->nq);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(grad_acc, s->stage, (size_t)s->ncell * 24,
                              device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost, s->stream));
    }
    if (grad_face && fface)
        HIPCHK(hipMemcpyAsync(grad_face, s->gface, fface * 24, device ? hipMemcpyDeviceToDevice : hipMemcpyDeviceToHost,
                              s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

int lb_adj_backward(lb_solver* s, int face, double* grad_acc, double* grad_face) {
    return adj_backward(s, face, grad_acc, grad_face, false);
}

int lb_adj_backward_device(lb_solver* s, int face, double* grad_acc_dev, double* grad_face_dev) {
    return adj_backward(s, face, grad_acc_dev, grad_face_dev, true);
}

int lb_set_face_data_device(lb_solver* s, int face, const double* den_dev, const double* vel_dev) {
    if (!s || face < 0 || face > 5 || !den_dev || !vel_dev) return fail(LB_EINVAL, "invalid argument");
    inputs_changed(s);
    if (!s->fden[face]) return fail(LB_EINVAL, "face %d has no FIXED_DEN/FIXED_VEL condition", face);
    const size_t n = (size_t)(face < 2 ? s->NJ * s->NK : face < 4 ? s->NI * s->NK : s->NI * s->NJ);
    HIPCHK(hipMemcpyAsync(s->fden[face], den_dev, n * 8, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->fvel[face], vel_dev, 3 * n * 8, hipMemcpyDeviceToDevice, s->stream));
    return LB_OK;
}

int lb_snapshot(lb_solver* s) {
    if (!s) return fail(LB_EINVAL, "null handle");
    if (!s->finalized) return fail(LB_ESTATE, "solver not finalized");
    const size_t fbytes = (size_t)s->fsize * ts_size(s->prec);
    int rc;
    if (!s->snap && (rc = dalloc(s, &s->snap, fbytes))) return rc;
    if (!s->snap_dv && (rc = dalloc(s, (void**)&s->snap_dv, (size_t)4 * s->nq * 8))) return rc;
    HIPCHK(stream_copy(s->snap, s->f, fbytes, s->stream));
    HIPCHK(hipMemcpyAsync(s->snap_dv, s->den, (size_t)s->nq * 8, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->snap_dv + s->nq, s->vel, (size_t)3 * s->nq * 8, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->snap_steps = s->steps;
    s->snap_odd = s->next_odd;
    return LB_OK;
}

int lb_restore(lb_solver* s) {
    if (!s) return fail(LB_EINVAL, "null handle");
    inputs_changed(s);
    if (!s->snap) return fail(LB_ESTATE, "no snapshot taken");
    const size_t fbytes = (size_t)s->fsize * ts_size(s->prec);
    HIPCHK(stream_copy(s->f, s->snap, fbytes, s->stream));
    HIPCHK(hipMemcpyAsync(s->den, s->snap_dv, (size_t)s->nq * 8, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipMemcpyAsync(s->vel, s->snap_dv + s->nq, (size_t)3 * s->nq * 8, hipMemcpyDeviceToDevice, s->stream));
    HIPCHK(hipStreamSynchronize(s->stream));
    s->steps = s->snap_steps;
    s->next_odd = s->snap_odd;
    s->adj_ready = false;
    return LB_OK;
}

int lb_info(lb_solver* s, uint64_t* steps, int* next_odd, uint64_t* bytes, double* scale) {
    if (!s) return fail(LB_EINVAL, "null handle");
    if (steps) *steps = s->steps;
    if (next_odd) *next_odd = s->next_odd;
    if (bytes) *bytes = s->bytes;
    if (scale) *scale = s->scale;
    return LB_OK;
}

int lb_sparse_codes(lb_solver* s, uint64_t* counts) {
    if (!s || !counts) return fail(LB_EINVAL, "null argument");
    if (!s->finalized || s->layout != LB_LAYOUT_SPARSE)
        return fail(LB_ESTATE, "lb_sparse_codes needs a finalized solver in the sparse layout");
    for (int k = 0; k < 4; ++k) counts[k] = s->sp_kinds[k];
    return LB_OK;
}

int lb_pull_steps(lb_solver* s, int* on) {
    if (!s || !on) return fail(LB_EINVAL, "null argument");
    *on = s->pull_active ? 1 : 0;
    return LB_OK;
}

int lb_layout(lb_solver* s, int* layout) {
    if (!s || !layout) return fail(LB_EINVAL, "null argument");
    *layout = s->layout;
    return LB_OK;
}

int lb_stream_copy(void* dst, const void* src, size_t bytes, void* hip_stream) {
    if (!dst || !src) return fail(LB_EINVAL, "null argument");
    if (bytes == 0) return LB_OK;
    HIPCHK(stream_copy(dst, src, bytes, (hipStream_t)hip_stream));
    return LB_OK;
}

int lb_synchronize(lb_solver* s) {
    if (!s) return fail(LB_EINVAL, "null handle");
    HIPCHK(hipStreamSynchronize(s->stream));
    return LB_OK;
}

}  // extern "C"

