// Per-site D3Q19 physics for the fused stream+collide kernel: moment maps,
// BGK / TRT / regularized collision with Guo-type forcing in moment space,
// and the Zou-He fixed-density / fixed-velocity faces.
//
// T is the compute type. SH selects the shifted-population form used by the
// reduced-precision kernels: the stored populations are f~ = f - w*rho0 and
// every formula is rewritten so that only O(u, drho) quantities are stored
// and summed. With T = double and SH = false every expression below keeps
// the reference's operation order term by term (D3Q19.pyx:348-940,
// bcs.pyx:31-716), so the FP64 kernel, compiled without FMA contraction,
// reproduces the reference bit for bit.
#pragma once
#include "lb_lattice.h"

namespace lb {

#define LBD __device__ __forceinline__

// f_to_den_vel, D3Q19.pyx:348-390. Returns sum f (= den, or drho when SH);
// `den` receives the full density.
template <class T, bool SH>
LBD T moments_den_vel(const T* f, T rho0, T& den, T* vel) {
    T den_1x = f[0] + f[1] + f[2] + f[3] + f[4];
    T den_x1 = f[14] + f[15] + f[16] + f[17] + f[18];
    T den_x0_1y = f[5] + f[6] + f[7];
    T den_x0_y1 = f[11] + f[12] + f[13];
    T s = den_1x + den_x1 + den_x0_1y + den_x0_y1 + f[8] + f[9] + f[10];
    T jx = den_x1 - den_1x;
    T jy = f[4] + den_x0_y1 + f[18] - f[0] - den_x0_1y - f[14];
    T jz = f[3] + f[7] + f[10] + f[13] + f[17] - f[1] - f[5] - f[8] - f[11] - f[15];
    den = SH ? rho0 + s : s;
    T inv_den = T(1.0) / den;
    vel[0] = inv_den * jx;
    vel[1] = inv_den * jy;
    vel[2] = inv_den * jz;
    return s;
}

// f_to_h, D3Q19.pyx:395-452 (h[0] is drho when SH; h2 is the same either way)
template <class T>
LBD void moments_h(const T* f, T* h) {
    T den_1x = f[0] + f[1] + f[2] + f[3] + f[4];
    T den_x1 = f[14] + f[15] + f[16] + f[17] + f[18];
    T den_x0_1y = f[5] + f[6] + f[7];
    T den_x0_y1 = f[11] + f[12] + f[13];
    T den_1y = f[0] + den_x0_1y + f[14];
    T den_y1 = f[4] + den_x0_y1 + f[18];
    T den_1z = f[1] + f[5] + f[8] + f[11] + f[15];
    T den_z1 = f[3] + f[7] + f[10] + f[13] + f[17];
    T den = den_1x + den_x1 + den_x0_1y + den_x0_y1 + f[8] + f[9] + f[10];
    const T INV_AS2 = T(1.0) / T(3.0);
    h[0] = den;
    h[1] = den_x1 - den_1x;
    h[2] = den_y1 - den_1y;
    h[3] = den_z1 - den_1z;
    h[4] = (den_1x + den_x1) - INV_AS2 * den;
    h[5] = f[0] + f[18] - f[4] - f[14];
    h[6] = f[1] + f[17] - f[3] - f[15];
    h[7] = (den_1y + den_y1) - INV_AS2 * den;
    h[8] = f[5] + f[13] - f[7] - f[11];
    h[9] = (den_1z + den_z1) - INV_AS2 * den;
}

// Kinetic projection P(h): h_to_f (D3Q19.pyx:460-527) when ADD is false,
// add_h_to_f (D3Q19.pyx:532-599) when true. DIV corrects the products with
// the rounded reciprocals of 72 and 3 to the quotients.
template <class T, bool ADD, bool DIV = false>
LBD void project(const T* h, T* f) {
    T e_xy = T(2) * h[0] + T(6) * h[4] + T(6) * h[7] - T(3) * h[9];
    T e_1x_1y = e_xy + T(18) * h[5];
    T o_1x_1y = T(-6) * h[1] - T(6) * h[2];
    T e_1x_y1 = e_xy - T(18) * h[5];
    T o_1x_y1 = T(-6) * h[1] + T(6) * h[2];
    T e_xz = T(2) * h[0] + T(6) * h[4] - T(3) * h[7] + T(6) * h[9];
    T e_1x_1z = e_xz + T(18) * h[6];
    T o_1x_1z = T(-6) * h[1] - T(6) * h[3];
    T e_1x_z1 = e_xz - T(18) * h[6];
    T o_1x_z1 = T(-6) * h[1] + T(6) * h[3];
    T e_yz = T(2) * h[0] - T(3) * h[4] + T(6) * h[7] + T(6) * h[9];
    T e_1y_1z = e_yz + T(18) * h[8];
    T o_1y_1z = T(-6) * h[2] - T(6) * h[3];
    T e_1y_z1 = e_yz - T(18) * h[8];
    T o_1y_z1 = T(-6) * h[2] + T(6) * h[3];
    T e_1x = T(4) * h[0] + T(12) * h[4] - T(6) * h[7] - T(6) * h[9];
    T e_1y = T(4) * h[0] - T(6) * h[4] + T(12) * h[7] - T(6) * h[9];
    T e_1z = T(4) * h[0] - T(6) * h[4] - T(6) * h[7] + T(12) * h[9];
    // In FP32 the rounded 1/72 and 1/3 are 7.5e-9 and 3e-8 too large, a bias
    // that every step applies to every moment. DIV removes it: the product
    // with the rounded reciprocal is corrected by its remainder v - q d, which
    // an FMA computes exactly (one Newton step on the quotient). FP64 keeps
    // the reference's multiplications.
    auto quo = [](T v, T d) -> T {
        const T r = T(1.0) / d;
        if constexpr (!DIV) {
            return r * v;
        } else {
            static_assert(sizeof(T) == 4, "DIV is an FP32 form");
            const float q = v * r;
            return __builtin_fmaf(__builtin_fmaf(-q, d, v), r, q);
        }
    };
    auto c72 = [&](T v) { return quo(v, T(72.0)); };
    T p[19];
    p[0] = c72(e_1x_1y + o_1x_1y);
    p[18] = c72(e_1x_1y - o_1x_1y);
    p[1] = c72(e_1x_1z + o_1x_1z);
    p[17] = c72(e_1x_1z - o_1x_1z);
    p[2] = c72(e_1x - T(12) * h[1]);
    p[16] = c72(e_1x + T(12) * h[1]);
    p[3] = c72(e_1x_z1 + o_1x_z1);
    p[15] = c72(e_1x_z1 - o_1x_z1);
    p[4] = c72(e_1x_y1 + o_1x_y1);
    p[14] = c72(e_1x_y1 - o_1x_y1);
    p[5] = c72(e_1y_1z + o_1y_1z);
    p[13] = c72(e_1y_1z - o_1y_1z);
    p[6] = c72(e_1y - T(12) * h[2]);
    p[12] = c72(e_1y + T(12) * h[2]);
    p[7] = c72(e_1y_z1 + o_1y_z1);
    p[11] = c72(e_1y_z1 - o_1y_z1);
    p[8] = c72(e_1z - T(12) * h[3]);
    p[10] = c72(e_1z + T(12) * h[3]);
    p[9] = quo(h[0], T(3.0)) - T(0.5) * (h[4] + h[7] + h[9]);
#pragma unroll
    for (int l = 0; l < Q; ++l) f[l] = ADD ? f[l] + p[l] : p[l];
}

// den_vel_to_heq, D3Q19.pyx:604-630
template <class T>
LBD void heq(T den, const T* vel, T* h) {
    T jx = den * vel[0], jy = den * vel[1], jz = den * vel[2];
    h[0] = den;
    h[1] = jx;
    h[2] = jy;
    h[3] = jz;
    h[4] = jx * vel[0];
    h[5] = jx * vel[1];
    h[6] = jx * vel[2];
    h[7] = jy * vel[1];
    h[8] = jy * vel[2];
    h[9] = jz * vel[2];
}

// forcing moments: D3Q19.pyx:717-729 (BGK), 812-824 (TRT), 927-935 (regul.)
template <class T>
LBD void forcing(T* h, const T* acc, T den, T jx, T jy, T jz, T c1, T c2) {
    h[1] += c1 * acc[0] * den;
    h[2] += c1 * acc[1] * den;
    h[3] += c1 * acc[2] * den;
    h[4] += T(2.0) * c2 * acc[0] * jx;
    h[5] += c2 * (acc[0] * jy + acc[1] * jx);
    h[6] += c2 * (acc[0] * jz + acc[2] * jx);
    h[7] += T(2.0) * c2 * acc[1] * jy;
    h[8] += c2 * (acc[1] * jz + acc[2] * jy);
    h[9] += T(2.0) * c2 * acc[2] * jz;
}

// forw_euler_BGK, D3Q19.pyx:655-740. prm[0] = 1/tau. Returns den; vel gets
// the half-force shifted velocity, ds the density as the populations hold it
// (den, or den - rho0 when SH: the caller can add rho0 back in FP64, where
// den itself carries only the compute type's precision near rho0).
template <class T, bool SH, bool DIV = false>
LBD T collide_bgk(T* f, const T* acc, const T* prm, T rho0, T* vel, T& ds) {
    T inv_tau = prm[0], rlx_cff1 = T(1.0) - inv_tau, rlx_cff2 = (T(1.0) - T(0.5) * inv_tau);
    T h[10], den;
    T s = moments_den_vel<T, SH>(f, rho0, den, vel);
    ds = s;
    vel[0] += T(0.5) * acc[0];
    vel[1] += T(0.5) * acc[1];
    vel[2] += T(0.5) * acc[2];
    heq<T>(den, vel, h);
    if (SH) h[0] = s;
#pragma unroll
    for (int m = 0; m < 10; ++m) h[m] *= inv_tau;
    T jx = den * vel[0], jy = den * vel[1], jz = den * vel[2];
    forcing<T>(h, acc, den, jx, jy, jz, rlx_cff2, rlx_cff2);
#pragma unroll
    for (int l = 0; l < Q; ++l) f[l] *= rlx_cff1;
    project<T, true, DIV>(h, f);
    return den;
}

// forw_euler_TRT, D3Q19.pyx:745-853. prm[0] = 1/tau_even, prm[1] = 1/tau_odd.
template <class T, bool SH, bool DIV = false>
LBD T collide_trt(T* f, const T* acc, const T* prm, T rho0, T* vel, T& ds) {
    T inv_tau_e = prm[0], inv_tau_o = prm[1];
    T rlx_cff1 = T(1.0) - T(0.5) * inv_tau_o;
    T rlx_cff2 = T(1.0) - T(0.5) * inv_tau_e;
    T rlx_cff3 = T(1.0) - T(0.5) * (inv_tau_e + inv_tau_o);
    T rlx_cff4 = T(0.5) * (inv_tau_e - inv_tau_o);
    T h[10], den;
    T s = moments_den_vel<T, SH>(f, rho0, den, vel);
    ds = s;
    vel[0] += T(0.5) * acc[0];
    vel[1] += T(0.5) * acc[1];
    vel[2] += T(0.5) * acc[2];
    heq<T>(den, vel, h);
    if (SH) h[0] = s;
    h[0] *= inv_tau_e;
    h[1] *= inv_tau_o;
    h[2] *= inv_tau_o;
    h[3] *= inv_tau_o;
#pragma unroll
    for (int m = 4; m < 10; ++m) h[m] *= inv_tau_e;
    T jx = den * vel[0], jy = den * vel[1], jz = den * vel[2];
    forcing<T>(h, acc, den, jx, jy, jz, rlx_cff1, rlx_cff2);
#pragma unroll
    for (int l = 0; l < 9; ++l) {
        T faux = f[l];
        f[l] = rlx_cff3 * faux - rlx_cff4 * f[18 - l];
        f[18 - l] = rlx_cff3 * f[18 - l] - rlx_cff4 * faux;
    }
    f[9] *= (T(1.0) - inv_tau_e);
    project<T, true, DIV>(h, f);
    return den;
}

// forw_euler_regul, D3Q19.pyx:859-940
template <class T, bool SH, bool DIV = false>
LBD T collide_reg(T* f, const T* acc, const T* prm, T rho0, T* vel, T& ds) {
    T inv_tau = prm[0], rlx_cff1 = T(1.0) - inv_tau, rlx_cff2 = (T(1.0) - T(0.5) * inv_tau);
    T he[10], h[10];
    moments_h<T>(f, h);
    ds = h[0];
    T den = SH ? rho0 + h[0] : h[0];
    T inv_den = T(1.0) / den;
    T jx = h[1] + T(0.5) * acc[0] * den;
    T jy = h[2] + T(0.5) * acc[1] * den;
    T jz = h[3] + T(0.5) * acc[2] * den;
    vel[0] = inv_den * jx;
    vel[1] = inv_den * jy;
    vel[2] = inv_den * jz;
    heq<T>(den, vel, he);
    if (SH) he[0] = h[0];
#pragma unroll
    for (int m = 0; m < 10; ++m) h[m] *= rlx_cff1;
#pragma unroll
    for (int m = 0; m < 10; ++m) h[m] += inv_tau * he[m];
    forcing<T>(h, acc, den, jx, jy, jz, rlx_cff2, rlx_cff2);
    project<T, false, DIV>(h, f);
    return den;
}

// DIV: the projection's quotients are corrected (the FP32-shifted variant:
// its error against FP64 on a 512x128x128 duct drops from 2.4e-5 to 1.2e-6;
// FP16 storage, whose error is its storage, keeps the plain products)
template <int RLX, class T, bool SH, bool DIV = false>
LBD T collide(T* f, const T* acc, const T* prm, T rho0, T* vel, T& ds) {
    if constexpr (RLX == BGK) return collide_bgk<T, SH, DIV>(f, acc, prm, rho0, vel, ds);
    else if constexpr (RLX == TRT) return collide_trt<T, SH, DIV>(f, acc, prm, rho0, vel, ds);
    else return collide_reg<T, SH, DIV>(f, acc, prm, rho0, vel, ds);
}

// ---- Zou-He faces ----------------------------------------------------------
// _assign_f_1x .. _assign_f_z1, bcs.pyx:487-716. The weights of every
// assigned population equal those of its partner and the transverse
// corrections are weight-balanced, so the formulas hold unchanged for
// shifted populations.
template <class T>
LBD void assign_face(int face, T jx, T jy, T jz, T* f) {
    const T s6 = T(1.0) / T(6.0), s3 = T(1.0) / T(3.0), h = T(0.5);
    switch (face) {
    case F1X:
        f[14] = (f[4] + s6 * jx + h * (f[11] + f[12] + f[13] - f[5] - f[6] - f[7]) - h * jy);
        f[15] = (f[3] + s6 * jx + h * (f[7] + f[10] + f[13] - f[5] - f[8] - f[11]) - h * jz);
        f[16] = f[2] + s3 * jx;
        f[17] = (f[1] + s6 * jx - h * (f[7] + f[10] + f[13] - f[5] - f[8] - f[11]) + h * jz);
        f[18] = (f[0] + s6 * jx - h * (f[11] + f[12] + f[13] - f[5] - f[6] - f[7]) + h * jy);
        break;
    case FX1:
        f[0] = (f[18] - s6 * jx + h * (f[11] + f[12] + f[13] - f[5] - f[6] - f[7]) - h * jy);
        f[1] = (f[17] - s6 * jx + h * (f[7] + f[10] + f[13] - f[5] - f[8] - f[11]) - h * jz);
        f[2] = f[16] - s3 * jx;
        f[3] = (f[15] - s6 * jx - h * (f[7] + f[10] + f[13] - f[5] - f[8] - f[11]) + h * jz);
        f[4] = (f[14] - s6 * jx - h * (f[11] + f[12] + f[13] - f[5] - f[6] - f[7]) + h * jy);
        break;
    case F1Y:
        f[4] = (f[14] + s6 * jy + h * (f[15] + f[16] + f[17] - f[1] - f[2] - f[3]) - h * jx);
        f[11] = (f[7] + s6 * jy + h * (f[17] + f[10] + f[3] - f[15] - f[8] - f[1]) - h * jz);
        f[12] = f[6] + s3 * jy;
        f[13] = (f[5] + s6 * jy - h * (f[17] + f[10] + f[3] - f[15] - f[8] - f[1]) + h * jz);
        f[18] = (f[0] + s6 * jy - h * (f[15] + f[16] + f[17] - f[1] - f[2] - f[3]) + h * jx);
        break;
    case FY1:
        f[0] = (f[18] - s6 * jy + h * (f[15] + f[16] + f[17] - f[1] - f[2] - f[3]) - h * jx);
        f[5] = (f[13] - s6 * jy + h * (f[17] + f[10] + f[3] - f[15] - f[8] - f[1]) - h * jz);
        f[6] = f[12] - s3 * jy;
        f[7] = (f[11] - s6 * jy - h * (f[17] + f[10] + f[3] - f[15] - f[8] - f[1]) + h * jz);
        f[14] = (f[4] - s6 * jy - h * (f[15] + f[16] + f[17] - f[1] - f[2] - f[3]) + h * jx);
        break;
    case F1Z:
        f[3] = (f[15] + s6 * jz + h * (f[18] + f[16] + f[14] - f[4] - f[2] - f[0]) - h * jx);
        f[7] = (f[11] + s6 * jz + h * (f[18] + f[12] + f[4] - f[14] - f[6] - f[0]) - h * jy);
        f[10] = f[8] + s3 * jz;
        f[13] = (f[5] + s6 * jz - h * (f[18] + f[12] + f[4] - f[14] - f[6] - f[0]) + h * jy);
        f[17] = (f[1] + s6 * jz - h * (f[18] + f[16] + f[14] - f[4] - f[2] - f[0]) + h * jx);
        break;
    default:  // FZ1
        f[1] = (f[17] - s6 * jz + h * (f[18] + f[16] + f[14] - f[4] - f[2] - f[0]) - h * jx);
        f[5] = (f[13] - s6 * jz + h * (f[18] + f[12] + f[4] - f[14] - f[6] - f[0]) - h * jy);
        f[8] = f[10] - s3 * jz;
        f[11] = (f[7] - s6 * jz - h * (f[18] + f[12] + f[4] - f[14] - f[6] - f[0]) + h * jy);
        f[15] = (f[3] - s6 * jz - h * (f[18] + f[16] + f[14] - f[4] - f[2] - f[0]) + h * jx);
        break;
    }
}

// Sums of the 14 known populations (den_kn) and of the 5 populations leaving
// through the face (j_out), in the reference's summation order (bcs.pyx).
template <class T>
LBD void zh_sums(int face, const T* f, T& den_kn, T& jout) {
    switch (face) {
    case F1X:
        den_kn = (f[9] + f[2] + f[6] + f[12] + f[8] + f[5] + f[11] + f[1] + f[0] + f[4] + f[10] + f[7] + f[13] + f[3]);
        jout = (f[2] + f[1] + f[0] + f[4] + f[3]);
        break;
    case FX1:
        den_kn = (f[9] + f[16] + f[6] + f[12] + f[8] + f[5] + f[11] + f[15] + f[14] + f[18] + f[10] + f[7] + f[13] + f[17]);
        jout = (f[16] + f[15] + f[14] + f[18] + f[17]);
        break;
    case F1Y:
        den_kn = (f[9] + f[2] + f[16] + f[6] + f[8] + f[5] + f[1] + f[15] + f[0] + f[14] + f[10] + f[7] + f[3] + f[17]);
        jout = (f[6] + f[5] + f[0] + f[14] + f[7]);
        break;
    case FY1:
        den_kn = (f[9] + f[2] + f[16] + f[12] + f[8] + f[11] + f[1] + f[15] + f[4] + f[18] + f[10] + f[13] + f[3] + f[17]);
        jout = (f[12] + f[11] + f[4] + f[18] + f[13]);
        break;
    case F1Z:
        den_kn = (f[9] + f[2] + f[16] + f[6] + f[12] + f[8] + f[5] + f[11] + f[1] + f[15] + f[0] + f[18] + f[14] + f[4]);
        jout = (f[8] + f[5] + f[11] + f[1] + f[15]);
        break;
    default:
        den_kn = (f[9] + f[2] + f[16] + f[6] + f[12] + f[10] + f[7] + f[13] + f[3] + f[17] + f[0] + f[18] + f[14] + f[4]);
        jout = (f[10] + f[7] + f[13] + f[3] + f[17]);
        break;
    }
}

// fix_den_face_*, bcs.pyx:31-260: the density and the tangential velocity are
// enforced, the normal momentum follows from mass balance. `ddev` is den -
// rho0, formed by the caller before rounding to T (the shifted sums miss
// rho0 * (5/6) and rho0 * (1/6), which it restores; rounding den itself to
// FP32 would cost the deviation up to 6e-8); unused without SH.
template <class T, bool SH>
LBD void zh_fix_den(int face, T den, T ddev, const T* vel, T* f) {
    T den_kn, jout;
    zh_sums<T>(face, f, den_kn, jout);
    T dd = SH ? ddev : den;
    int axis = face >> 1;
    bool low = (face & 1) == 0;
    T jn;
    if (low) jn = T(1.0) * (T(1.0) * (dd - den_kn) - jout);        // j_pos - j_neg
    else jn = T(1.0) * (jout - T(1.0) * (dd - den_kn));            // j_pos - j_neg
    T jx, jy, jz;
    if (axis == 0) { jx = jn; jy = T(1.0) * den * vel[1]; jz = T(1.0) * den * vel[2]; }
    else if (axis == 1) { jy = jn; jx = T(1.0) * den * vel[0]; jz = T(1.0) * den * vel[2]; }
    else { jz = jn; jx = T(1.0) * den * vel[0]; jy = T(1.0) * den * vel[1]; }
    assign_face<T>(face, jx, jy, jz, f);
}

// fix_vel_face_*, bcs.pyx:265-482: the full velocity is enforced, the density
// follows from the known populations.
template <class T, bool SH>
LBD void zh_fix_vel(int face, const T* vel, T rho0, T* f) {
    T den_kn, jout;
    zh_sums<T>(face, f, den_kn, jout);
    int axis = face >> 1;
    bool low = (face & 1) == 0;
    T num = T(1.0) * den_kn + jout;
    if (SH) num = num + rho0;
    T den_bc = low ? num / (T(1.0) - vel[axis]) : num / (T(1.0) + vel[axis]);
    T jx = den_bc * T(1.0) * vel[0];
    T jy = den_bc * T(1.0) * vel[1];
    T jz = den_bc * T(1.0) * vel[2];
    assign_face<T>(face, jx, jy, jz, f);
}

}  // namespace lb
