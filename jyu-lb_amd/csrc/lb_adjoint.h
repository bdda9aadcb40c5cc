// Vector-Jacobian products (reverse mode) of the per-site operators of the
// fused step, for the adjoint AA step.
//
// Forward model at a fluid site (lb_physics.h): g' = Z(g; face data) on a
// Zou-He face (identity elsewhere), then f* = C(g'; a) with density
// rho = sum g', momentum j = sum c g', shifted momentum j* = j + rho a/2,
// u* = j*/rho. Every collision writes f* = D g' + P(H) with a linear part D
// (BGK: (1-w) I; TRT: the pair mixing; regularized: (1-w) P h(.)) and the
// projection P of the moments
//   H0 = wE rho, H1 = wO j* + c1 rho a, H2 = wE j* j*/rho + c2 (a j* + j* a)
// (D3Q19.pyx:655-940). The Jacobian therefore depends on (rho, u*, a) only,
// which is what the forward recompute records per site and step.
#pragma once
#include "lb_lattice.h"
#include "lb_physics.h"

namespace lb {

__host__ __device__ constexpr double wgt(int l) { return l == REST ? 12.0 / 36.0 : nnz(l) == 1 ? 2.0 / 36.0 : 1.0 / 36.0; }

// Adjoint moments of lam under the projection P(h)_l =
// w_l [h0 + 3 c.h1 + 4.5 sum_a (c_a^2 - 1/3) h_aa + 9 sum_{a<b} c_a c_b h_ab]:
// sum_l lam_l P(H)_l = A0 H0 + A1.H1 + sum_ab B_ab H2_ab (B symmetric, full sum).
template <class T>
LBD void proj_adjoint(const T* lam, T& A0, T* A1, T B[3][3]) {
    A0 = T(0);
    A1[0] = A1[1] = A1[2] = T(0);
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) B[a][b] = T(0);
#pragma unroll
    for (int l = 0; l < Q; ++l) {
        const T wl = (T)wgt(l) * lam[l];
        const int c[3] = {CI[l], CJ[l], CK[l]};
        A0 += wl;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            A1[a] += T(3) * T(c[a]) * wl;
            B[a][a] += T(4.5) * (T(c[a] * c[a]) - T(1.0 / 3.0)) * wl;
#pragma unroll
            for (int b = a + 1; b < 3; ++b) {
                const T v = T(4.5) * T(c[a] * c[b]) * wl;  // 9 c_a c_b split over B_ab and B_ba
                B[a][b] += v;
                B[b][a] += v;
            }
        }
    }
}

// VJP of the collision. lam: cotangent of f* (19); rho, us: recorded density
// and half-force shifted velocity; acc: acceleration; prm: relaxation rates.
// Writes mu (cotangent of g', 19) and adds the acceleration cotangent to ga.
template <int RLX, class T>
LBD void collide_vjp(const T* lam, T rho, const T* us, const T* acc, const T* prm, T* mu, T* ga) {
    T wE, wO;
    if (RLX == TRT) {
        wE = prm[0];
        wO = prm[1];
    } else {
        wE = wO = prm[0];
    }
    const T c1 = T(1) - T(0.5) * wO, c2 = T(1) - T(0.5) * wE;
    T A0, A1[3], B[3][3];
    proj_adjoint<T>(lam, A0, A1, B);
    const T js[3] = {rho * us[0], rho * us[1], rho * us[2]};
    T Bj[3], Ba[3], jBj = T(0);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        Bj[a] = B[a][0] * js[0] + B[a][1] * js[1] + B[a][2] * js[2];
        Ba[a] = B[a][0] * acc[0] + B[a][1] * acc[1] + B[a][2] * acc[2];
        jBj += js[a] * Bj[a];
    }
    // derivatives w.r.t. j* (shared), rho and a with j* held fixed
    T Gjs[3], Ga[3];
    T Grho = A0 * wE + c1 * (A1[0] * acc[0] + A1[1] * acc[1] + A1[2] * acc[2]) - wE * jBj / (rho * rho);
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        Gjs[a] = wO * A1[a] + T(2) * wE * Bj[a] / rho + T(2) * c2 * Ba[a];
        Ga[a] = c1 * rho * A1[a] + T(2) * c2 * Bj[a];
    }
    // chain through j* = j + rho a / 2
    Grho += T(0.5) * (Gjs[0] * acc[0] + Gjs[1] * acc[1] + Gjs[2] * acc[2]);
#pragma unroll
    for (int a = 0; a < 3; ++a) ga[a] += Ga[a] + T(0.5) * rho * Gjs[a];
    // linear part D^T lam
    if (RLX == BGK) {
        const T r = T(1) - wE;
#pragma unroll
        for (int l = 0; l < Q; ++l) mu[l] = r * lam[l];
    } else if (RLX == TRT) {
        const T c3 = T(1) - T(0.5) * (wE + wO), c4 = T(0.5) * (wE - wO);
#pragma unroll
        for (int l = 0; l < Q; ++l) mu[l] = l == REST ? (T(1) - wE) * lam[l] : c3 * lam[l] - c4 * lam[opp(l)];
    } else {
        // f* = P((1-w) h(g') + ...): d/dg'_l of (1-w) (A0 h0 + A1.h1 + B:h2),
        // h2_ab = sum c_a c_b g - delta_ab rho / 3
        const T r = T(1) - wE;
        const T trB = B[0][0] + B[1][1] + B[2][2];
#pragma unroll
        for (int l = 0; l < Q; ++l) {
            const int c[3] = {CI[l], CJ[l], CK[l]};
            T q = A0 - trB / T(3);
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                q += A1[a] * T(c[a]);
#pragma unroll
                for (int b = 0; b < 3; ++b) q += B[a][b] * T(c[a] * c[b]);
            }
            mu[l] = r * q;
        }
    }
    // moments rho = sum g', j = sum c g'
#pragma unroll
    for (int l = 0; l < Q; ++l) mu[l] += Grho + T(CI[l]) * Gjs[0] + T(CJ[l]) * Gjs[1] + T(CK[l]) * Gjs[2];
}

// VJP of one Zou-He face. mu: cotangent of the populations after the face
// condition; on return nu: cotangent of the populations before it. The map
// is linear in the populations: its transpose is assembled from unit
// evaluations of the forward assignment (assign_face, bcs.pyx:487-716).
// kind FIXED_DEN: den = face density; FIXED_VEL: den = density produced by
// the condition (the site's recorded density). Adds the cotangent of the face
// velocity to gv[3].
// FACE is a template parameter so that the unit evaluations below fold to
// constants at compile time; zh_vjp dispatches on the runtime face.
template <class T, int FACE>
LBD void zh_vjp_face(int kind, T den, const T* vel, const T* mu, T* nu, T* gv) {
    constexpr int face = FACE;
    constexpr int axis = face >> 1;
    constexpr int sgn_in = (face & 1) ? -1 : 1;  // component of the populations entering the domain
    bool unk[Q];
#pragma unroll
    for (int l = 0; l < Q; ++l) unk[l] = ccomp(l, axis) == sgn_in;
    // cotangents of (jx, jy, jz) through the assignment
    T cj[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        T fz[Q];
#pragma unroll
        for (int l = 0; l < Q; ++l) fz[l] = T(0);
        assign_face<T>(face, T(a == 0), T(a == 1), T(a == 2), fz);
        T s = T(0);
#pragma unroll
        for (int l = 0; l < Q; ++l)
            if (unk[l]) s += mu[l] * fz[l];
        cj[a] = s;
    }
    // direct linear part: known k -> unknown u
#pragma unroll
    for (int k = 0; k < Q; ++k) {
        if (unk[k]) {
            nu[k] = T(0);
            continue;
        }
        T fz[Q];
#pragma unroll
        for (int l = 0; l < Q; ++l) fz[l] = T(0);
        fz[k] = T(1);
        assign_face<T>(face, T(0), T(0), T(0), fz);
        T s = mu[k];
#pragma unroll
        for (int l = 0; l < Q; ++l)
            if (unk[l]) s += mu[l] * fz[l];
        nu[k] = s;
    }
    // the momentum depends on the known populations: den_kn covers every
    // known one, the outgoing ones (c_axis = -sgn_in) are counted again in j_out
    if (kind == FIXED_DEN) {
        // low face: j_n = den - den_kn - j_out; high face: j_n = j_out - den + den_kn
        const T s = sgn_in > 0 ? T(-1) : T(1);
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            if (unk[k]) continue;
            const T cnt = ccomp(k, axis) == -sgn_in ? T(2) : T(1);
            nu[k] += cj[axis] * s * cnt;
        }
        // tangential momentum den * v_t
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (a != axis) gv[a] += cj[a] * den;
    } else {
        // den_bc = (den_kn + j_out) / (1 - sgn_in v_n), j = den_bc v
        const T dn = T(1) - T(sgn_in) * vel[axis];
        const T cv = cj[0] * vel[0] + cj[1] * vel[1] + cj[2] * vel[2];
#pragma unroll
        for (int k = 0; k < Q; ++k) {
            if (unk[k]) continue;
            const T cnt = ccomp(k, axis) == -sgn_in ? T(2) : T(1);
            nu[k] += cv * cnt / dn;
        }
#pragma unroll
        for (int a = 0; a < 3; ++a) gv[a] += cj[a] * den;
        gv[axis] += cv * den * T(sgn_in) / dn;
    }
}

template <class T>
LBD void zh_vjp(int face, int kind, T den, const T* vel, const T* mu, T* nu, T* gv) {
    switch (face) {
    case 0: zh_vjp_face<T, 0>(kind, den, vel, mu, nu, gv); break;
    case 1: zh_vjp_face<T, 1>(kind, den, vel, mu, nu, gv); break;
    case 2: zh_vjp_face<T, 2>(kind, den, vel, mu, nu, gv); break;
    case 3: zh_vjp_face<T, 3>(kind, den, vel, mu, nu, gv); break;
    case 4: zh_vjp_face<T, 4>(kind, den, vel, mu, nu, gv); break;
    default: zh_vjp_face<T, 5>(kind, den, vel, mu, nu, gv); break;
    }
}

}  // namespace lb
