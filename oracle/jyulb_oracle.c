/*
 * CPU FP64 oracle for the jyulb isothermal evolve path -- TEST INFRASTRUCTURE.
 *
 * This file is the checker, not the product: only tests/, the smoke check and
 * the CPU-baseline leg of bench.py load it. It restates, in plain C with
 * OpenMP, the two-lattice ("A-B") pull scheme of the reference solver
 * (jyulb/isothermal.pyx, jyulb/D3Q19.pyx, jyulb/bcs.pyx) on halo-padded
 * dense arrays, keeping the reference's floating-point operation order
 * term by term so that, compiled without FMA contraction, it reproduces the
 * reference bit for bit. Differences from the reference that do not change
 * results: 64-bit indices (the reference's int indices overflow for
 * NCNT > 113,025,455, isothermal.pxd:136) and a population-major layout
 * f[l*NCNT + n] instead of Bundle C (isothermal.pyx:1536-1704).
 *
 * Array conventions (C order, halo layer of one site on every face):
 *   phase u8 [NI][NJ][NK], den f64 [NI][NJ][NK], vel/acc f64 [NI][NJ][NK][3],
 *   f1/f2 f64 [19][NI*NJ*NK], face data as in isothermal.pyx:243-263.
 * Lattice vectors l = 0..18 in the order of D3Q19.pxd:59-79.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

enum { WALL = 0, PERIODIC = 1, FIXED_DEN = 2, FIXED_VEL = 3, SYMMETRIC = 4 };
enum { SOLID = 0 };
enum { BGK = 0, TRT = 1, REGUL = 2 };
enum { F1X = 0, FX1 = 1, F1Y = 2, FY1 = 3, F1Z = 4, FZ1 = 5 };

/* D3Q19.pyx:66-79 (constants), 102-163 (index vectors), 192-211 (weights) */
static const double AS2 = 3.0;
static const double W0 = 12.0 / 36.0, W1 = 2.0 / 36.0, W2 = 1.0 / 36.0;
static const int LI[19] = {-1, -1, -1, -1, -1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1};
static const int LJ[19] = {-1, 0, 0, 0, 1, -1, -1, -1, 0, 0, 0, 1, 1, 1, -1, 0, 0, 0, 1};
static const int LK[19] = {0, -1, 0, 1, 0, -1, 0, 1, -1, 0, 1, -1, 0, 1, 0, -1, 0, 1, 0};
/* D3Q19.pyx:215-253: lattice vectors along the outward normal of each face */
static const int FDIR[6][5] = {
    {0, 1, 2, 3, 4}, {14, 15, 16, 17, 18}, {0, 5, 6, 7, 14},
    {4, 11, 12, 13, 18}, {1, 5, 8, 11, 15}, {3, 7, 10, 13, 17}};

typedef struct {
    int32_t NI, NJ, NK;
    int32_t rlx;
    int32_t bc[6];
    double g[3];
    double prm[2];
    uint8_t *phase;
    double *den, *vel, *acc;
    double *f1, *f2;
    double *bcden[6];
    double *bcvel[6];
    int64_t nfluid;
    const uint32_t *sites; /* [3][nfluid], interior coordinates (halo excluded) */
} orc_state;

static inline double lw(int l) {
    if (l == 9) return W0;
    int nz = (LI[l] != 0) + (LJ[l] != 0) + (LK[l] != 0);
    return nz == 1 ? W1 : W2;
}

#define IDX(s, i, j, k) (((int64_t)(i) * (s)->NJ + (j)) * (s)->NK + (k))
#define FOFF(s, l) ((int64_t)(l) * (int64_t)(s)->NI * (s)->NJ * (s)->NK)

/* ---- D3Q19 moment helpers ------------------------------------------------ */

/* f_to_den_vel, D3Q19.pyx:348-390 */
static inline double f_to_den_vel(const double *f, double *vel) {
    double den_1x = f[0] + f[1] + f[2] + f[3] + f[4];
    double den_x1 = f[14] + f[15] + f[16] + f[17] + f[18];
    double den_x0_1y = f[5] + f[6] + f[7];
    double den_x0_y1 = f[11] + f[12] + f[13];
    double den = den_1x + den_x1 + den_x0_1y + den_x0_y1 + f[8] + f[9] + f[10];
    double jx = den_x1 - den_1x;
    double jy = f[4] + den_x0_y1 + f[18] - f[0] - den_x0_1y - f[14];
    double jz = f[3] + f[7] + f[10] + f[13] + f[17] - f[1] - f[5] - f[8] - f[11] - f[15];
    double inv_den = 1.0 / den;
    vel[0] = inv_den * jx;
    vel[1] = inv_den * jy;
    vel[2] = inv_den * jz;
    return den;
}

/* f_to_h, D3Q19.pyx:395-452 */
static inline void f_to_h(const double *f, double *h) {
    double den_1x = f[0] + f[1] + f[2] + f[3] + f[4];
    double den_x1 = f[14] + f[15] + f[16] + f[17] + f[18];
    double den_x0_1y = f[5] + f[6] + f[7];
    double den_x0_y1 = f[11] + f[12] + f[13];
    double den_1y = f[0] + den_x0_1y + f[14];
    double den_y1 = f[4] + den_x0_y1 + f[18];
    double den_1z = f[1] + f[5] + f[8] + f[11] + f[15];
    double den_z1 = f[3] + f[7] + f[10] + f[13] + f[17];
    double den = den_1x + den_x1 + den_x0_1y + den_x0_y1 + f[8] + f[9] + f[10];
    double jx = den_x1 - den_1x;
    double jy = den_y1 - den_1y;
    double jz = den_z1 - den_1z;
    double m2_xx = den_1x + den_x1;
    double m2_yy = den_1y + den_y1;
    double m2_zz = den_1z + den_z1;
    double m2_xy = f[0] + f[18] - f[4] - f[14];
    double m2_xz = f[1] + f[17] - f[3] - f[15];
    double m2_yz = f[5] + f[13] - f[7] - f[11];
    const double INV_AS2 = 1.0 / AS2;
    h[0] = den;
    h[1] = jx;
    h[2] = jy;
    h[3] = jz;
    h[4] = m2_xx - INV_AS2 * den;
    h[5] = m2_xy;
    h[6] = m2_xz;
    h[7] = m2_yy - INV_AS2 * den;
    h[8] = m2_yz;
    h[9] = m2_zz - INV_AS2 * den;
}

/* Kinetic projection P(h); add != 0 accumulates into f.
 * h_to_f D3Q19.pyx:460-527 and add_h_to_f D3Q19.pyx:532-599 */
static inline void project(const double *h, double *f, int add) {
    double e_xy = 2 * h[0] + 6 * h[4] + 6 * h[7] - 3 * h[9];
    double e_1x_1y = e_xy + 18 * h[5];
    double o_1x_1y = -6 * h[1] - 6 * h[2];
    double e_1x_y1 = e_xy - 18 * h[5];
    double o_1x_y1 = -6 * h[1] + 6 * h[2];
    double e_xz = 2 * h[0] + 6 * h[4] - 3 * h[7] + 6 * h[9];
    double e_1x_1z = e_xz + 18 * h[6];
    double o_1x_1z = -6 * h[1] - 6 * h[3];
    double e_1x_z1 = e_xz - 18 * h[6];
    double o_1x_z1 = -6 * h[1] + 6 * h[3];
    double e_yz = 2 * h[0] - 3 * h[4] + 6 * h[7] + 6 * h[9];
    double e_1y_1z = e_yz + 18 * h[8];
    double o_1y_1z = -6 * h[2] - 6 * h[3];
    double e_1y_z1 = e_yz - 18 * h[8];
    double o_1y_z1 = -6 * h[2] + 6 * h[3];
    double e_1x = 4 * h[0] + 12 * h[4] - 6 * h[7] - 6 * h[9];
    double e_1y = 4 * h[0] - 6 * h[4] + 12 * h[7] - 6 * h[9];
    double e_1z = 4 * h[0] - 6 * h[4] - 6 * h[7] + 12 * h[9];
    const double c72 = 1.0 / 72.0;
    double p[19];
    p[0] = c72 * (e_1x_1y + o_1x_1y);
    p[18] = c72 * (e_1x_1y - o_1x_1y);
    p[1] = c72 * (e_1x_1z + o_1x_1z);
    p[17] = c72 * (e_1x_1z - o_1x_1z);
    p[2] = c72 * (e_1x - 12 * h[1]);
    p[16] = c72 * (e_1x + 12 * h[1]);
    p[3] = c72 * (e_1x_z1 + o_1x_z1);
    p[15] = c72 * (e_1x_z1 - o_1x_z1);
    p[4] = c72 * (e_1x_y1 + o_1x_y1);
    p[14] = c72 * (e_1x_y1 - o_1x_y1);
    p[5] = c72 * (e_1y_1z + o_1y_1z);
    p[13] = c72 * (e_1y_1z - o_1y_1z);
    p[6] = c72 * (e_1y - 12 * h[2]);
    p[12] = c72 * (e_1y + 12 * h[2]);
    p[7] = c72 * (e_1y_z1 + o_1y_z1);
    p[11] = c72 * (e_1y_z1 - o_1y_z1);
    p[8] = c72 * (e_1z - 12 * h[3]);
    p[10] = c72 * (e_1z + 12 * h[3]);
    p[9] = (1.0 / 3.0) * h[0] - 0.5 * (h[4] + h[7] + h[9]);
    if (add)
        for (int l = 0; l < 19; ++l) f[l] += p[l];
    else
        for (int l = 0; l < 19; ++l) f[l] = p[l];
}

/* den_vel_to_heq, D3Q19.pyx:604-630 */
static inline void den_vel_to_heq(double den, const double *vel, double *heq) {
    double jx = den * vel[0], jy = den * vel[1], jz = den * vel[2];
    heq[0] = den;
    heq[1] = jx;
    heq[2] = jy;
    heq[3] = jz;
    heq[4] = jx * vel[0];
    heq[5] = jx * vel[1];
    heq[6] = jx * vel[2];
    heq[7] = jy * vel[1];
    heq[8] = jy * vel[2];
    heq[9] = jz * vel[2];
}

/* Forcing moments shared by the three schemes (D3Q19.pyx:717-729, 812-824,
 * 927-935): h1 += c1*acc*den, h2 += c2*(acc j + j acc). */
static inline void add_forcing(double *h, const double *acc, double den, double jx,
                               double jy, double jz, double c1, double c2) {
    h[1] += c1 * acc[0] * den;
    h[2] += c1 * acc[1] * den;
    h[3] += c1 * acc[2] * den;
    h[4] += 2.0 * c2 * acc[0] * jx;
    h[5] += c2 * (acc[0] * jy + acc[1] * jx);
    h[6] += c2 * (acc[0] * jz + acc[2] * jx);
    h[7] += 2.0 * c2 * acc[1] * jy;
    h[8] += c2 * (acc[1] * jz + acc[2] * jy);
    h[9] += 2.0 * c2 * acc[2] * jz;
}

/* forw_euler_BGK, D3Q19.pyx:655-740 */
static double coll_bgk(double *f, const double *acc, const double *prm, double *vel) {
    double inv_tau = prm[0], rlx_cff1 = 1.0 - inv_tau, rlx_cff2 = (1.0 - 0.5 * inv_tau);
    double h[10];
    double den = f_to_den_vel(f, vel);
    vel[0] += 0.5 * acc[0];
    vel[1] += 0.5 * acc[1];
    vel[2] += 0.5 * acc[2];
    den_vel_to_heq(den, vel, h);
    for (int m = 0; m < 10; ++m) h[m] *= inv_tau;
    double jx = den * vel[0], jy = den * vel[1], jz = den * vel[2];
    add_forcing(h, acc, den, jx, jy, jz, rlx_cff2, rlx_cff2);
    for (int l = 0; l < 19; ++l) f[l] *= rlx_cff1;
    project(h, f, 1);
    return den;
}

/* forw_euler_TRT, D3Q19.pyx:745-853 */
static double coll_trt(double *f, const double *acc, const double *prm, double *vel) {
    double inv_tau_e = prm[0], inv_tau_o = prm[1];
    double rlx_cff1 = 1.0 - 0.5 * inv_tau_o;
    double rlx_cff2 = 1.0 - 0.5 * inv_tau_e;
    double rlx_cff3 = 1.0 - 0.5 * (inv_tau_e + inv_tau_o);
    double rlx_cff4 = 0.5 * (inv_tau_e - inv_tau_o);
    double h[10];
    double den = f_to_den_vel(f, vel);
    vel[0] += 0.5 * acc[0];
    vel[1] += 0.5 * acc[1];
    vel[2] += 0.5 * acc[2];
    den_vel_to_heq(den, vel, h);
    h[0] *= inv_tau_e;
    h[1] *= inv_tau_o;
    h[2] *= inv_tau_o;
    h[3] *= inv_tau_o;
    for (int m = 4; m < 10; ++m) h[m] *= inv_tau_e;
    double jx = den * vel[0], jy = den * vel[1], jz = den * vel[2];
    add_forcing(h, acc, den, jx, jy, jz, rlx_cff1, rlx_cff2);
    for (int l = 0; l < 9; ++l) {
        int o = 18 - l;
        double faux = f[l];
        f[l] = rlx_cff3 * faux - rlx_cff4 * f[o];
        f[o] = rlx_cff3 * f[o] - rlx_cff4 * faux;
    }
    f[9] *= (1.0 - inv_tau_e);
    project(h, f, 1);
    return den;
}

/* forw_euler_regul, D3Q19.pyx:859-940 */
static double coll_reg(double *f, const double *acc, const double *prm, double *vel) {
    double inv_tau = prm[0], rlx_cff1 = 1.0 - inv_tau, rlx_cff2 = (1.0 - 0.5 * inv_tau);
    double heq[10], h[10];
    f_to_h(f, h);
    double den = h[0];
    double inv_den = 1.0 / den;
    double jx = h[1] + 0.5 * acc[0] * den;
    double jy = h[2] + 0.5 * acc[1] * den;
    double jz = h[3] + 0.5 * acc[2] * den;
    vel[0] = inv_den * jx;
    vel[1] = inv_den * jy;
    vel[2] = inv_den * jz;
    den_vel_to_heq(den, vel, heq);
    for (int m = 0; m < 10; ++m) h[m] *= rlx_cff1;
    for (int m = 0; m < 10; ++m) h[m] += inv_tau * heq[m];
    add_forcing(h, acc, den, jx, jy, jz, rlx_cff2, rlx_cff2);
    project(h, f, 0);
    return den;
}

/* ---- Zou-He faces, bcs.pyx:31-716 ----------------------------------------- */

/* _assign_f_1x .. _assign_f_z1, bcs.pyx:487-716 */
static inline void assign_1x(double jx, double jy, double jz, double *f) {
    f[14] = (f[4] + (1.0 / 6.0) * jx + 0.5 * (f[11] + f[12] + f[13] - f[5] - f[6] - f[7]) - 0.5 * jy);
    f[15] = (f[3] + (1.0 / 6.0) * jx + 0.5 * (f[7] + f[10] + f[13] - f[5] - f[8] - f[11]) - 0.5 * jz);
    f[16] = f[2] + (1.0 / 3.0) * jx;
    f[17] = (f[1] + (1.0 / 6.0) * jx - 0.5 * (f[7] + f[10] + f[13] - f[5] - f[8] - f[11]) + 0.5 * jz);
    f[18] = (f[0] + (1.0 / 6.0) * jx - 0.5 * (f[11] + f[12] + f[13] - f[5] - f[6] - f[7]) + 0.5 * jy);
}
static inline void assign_x1(double jx, double jy, double jz, double *f) {
    f[0] = (f[18] - (1.0 / 6.0) * jx + 0.5 * (f[11] + f[12] + f[13] - f[5] - f[6] - f[7]) - 0.5 * jy);
    f[1] = (f[17] - (1.0 / 6.0) * jx + 0.5 * (f[7] + f[10] + f[13] - f[5] - f[8] - f[11]) - 0.5 * jz);
    f[2] = f[16] - (1.0 / 3.0) * jx;
    f[3] = (f[15] - (1.0 / 6.0) * jx - 0.5 * (f[7] + f[10] + f[13] - f[5] - f[8] - f[11]) + 0.5 * jz);
    f[4] = (f[14] - (1.0 / 6.0) * jx - 0.5 * (f[11] + f[12] + f[13] - f[5] - f[6] - f[7]) + 0.5 * jy);
}
static inline void assign_1y(double jx, double jy, double jz, double *f) {
    f[4] = (f[14] + (1.0 / 6.0) * jy + 0.5 * (f[15] + f[16] + f[17] - f[1] - f[2] - f[3]) - 0.5 * jx);
    f[11] = (f[7] + (1.0 / 6.0) * jy + 0.5 * (f[17] + f[10] + f[3] - f[15] - f[8] - f[1]) - 0.5 * jz);
    f[12] = f[6] + (1.0 / 3.0) * jy;
    f[13] = (f[5] + (1.0 / 6.0) * jy - 0.5 * (f[17] + f[10] + f[3] - f[15] - f[8] - f[1]) + 0.5 * jz);
    f[18] = (f[0] + (1.0 / 6.0) * jy - 0.5 * (f[15] + f[16] + f[17] - f[1] - f[2] - f[3]) + 0.5 * jx);
}
static inline void assign_y1(double jx, double jy, double jz, double *f) {
    f[0] = (f[18] - (1.0 / 6.0) * jy + 0.5 * (f[15] + f[16] + f[17] - f[1] - f[2] - f[3]) - 0.5 * jx);
    f[5] = (f[13] - (1.0 / 6.0) * jy + 0.5 * (f[17] + f[10] + f[3] - f[15] - f[8] - f[1]) - 0.5 * jz);
    f[6] = f[12] - (1.0 / 3.0) * jy;
    f[7] = (f[11] - (1.0 / 6.0) * jy - 0.5 * (f[17] + f[10] + f[3] - f[15] - f[8] - f[1]) + 0.5 * jz);
    f[14] = (f[4] - (1.0 / 6.0) * jy - 0.5 * (f[15] + f[16] + f[17] - f[1] - f[2] - f[3]) + 0.5 * jx);
}
static inline void assign_1z(double jx, double jy, double jz, double *f) {
    f[3] = (f[15] + (1.0 / 6.0) * jz + 0.5 * (f[18] + f[16] + f[14] - f[4] - f[2] - f[0]) - 0.5 * jx);
    f[7] = (f[11] + (1.0 / 6.0) * jz + 0.5 * (f[18] + f[12] + f[4] - f[14] - f[6] - f[0]) - 0.5 * jy);
    f[10] = f[8] + (1.0 / 3.0) * jz;
    f[13] = (f[5] + (1.0 / 6.0) * jz - 0.5 * (f[18] + f[12] + f[4] - f[14] - f[6] - f[0]) + 0.5 * jy);
    f[17] = (f[1] + (1.0 / 6.0) * jz - 0.5 * (f[18] + f[16] + f[14] - f[4] - f[2] - f[0]) + 0.5 * jx);
}
static inline void assign_z1(double jx, double jy, double jz, double *f) {
    f[1] = (f[17] - (1.0 / 6.0) * jz + 0.5 * (f[18] + f[16] + f[14] - f[4] - f[2] - f[0]) - 0.5 * jx);
    f[5] = (f[13] - (1.0 / 6.0) * jz + 0.5 * (f[18] + f[12] + f[4] - f[14] - f[6] - f[0]) - 0.5 * jy);
    f[8] = f[10] - (1.0 / 3.0) * jz;
    f[11] = (f[7] - (1.0 / 6.0) * jz - 0.5 * (f[18] + f[12] + f[4] - f[14] - f[6] - f[0]) + 0.5 * jy);
    f[15] = (f[3] - (1.0 / 6.0) * jz - 0.5 * (f[18] + f[16] + f[14] - f[4] - f[2] - f[0]) + 0.5 * jx);
}

/* fix_den_face_*, bcs.pyx:31-260 */
static void fix_den(int face, double den, const double *vel, double *f) {
    double den_kn, jp, jn, jx, jy, jz;
    switch (face) {
    case F1X:
        den_kn = (f[9] + f[2] + f[6] + f[12] + f[8] + f[5] + f[11] + f[1] + f[0] + f[4] + f[10] + f[7] + f[13] + f[3]);
        jp = 1.0 * (den - den_kn);
        jn = 1.0 * (f[2] + f[1] + f[0] + f[4] + f[3]);
        jx = 1.0 * (jp - jn);
        jy = 1.0 * den * vel[1], jz = 1.0 * den * vel[2];
        assign_1x(jx, jy, jz, f);
        break;
    case FX1:
        den_kn = (f[9] + f[16] + f[6] + f[12] + f[8] + f[5] + f[11] + f[15] + f[14] + f[18] + f[10] + f[7] + f[13] + f[17]);
        jn = 1.0 * (den - den_kn);
        jp = 1.0 * (f[16] + f[15] + f[14] + f[18] + f[17]);
        jx = 1.0 * (jp - jn);
        jy = 1.0 * den * vel[1], jz = 1.0 * den * vel[2];
        assign_x1(jx, jy, jz, f);
        break;
    case F1Y:
        den_kn = (f[9] + f[2] + f[16] + f[6] + f[8] + f[5] + f[1] + f[15] + f[0] + f[14] + f[10] + f[7] + f[3] + f[17]);
        jp = 1.0 * (den - den_kn);
        jn = 1.0 * (f[6] + f[5] + f[0] + f[14] + f[7]);
        jy = 1.0 * (jp - jn);
        jx = 1.0 * den * vel[0], jz = 1.0 * den * vel[2];
        assign_1y(jx, jy, jz, f);
        break;
    case FY1:
        den_kn = (f[9] + f[2] + f[16] + f[12] + f[8] + f[11] + f[1] + f[15] + f[4] + f[18] + f[10] + f[13] + f[3] + f[17]);
        jn = 1.0 * (den - den_kn);
        jp = 1.0 * (f[12] + f[11] + f[4] + f[18] + f[13]);
        jy = 1.0 * (jp - jn);
        jx = 1.0 * den * vel[0], jz = 1.0 * den * vel[2];
        assign_y1(jx, jy, jz, f);
        break;
    case F1Z:
        den_kn = (f[9] + f[2] + f[16] + f[6] + f[12] + f[8] + f[5] + f[11] + f[1] + f[15] + f[0] + f[18] + f[14] + f[4]);
        jp = 1.0 * (den - den_kn);
        jn = 1.0 * (f[8] + f[5] + f[11] + f[1] + f[15]);
        jz = 1.0 * (jp - jn);
        jx = 1.0 * den * vel[0], jy = 1.0 * den * vel[1];
        assign_1z(jx, jy, jz, f);
        break;
    default: /* FZ1 */
        den_kn = (f[9] + f[2] + f[16] + f[6] + f[12] + f[10] + f[7] + f[13] + f[3] + f[17] + f[0] + f[18] + f[14] + f[4]);
        jn = 1.0 * (den - den_kn);
        jp = 1.0 * (f[10] + f[7] + f[13] + f[3] + f[17]);
        jz = 1.0 * (jp - jn);
        jx = 1.0 * den * vel[0], jy = 1.0 * den * vel[1];
        assign_z1(jx, jy, jz, f);
        break;
    }
}

/* fix_vel_face_*, bcs.pyx:265-482 */
static void fix_vel(int face, const double *vel, double *f) {
    double den_kn, jn, jp, den_bc, jx, jy, jz;
    switch (face) {
    case F1X:
        den_kn = (f[9] + f[2] + f[6] + f[12] + f[8] + f[5] + f[11] + f[1] + f[0] + f[4] + f[10] + f[7] + f[13] + f[3]);
        jn = 1.0 * (f[2] + f[1] + f[0] + f[4] + f[3]);
        den_bc = (1.0 * den_kn + jn) / (1.0 - vel[0]);
        break;
    case FX1:
        den_kn = (f[9] + f[16] + f[6] + f[12] + f[8] + f[5] + f[11] + f[15] + f[14] + f[18] + f[10] + f[7] + f[13] + f[17]);
        jp = 1.0 * (f[16] + f[15] + f[14] + f[18] + f[17]);
        den_bc = (1.0 * den_kn + jp) / (1.0 + vel[0]);
        break;
    case F1Y:
        den_kn = (f[9] + f[2] + f[16] + f[6] + f[8] + f[5] + f[1] + f[15] + f[0] + f[14] + f[10] + f[7] + f[3] + f[17]);
        jn = 1.0 * (f[6] + f[5] + f[0] + f[14] + f[7]);
        den_bc = (1.0 * den_kn + jn) / (1.0 - vel[1]);
        break;
    case FY1:
        den_kn = (f[9] + f[2] + f[16] + f[12] + f[8] + f[11] + f[1] + f[15] + f[4] + f[18] + f[10] + f[13] + f[3] + f[17]);
        jp = 1.0 * (f[12] + f[11] + f[4] + f[18] + f[13]);
        den_bc = (1.0 * den_kn + jp) / (1.0 + vel[1]);
        break;
    case F1Z:
        den_kn = (f[9] + f[2] + f[16] + f[6] + f[12] + f[8] + f[5] + f[11] + f[1] + f[15] + f[0] + f[18] + f[14] + f[4]);
        jn = 1.0 * (f[8] + f[5] + f[11] + f[1] + f[15]);
        den_bc = (1.0 * den_kn + jn) / (1.0 - vel[2]);
        break;
    default: /* FZ1 */
        den_kn = (f[9] + f[2] + f[16] + f[6] + f[12] + f[10] + f[7] + f[13] + f[3] + f[17] + f[0] + f[18] + f[14] + f[4]);
        jp = 1.0 * (f[10] + f[7] + f[13] + f[3] + f[17]);
        den_bc = (1.0 * den_kn + jp) / (1.0 + vel[2]);
        break;
    }
    jx = den_bc * 1.0 * vel[0];
    jy = den_bc * 1.0 * vel[1];
    jz = den_bc * 1.0 * vel[2];
    switch (face) {
    case F1X: assign_1x(jx, jy, jz, f); break;
    case FX1: assign_x1(jx, jy, jz, f); break;
    case F1Y: assign_1y(jx, jy, jz, f); break;
    case FY1: assign_y1(jx, jy, jz, f); break;
    case F1Z: assign_1z(jx, jy, jz, f); break;
    default: assign_z1(jx, jy, jz, f); break;
    }
}

/* ---- Solver steps -------------------------------------------------------- */

/* Solver._set_bc_data, isothermal.pyx:476-502, including the face-Z1
 * condition that tests bc[FACE_1Z] for FIXED_VEL (isothermal.pyx:500). */
void orc_set_bc_data(orc_state *s) {
    int NI = s->NI, NJ = s->NJ, NK = s->NK;
    int fixed[6];
    for (int f = 0; f < 6; ++f) fixed[f] = (s->bc[f] == FIXED_DEN || s->bc[f] == FIXED_VEL);
    fixed[FZ1] = (s->bc[FZ1] == FIXED_DEN || s->bc[F1Z] == FIXED_VEL);
    for (int f = 0; f < 6; ++f) {
        if (!fixed[f]) continue;
        int a = f / 2, lay = (f % 2 == 0) ? 1 : (a == 0 ? NI - 2 : a == 1 ? NJ - 2 : NK - 2);
        int n1 = a == 0 ? NJ : NI, n2 = a == 2 ? NJ : NK;
        for (int p = 0; p < n1; ++p)
            for (int q = 0; q < n2; ++q) {
                int64_t c = a == 0 ? IDX(s, lay, p, q) : a == 1 ? IDX(s, p, lay, q) : IDX(s, p, q, lay);
                int64_t fi = (int64_t)p * n2 + q;
                s->bcden[f][fi] = s->den[c];
                for (int d = 0; d < 3; ++d) s->bcvel[f][3 * fi + d] = s->vel[3 * c + d];
            }
    }
}

/* Solver._enforce_phase_per_sym_bc, isothermal.pyx:509-577 */
void orc_enforce_phase(orc_state *s) {
    int NI = s->NI, NJ = s->NJ, NK = s->NK;
    uint8_t *ph = s->phase;
    if (s->bc[F1X] == PERIODIC || s->bc[F1X] == SYMMETRIC)
        for (int j = 1; j < NJ - 1; ++j)
            for (int k = 1; k < NK - 1; ++k) {
                if (s->bc[F1X] == PERIODIC) ph[IDX(s, NI - 1, j, k)] = ph[IDX(s, 1, j, k)];
                else ph[IDX(s, 0, j, k)] = ph[IDX(s, 1, j, k)];
            }
    if (s->bc[FX1] == PERIODIC || s->bc[FX1] == SYMMETRIC)
        for (int j = 1; j < NJ - 1; ++j)
            for (int k = 1; k < NK - 1; ++k) {
                if (s->bc[FX1] == PERIODIC) ph[IDX(s, 0, j, k)] = ph[IDX(s, NI - 2, j, k)];
                else ph[IDX(s, NI - 1, j, k)] = ph[IDX(s, NI - 2, j, k)];
            }
    if (s->bc[F1Y] == PERIODIC || s->bc[F1Y] == SYMMETRIC)
        for (int i = 0; i < NI; ++i)
            for (int k = 1; k < NK - 1; ++k) {
                if (s->bc[F1Y] == PERIODIC) ph[IDX(s, i, NJ - 1, k)] = ph[IDX(s, i, 1, k)];
                else ph[IDX(s, i, 0, k)] = ph[IDX(s, i, 1, k)];
            }
    if (s->bc[FY1] == PERIODIC || s->bc[FY1] == SYMMETRIC)
        for (int i = 0; i < NI; ++i)
            for (int k = 1; k < NK - 1; ++k) {
                if (s->bc[FY1] == PERIODIC) ph[IDX(s, i, 0, k)] = ph[IDX(s, i, NJ - 2, k)];
                else ph[IDX(s, i, NJ - 1, k)] = ph[IDX(s, i, NJ - 2, k)];
            }
    if (s->bc[F1Z] == PERIODIC || s->bc[F1Z] == SYMMETRIC)
        for (int i = 0; i < NI; ++i)
            for (int j = 0; j < NJ; ++j) {
                if (s->bc[F1Z] == PERIODIC) ph[IDX(s, i, j, NK - 1)] = ph[IDX(s, i, j, 1)];
                else ph[IDX(s, i, j, 0)] = ph[IDX(s, i, j, 1)];
            }
    if (s->bc[FZ1] == PERIODIC || s->bc[FZ1] == SYMMETRIC)
        for (int i = 0; i < NI; ++i)
            for (int j = 0; j < NJ; ++j) {
                if (s->bc[FZ1] == PERIODIC) ph[IDX(s, i, j, 0)] = ph[IDX(s, i, j, NK - 2)];
                else ph[IDX(s, i, j, NK - 1)] = ph[IDX(s, i, j, NK - 2)];
            }
}

/* copy den/vel of cell `src` into `dst`, negating component `neg` (-1: none) */
static inline void cp_dv(orc_state *s, int64_t dst, int64_t src, int neg) {
    s->den[dst] = s->den[src];
    for (int d = 0; d < 3; ++d) s->vel[3 * dst + d] = (d == neg) ? -s->vel[3 * src + d] : s->vel[3 * src + d];
}

/* Solver._enforce_den_vel_per_sym_bc, isothermal.pyx:582-686 */
void orc_enforce_den_vel(orc_state *s) {
    int NI = s->NI, NJ = s->NJ, NK = s->NK;
    if (s->bc[F1X] == PERIODIC || s->bc[F1X] == SYMMETRIC)
        for (int j = 1; j < NJ - 1; ++j)
            for (int k = 1; k < NK - 1; ++k) {
                if (s->bc[F1X] == PERIODIC) cp_dv(s, IDX(s, NI - 1, j, k), IDX(s, 1, j, k), -1);
                else cp_dv(s, IDX(s, 0, j, k), IDX(s, 1, j, k), 0);
            }
    if (s->bc[FX1] == PERIODIC || s->bc[FX1] == SYMMETRIC)
        for (int j = 1; j < NJ - 1; ++j)
            for (int k = 1; k < NK - 1; ++k) {
                if (s->bc[FX1] == PERIODIC) cp_dv(s, IDX(s, 0, j, k), IDX(s, NI - 2, j, k), -1);
                else cp_dv(s, IDX(s, NI - 1, j, k), IDX(s, NI - 2, j, k), 0);
            }
    if (s->bc[F1Y] == PERIODIC || s->bc[F1Y] == SYMMETRIC)
        for (int i = 0; i < NI; ++i)
            for (int k = 1; k < NK - 1; ++k) {
                if (s->bc[F1Y] == PERIODIC) cp_dv(s, IDX(s, i, NJ - 1, k), IDX(s, i, 1, k), -1);
                else cp_dv(s, IDX(s, i, 0, k), IDX(s, i, 1, k), 1);
            }
    if (s->bc[FY1] == PERIODIC || s->bc[FY1] == SYMMETRIC)
        for (int i = 0; i < NI; ++i)
            for (int k = 1; k < NK - 1; ++k) {
                if (s->bc[FY1] == PERIODIC) cp_dv(s, IDX(s, i, 0, k), IDX(s, i, NJ - 2, k), -1);
                else cp_dv(s, IDX(s, i, NJ - 1, k), IDX(s, i, NJ - 2, k), 1);
            }
    if (s->bc[F1Z] == PERIODIC || s->bc[F1Z] == SYMMETRIC)
        for (int i = 0; i < NI; ++i)
            for (int j = 0; j < NJ; ++j) {
                if (s->bc[F1Z] == PERIODIC) cp_dv(s, IDX(s, i, j, NK - 1), IDX(s, i, j, 1), -1);
                else cp_dv(s, IDX(s, i, j, 0), IDX(s, i, j, 1), 2);
            }
    if (s->bc[FZ1] == PERIODIC || s->bc[FZ1] == SYMMETRIC)
        for (int i = 0; i < NI; ++i)
            for (int j = 0; j < NJ; ++j) {
                if (s->bc[FZ1] == PERIODIC) cp_dv(s, IDX(s, i, j, 0), IDX(s, i, j, NK - 2), -1);
                else cp_dv(s, IDX(s, i, j, NK - 1), IDX(s, i, j, NK - 2), 2);
            }
}

/* populations crossing a face, by axis and sign of their component:
 * the c_a = -1 and c_a = +1 sets (cf. FDIR, D3Q19.pyx:215-253) */
static void face_pops(int axis, int sign, int *out) {
    int n = 0;
    for (int l = 0; l < 19; ++l) {
        int c = axis == 0 ? LI[l] : axis == 1 ? LJ[l] : LK[l];
        if (c == sign) out[n++] = l;
    }
}

/* mirror of lattice vector l across an axis */
static int mirror_l(int l, int axis) {
    int ci = LI[l], cj = LJ[l], ck = LK[l];
    if (axis == 0) ci = -ci;
    else if (axis == 1) cj = -cj;
    else ck = -ck;
    for (int m = 0; m < 19; ++m)
        if (LI[m] == ci && LJ[m] == cj && LK[m] == ck) return m;
    return -1;
}

/* One face of Solver._enforce_distr_per_sym_bc, isothermal.pyx:691-837.
 * Periodic: the 5 populations leaving through the face are copied into the
 * opposite halo. Symmetric: they are mirrored into the adjacent halo. Source
 * sites that are SOLID are skipped. */
static void distr_face(orc_state *s, int face) {
    int NI = s->NI, NJ = s->NJ, NK = s->NK;
    int a = face / 2, lo = (face % 2 == 0);
    int N[3] = {NI, NJ, NK};
    int src = lo ? 1 : N[a] - 2;
    int per = s->bc[face] == PERIODIC;
    int dst = per ? (lo ? N[a] - 1 : 0) : (lo ? 0 : N[a] - 1);
    int pops[5];
    face_pops(a, lo ? -1 : 1, pops);
    /* ranges of the two other axes (x pass: j,k interior; y pass: all i, k
     * interior; z pass: all i, j) */
    int r0[3], r1[3];
    r0[0] = a == 0 ? 0 : 0; r1[0] = NI;
    r0[1] = 0; r1[1] = NJ;
    r0[2] = 0; r1[2] = NK;
    if (a == 0) { r0[1] = 1; r1[1] = NJ - 1; r0[2] = 1; r1[2] = NK - 1; }
    if (a == 1) { r0[2] = 1; r1[2] = NK - 1; }
    int b = a == 0 ? 1 : 0, c = a == 2 ? 1 : 2;
    for (int p = r0[b]; p < r1[b]; ++p)
        for (int q = r0[c]; q < r1[c]; ++q) {
            int rs[3], ws[3];
            rs[a] = src; rs[b] = p; rs[c] = q;
            ws[a] = dst; ws[b] = p; ws[c] = q;
            int64_t rn = IDX(s, rs[0], rs[1], rs[2]);
            if (s->phase[rn] == SOLID) continue;
            int64_t wn = IDX(s, ws[0], ws[1], ws[2]);
            for (int t = 0; t < 5; ++t) {
                int l = pops[t];
                int lw_ = per ? l : mirror_l(l, a);
                s->f1[FOFF(s, lw_) + wn] = s->f1[FOFF(s, l) + rn];
            }
        }
}

void orc_enforce_distr(orc_state *s) {
    for (int f = 0; f < 6; ++f)
        if (s->bc[f] == PERIODIC || s->bc[f] == SYMMETRIC) distr_face(s, f);
}

/* Solver._comp_der_vel, isothermal.pyx:958-1023 */
static void comp_der_vel(const orc_state *s, int i, int j, int k, int fdir, const double *vel,
                         double *der) {
    der[0] = der[1] = der[2] = 0.0;
    for (int fl = 0; fl < 5; ++fl) {
        int l = FDIR[fdir][fl];
        int64_t nn = IDX(s, i + LI[l], j + LJ[l], k + LK[l]);
        int64_t no = IDX(s, i - LI[l], j - LJ[l], k - LK[l]);
        double ux = s->vel[3 * nn], uy = s->vel[3 * nn + 1], uz = s->vel[3 * nn + 2];
        double uxo = s->vel[3 * no], uyo = s->vel[3 * no + 1], uzo = s->vel[3 * no + 2];
        double c1 = 0.5, c2 = 0.0, c3 = -0.5;
        if (s->phase[nn] == SOLID) {
            if (s->phase[no] == SOLID) {
                c1 = 1.0;
                c3 = -1.0;
            } else {
                c1 = 4.0 / 3.0;
                c2 = -1.0;
                c3 = -1.0 / 3.0;
            }
        } else if (s->phase[no] == SOLID) {
            c1 = 1.0 / 3.0;
            c2 = 1.0;
            c3 = -4.0 / 3.0;
        }
        double c0 = 2.0 * lw(l) * AS2;
        der[0] += c0 * (c1 * ux + c2 * vel[0] + c3 * uxo);
        der[1] += c0 * (c1 * uy + c2 * vel[1] + c3 * uyo);
        der[2] += c0 * (c1 * uz + c2 * vel[2] + c3 * uzo);
    }
}

/* Solver._init_f, isothermal.pyx:1028-1105 */
void orc_init_f(orc_state *s) {
    double inv_tau = s->prm[0], tau = 1.0 / inv_tau, rlx_cff = 1.0 - inv_tau;
    const double INV_AS2 = 1.0 / AS2;
#pragma omp parallel for schedule(static)
    for (int64_t fn = 0; fn < s->nfluid; ++fn) {
        int i = (int)s->sites[fn] + 1, j = (int)s->sites[s->nfluid + fn] + 1,
            k = (int)s->sites[2 * s->nfluid + fn] + 1;
        int64_t n = IDX(s, i, j, k);
        double vel[3], h[10], f[19], dx[3], dy[3], dz[3];
        double den = s->den[n];
        vel[0] = s->vel[3 * n];
        vel[1] = s->vel[3 * n + 1];
        vel[2] = s->vel[3 * n + 2];
        den_vel_to_heq(den, vel, h);
        comp_der_vel(s, i, j, k, FX1, vel, dx);
        comp_der_vel(s, i, j, k, FY1, vel, dy);
        comp_der_vel(s, i, j, k, FZ1, vel, dz);
        double m2 = -2.0 * tau * INV_AS2 * den;
        h[4] += rlx_cff * m2 * dx[0];
        h[5] += rlx_cff * m2 * 0.5 * (dx[1] + dy[0]);
        h[6] += rlx_cff * m2 * 0.5 * (dx[2] + dz[0]);
        h[7] += rlx_cff * m2 * dy[1];
        h[8] += rlx_cff * m2 * 0.5 * (dy[2] + dz[1]);
        h[9] += rlx_cff * m2 * dz[2];
        project(h, f, 0);
        for (int l = 0; l < 19; ++l) s->f1[FOFF(s, l) + n] = f[l];
    }
}

/* Solver._stream_ijk, isothermal.pyx:1112-1248: pull from x - c_l, or
 * halfway bounce-back with Ladd's moving-wall term when x - c_l is SOLID.
 * The wall term is c*(u_a +/- u_b) with the sign pattern of the reference:
 * the first nonzero component of c_l sets the outer sign. */
static inline void stream_ijk(const orc_state *s, int i, int j, int k, int64_t n, double *f) {
    double den = s->den[n];
    double cff1 = 2.0 * den * W1 * AS2;
    double cff2 = 2.0 * den * W2 * AS2;
    for (int l = 0; l < 19; ++l) {
        if (l == 9) {
            f[9] = s->f1[FOFF(s, 9) + n];
            continue;
        }
        int ci = LI[l], cj = LJ[l], ck = LK[l];
        int64_t nb = IDX(s, i - ci, j - cj, k - ck);
        if (s->phase[nb] != SOLID) {
            f[l] = s->f1[FOFF(s, l) + nb];
            continue;
        }
        const double *u = s->vel + 3 * nb;
        int c[3] = {ci, cj, ck};
        int a = -1, b = -1;
        for (int d = 0; d < 3; ++d)
            if (c[d] != 0) {
                if (a < 0) a = d;
                else b = d;
            }
        double term;
        if (b < 0) term = cff1 * u[a];
        else if (c[a] == c[b]) term = cff2 * (u[a] + u[b]);
        else term = cff2 * (u[a] - u[b]);
        double fo = s->f1[FOFF(s, 18 - l) + n];
        f[l] = (c[a] > 0) ? fo + term : fo - term;
    }
}

/* Solver._update, isothermal.pyx:1253-1326 */
void orc_update(orc_state *s) {
    int NI = s->NI, NJ = s->NJ, NK = s->NK;
#pragma omp parallel for schedule(static)
    for (int64_t fn = 0; fn < s->nfluid; ++fn) {
        int i = (int)s->sites[fn] + 1, j = (int)s->sites[s->nfluid + fn] + 1,
            k = (int)s->sites[2 * s->nfluid + fn] + 1;
        int64_t n = IDX(s, i, j, k);
        double f[19], vel[3], acc[3];
        stream_ijk(s, i, j, k, n, f);
        /* Zou-He faces in the order 1X, X1, 1Y, Y1, 1Z, Z1 (isothermal.pyx:1280-1291);
         * face data indexed as in _face_*_bc (isothermal.pyx:842-951) */
        int on[6] = {i == 1, i == NI - 2, j == 1, j == NJ - 2, k == 1, k == NK - 2};
        int64_t fidx[6] = {(int64_t)j * NK + k, (int64_t)j * NK + k, (int64_t)i * NK + k,
                           (int64_t)i * NK + k, (int64_t)i * NJ + j, (int64_t)i * NJ + j};
        for (int fc = 0; fc < 6; ++fc) {
            if (!on[fc]) continue;
            if (s->bc[fc] == FIXED_DEN)
                fix_den(fc, s->bcden[fc][fidx[fc]], s->bcvel[fc] + 3 * fidx[fc], f);
            else if (s->bc[fc] == FIXED_VEL)
                fix_vel(fc, s->bcvel[fc] + 3 * fidx[fc], f);
        }
        acc[0] = s->g[0] + s->acc[3 * n];
        acc[1] = s->g[1] + s->acc[3 * n + 1];
        acc[2] = s->g[2] + s->acc[3 * n + 2];
        double den;
        if (s->rlx == BGK) den = coll_bgk(f, acc, s->prm, vel);
        else if (s->rlx == TRT) den = coll_trt(f, acc, s->prm, vel);
        else den = coll_reg(f, acc, s->prm, vel);
        s->den[n] = den;
        s->vel[3 * n] = vel[0];
        s->vel[3 * n + 1] = vel[1];
        s->vel[3 * n + 2] = vel[2];
        for (int l = 0; l < 19; ++l) s->f2[FOFF(s, l) + n] = f[l];
    }
}

/* The step loop of Solver.evolve, isothermal.pyx:1408-1415 (the caller does
 * the error checks and _set_bc_data). f1/f2 are swapped after every step;
 * on return s->f1 holds the newest populations. */
void orc_evolve_steps(orc_state *s, uint32_t tsteps) {
    for (uint32_t t = 0; t < tsteps; ++t) {
        orc_enforce_distr(s);
        orc_enforce_den_vel(s);
        orc_update(s);
        double *sw = s->f1;
        s->f1 = s->f2;
        s->f2 = sw;
    }
}

/* Pure helpers exposed for closed-form tests of the moment maps and the
 * collision / Zou-He operators on single sites. */
void orc_project(const double *h, double *f) { project(h, f, 0); }
void orc_f_to_h(const double *f, double *h) { f_to_h(f, h); }
double orc_collide(int rlx, double *f, const double *acc, const double *prm, double *vel) {
    if (rlx == BGK) return coll_bgk(f, acc, prm, vel);
    if (rlx == TRT) return coll_trt(f, acc, prm, vel);
    return coll_reg(f, acc, prm, vel);
}
void orc_zouhe(int face, int kind, double den, const double *vel, double *f) {
    if (kind == FIXED_DEN) fix_den(face, den, vel, f);
    else fix_vel(face, vel, f);
}
