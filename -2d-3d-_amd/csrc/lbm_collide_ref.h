// lbm_collide_ref.h -- FP64 collision operators in the reference's exact expression order.
//
// These keep every sum, product and division of the reference's unrolled Fortran in the same
// order and association (2DRB/3DRB.F90:714-943 `collision`, :1121-1191 `collisionT`;
// 2DRB/2DRB.F90:686-759, :917-975), so that with contraction disabled (the translation unit
// is built with -ffp-contract=off) each IEEE-754 operation rounds exactly as on the
// reference's x86-64 build: the FP64 form reproduces the CPU path bit for bit.
#pragma once

#include "lbm_lattice.h"

namespace rblbm {

// Reference `collision` body for one node: F in (f after macro), P out (f_post).
__device__ __forceinline__ void ref_collide_d3q19(const double *F, double *P, double rhoLoc,
                                                  double uLoc, double vLoc, double wLoc, double T,
                                                  const KParams &k)
{
    double m[19], meq[19], mp[19], s[19], src[19];
    const double u2 = uLoc * uLoc + vLoc * vLoc + wLoc * wLoc;
    m[0] = F[0] + F[1] + F[2] + F[3] + F[4] + F[5] + F[6] + F[7] + F[8] + F[9] + F[10] + F[11] +
           F[12] + F[13] + F[14] + F[15] + F[16] + F[17] + F[18];
    m[1] = -30.0 * F[0] - 11.0 * (F[1] + F[2] + F[3] + F[4] + F[5] + F[6]) +
           8.0 * (F[7] + F[8] + F[9] + F[10] + F[11] + F[12] + F[13] + F[14] + F[15] + F[16] +
                  F[17] + F[18]);
    m[2] = 12.0 * F[0] - 4.0 * (F[1] + F[2] + F[3] + F[4] + F[5] + F[6]) +
           (F[7] + F[8] + F[9] + F[10] + F[11] + F[12] + F[13] + F[14] + F[15] + F[16] + F[17] +
            F[18]);
    m[3] = F[1] - F[2] + F[7] - F[8] + F[9] - F[10] + F[11] - F[12] + F[13] - F[14];
    m[4] = -4.0 * F[1] + 4.0 * F[2] + F[7] - F[8] + F[9] - F[10] + F[11] - F[12] + F[13] - F[14];
    m[5] = F[3] - F[4] + F[7] + F[8] - F[9] - F[10] + F[15] - F[16] + F[17] - F[18];
    m[6] = -4.0 * F[3] + 4.0 * F[4] + F[7] + F[8] - F[9] - F[10] + F[15] - F[16] + F[17] - F[18];
    m[7] = F[5] - F[6] + F[11] + F[12] - F[13] - F[14] + F[15] + F[16] - F[17] - F[18];
    m[8] = -4.0 * F[5] + 4.0 * F[6] + F[11] + F[12] - F[13] - F[14] + F[15] + F[16] - F[17] - F[18];
    m[9] = 2.0 * (F[1] + F[2]) - (F[3] + F[4] + F[5] + F[6]) +
           (F[7] + F[8] + F[9] + F[10] + F[11] + F[12] + F[13] + F[14]) -
           2.0 * (F[15] + F[16] + F[17] + F[18]);
    m[10] = -4.0 * (F[1] + F[2]) + 2.0 * (F[3] + F[4] + F[5] + F[6]) +
            (F[7] + F[8] + F[9] + F[10] + F[11] + F[12] + F[13] + F[14]) -
            2.0 * (F[15] + F[16] + F[17] + F[18]);
    m[11] = (F[3] + F[4]) - (F[5] + F[6]) + (F[7] + F[8] + F[9] + F[10]) -
            (F[11] + F[12] + F[13] + F[14]);
    m[12] = -2.0 * (F[3] + F[4]) + 2.0 * (F[5] + F[6]) + (F[7] + F[8] + F[9] + F[10]) -
            (F[11] + F[12] + F[13] + F[14]);
    m[13] = F[7] - F[8] - F[9] + F[10];
    m[14] = F[15] - F[16] - F[17] + F[18];
    m[15] = F[11] - F[12] - F[13] + F[14];
    m[16] = F[7] - F[8] + F[9] - F[10] - F[11] + F[12] - F[13] + F[14];
    m[17] = -F[7] - F[8] + F[9] + F[10] + F[15] - F[16] + F[17] - F[18];
    m[18] = F[11] + F[12] - F[13] - F[14] - F[15] - F[16] + F[17] + F[18];

    meq[0] = rhoLoc;
    meq[1] = rhoLoc * (-11.0 + 19.0 * u2);
    meq[2] = rhoLoc * (3.0 - 11.0 * u2 / 2.0);
    meq[3] = rhoLoc * uLoc;
    meq[4] = -2.0 * rhoLoc * uLoc / 3.0;
    meq[5] = rhoLoc * vLoc;
    meq[6] = -2.0 * rhoLoc * vLoc / 3.0;
    meq[7] = rhoLoc * wLoc;
    meq[8] = -2.0 * rhoLoc * wLoc / 3.0;
    meq[9] = rhoLoc * (2.0 * uLoc * uLoc - vLoc * vLoc - wLoc * wLoc);
    meq[10] = -0.5 * meq[9];
    meq[11] = rhoLoc * (vLoc * vLoc - wLoc * wLoc);
    meq[12] = -0.5 * meq[11];
    meq[13] = rhoLoc * uLoc * vLoc;
    meq[14] = rhoLoc * vLoc * wLoc;
    meq[15] = rhoLoc * uLoc * wLoc;
    meq[16] = 0.0;
    meq[17] = 0.0;
    meq[18] = 0.0;

    const double sv = k.sv, sq = k.sq;
    s[0] = 0.0; s[1] = sv; s[2] = sv; s[3] = 0.0; s[4] = sq; s[5] = 0.0; s[6] = sq; s[7] = 0.0;
    s[8] = sq; s[9] = sv; s[10] = sv; s[11] = sv; s[12] = sv; s[13] = sv; s[14] = sv; s[15] = sv;
    s[16] = sq; s[17] = sq; s[18] = sq;

    const double FxLoc = 0.0;
    const double FyLoc = rhoLoc * k.gbeta * (T - k.tref);
    const double FzLoc = 0.0;
    const double uDotF = uLoc * FxLoc + vLoc * FyLoc + wLoc * FzLoc;
    src[0] = 0.0;
    src[1] = (1.0 - 0.5 * s[1]) * 38.0 * uDotF;
    src[2] = (1.0 - 0.5 * s[2]) * (-11.0) * uDotF;
    src[3] = (1.0 - 0.5 * s[3]) * FxLoc;
    src[4] = (1.0 - 0.5 * s[4]) * (-2.0 / 3.0) * FxLoc;
    src[5] = (1.0 - 0.5 * s[5]) * FyLoc;
    src[6] = (1.0 - 0.5 * s[6]) * (-2.0 / 3.0) * FyLoc;
    src[7] = (1.0 - 0.5 * s[7]) * FzLoc;
    src[8] = (1.0 - 0.5 * s[8]) * (-2.0 / 3.0) * FzLoc;
    src[9] = (1.0 - 0.5 * s[9]) * (4.0 * uLoc * FxLoc - 2.0 * vLoc * FyLoc - 2.0 * wLoc * FzLoc);
    src[10] = (1.0 - 0.5 * s[10]) * (-2.0 * uLoc * FxLoc + vLoc * FyLoc + wLoc * FzLoc);
    src[11] = (1.0 - 0.5 * s[11]) * (2.0 * vLoc * FyLoc - 2.0 * wLoc * FzLoc);
    src[12] = (1.0 - 0.5 * s[12]) * (-vLoc * FyLoc + wLoc * FzLoc);
    src[13] = (1.0 - 0.5 * s[13]) * (uLoc * FyLoc + vLoc * FxLoc);
    src[14] = (1.0 - 0.5 * s[14]) * (vLoc * FzLoc + wLoc * FyLoc);
    src[15] = (1.0 - 0.5 * s[15]) * (uLoc * FzLoc + wLoc * FxLoc);
    src[16] = 0.0;
    src[17] = 0.0;
    src[18] = 0.0;

#pragma unroll
    for (int a = 0; a < 19; ++a) mp[a] = m[a] - s[a] * (m[a] - meq[a]) + src[a];

    P[0] = mp[0] / 19.0 - 5.0 * mp[1] / 399.0 + mp[2] / 21.0;
    P[1] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 + mp[3] / 10.0 - mp[4] / 10.0 +
           mp[9] / 18.0 - mp[10] / 18.0;
    P[2] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 - mp[3] / 10.0 + mp[4] / 10.0 +
           mp[9] / 18.0 - mp[10] / 18.0;
    P[3] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 + mp[5] / 10.0 - mp[6] / 10.0 -
           mp[9] / 36.0 + mp[10] / 36.0 + mp[11] / 12.0 - mp[12] / 12.0;
    P[4] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 - mp[5] / 10.0 + mp[6] / 10.0 -
           mp[9] / 36.0 + mp[10] / 36.0 + mp[11] / 12.0 - mp[12] / 12.0;
    P[5] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 + mp[7] / 10.0 - mp[8] / 10.0 -
           mp[9] / 36.0 + mp[10] / 36.0 - mp[11] / 12.0 + mp[12] / 12.0;
    P[6] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 - mp[7] / 10.0 + mp[8] / 10.0 -
           mp[9] / 36.0 + mp[10] / 36.0 - mp[11] / 12.0 + mp[12] / 12.0;
    P[7] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[3] / 10.0 + mp[4] / 40.0 +
           mp[5] / 10.0 + mp[6] / 40.0 + mp[9] / 36.0 + mp[10] / 72.0 + mp[11] / 12.0 +
           mp[12] / 24.0 + mp[13] / 4.0 + mp[16] / 8.0 - mp[17] / 8.0;
    P[8] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[3] / 10.0 - mp[4] / 40.0 +
           mp[5] / 10.0 + mp[6] / 40.0 + mp[9] / 36.0 + mp[10] / 72.0 + mp[11] / 12.0 +
           mp[12] / 24.0 - mp[13] / 4.0 - mp[16] / 8.0 - mp[17] / 8.0;
    P[9] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[3] / 10.0 + mp[4] / 40.0 -
           mp[5] / 10.0 - mp[6] / 40.0 + mp[9] / 36.0 + mp[10] / 72.0 + mp[11] / 12.0 +
           mp[12] / 24.0 - mp[13] / 4.0 + mp[16] / 8.0 + mp[17] / 8.0;
    P[10] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[3] / 10.0 - mp[4] / 40.0 -
            mp[5] / 10.0 - mp[6] / 40.0 + mp[9] / 36.0 + mp[10] / 72.0 + mp[11] / 12.0 +
            mp[12] / 24.0 + mp[13] / 4.0 - mp[16] / 8.0 + mp[17] / 8.0;
    P[11] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[3] / 10.0 + mp[4] / 40.0 +
            mp[7] / 10.0 + mp[8] / 40.0 + mp[9] / 36.0 + mp[10] / 72.0 - mp[11] / 12.0 -
            mp[12] / 24.0 + mp[15] / 4.0 - mp[16] / 8.0 + mp[18] / 8.0;
    P[12] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[3] / 10.0 - mp[4] / 40.0 +
            mp[7] / 10.0 + mp[8] / 40.0 + mp[9] / 36.0 + mp[10] / 72.0 - mp[11] / 12.0 -
            mp[12] / 24.0 - mp[15] / 4.0 + mp[16] / 8.0 + mp[18] / 8.0;
    P[13] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[3] / 10.0 + mp[4] / 40.0 -
            mp[7] / 10.0 - mp[8] / 40.0 + mp[9] / 36.0 + mp[10] / 72.0 - mp[11] / 12.0 -
            mp[12] / 24.0 - mp[15] / 4.0 - mp[16] / 8.0 - mp[18] / 8.0;
    P[14] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[3] / 10.0 - mp[4] / 40.0 -
            mp[7] / 10.0 - mp[8] / 40.0 + mp[9] / 36.0 + mp[10] / 72.0 - mp[11] / 12.0 -
            mp[12] / 24.0 + mp[15] / 4.0 + mp[16] / 8.0 - mp[18] / 8.0;
    P[15] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[5] / 10.0 + mp[6] / 40.0 +
            mp[7] / 10.0 + mp[8] / 40.0 - mp[9] / 18.0 - mp[10] / 36.0 + mp[14] / 4.0 +
            mp[17] / 8.0 - mp[18] / 8.0;
    P[16] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[5] / 10.0 - mp[6] / 40.0 +
            mp[7] / 10.0 + mp[8] / 40.0 - mp[9] / 18.0 - mp[10] / 36.0 - mp[14] / 4.0 -
            mp[17] / 8.0 - mp[18] / 8.0;
    P[17] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[5] / 10.0 + mp[6] / 40.0 -
            mp[7] / 10.0 - mp[8] / 40.0 - mp[9] / 18.0 - mp[10] / 36.0 - mp[14] / 4.0 +
            mp[17] / 8.0 + mp[18] / 8.0;
    P[18] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[5] / 10.0 - mp[6] / 40.0 -
            mp[7] / 10.0 - mp[8] / 40.0 - mp[9] / 18.0 - mp[10] / 36.0 + mp[14] / 4.0 -
            mp[17] / 8.0 + mp[18] / 8.0;
}

// Reference `collisionT` body: G in, P out; T the macroT value; u, v, w the new velocity;
// bp the M1G history (read, then overwritten when use_g).
__device__ __forceinline__ void ref_collide_d3q7(const double *G, double *P, double Tl, double u,
                                                 double v, double w, double *bp, const KParams &k)
{
    double nm[7], neq[7], q[7], np[7];
    const double Bx = u * Tl, By = v * Tl, Bz = w * Tl;
    double dBx = 0.0, dBy = 0.0, dBz = 0.0;
    if (k.use_g) {
        dBx = Bx - bp[0];
        dBy = By - bp[1];
        dBz = Bz - bp[2];
        bp[0] = Bx;
        bp[1] = By;
        bp[2] = Bz;
    }
    nm[0] = G[0] + G[1] + G[2] + G[3] + G[4] + G[5] + G[6];
    nm[1] = G[1] - G[2];
    nm[2] = G[3] - G[4];
    nm[3] = G[5] - G[6];
    nm[4] = -6.0 * G[0] + G[1] + G[2] + G[3] + G[4] + G[5] + G[6];
    nm[5] = 2.0 * G[1] + 2.0 * G[2] - G[3] - G[4] - G[5] - G[6];
    nm[6] = G[3] + G[4] - G[5] - G[6];
    neq[0] = Tl;
    neq[1] = Bx;
    neq[2] = By;
    neq[3] = Bz;
    neq[4] = k.neqc * Tl;   // -0.75 T (legacy scheme: paraA T, 3DRB.F90:1153-1157)
    neq[5] = 0.0;
    neq[6] = 0.0;
    q[0] = 0.0; q[1] = k.qk; q[2] = k.qk; q[3] = k.qk; q[4] = k.qnu; q[5] = k.qnu; q[6] = k.qnu;
    const double SG = k.sg;
    np[0] = nm[0] - q[0] * (nm[0] - neq[0]);
    np[1] = nm[1] - q[1] * (nm[1] - neq[1]) + SG * dBx;
    np[2] = nm[2] - q[2] * (nm[2] - neq[2]) + SG * dBy;
    np[3] = nm[3] - q[3] * (nm[3] - neq[3]) + SG * dBz;
    np[4] = nm[4] - q[4] * (nm[4] - neq[4]);
    np[5] = nm[5] - q[5] * (nm[5] - neq[5]);
    np[6] = nm[6] - q[6] * (nm[6] - neq[6]);
    P[0] = np[0] / 7.0 - np[4] / 7.0;
    P[1] = np[0] / 7.0 + np[1] / 2.0 + np[4] / 42.0 + np[5] / 6.0;
    P[2] = np[0] / 7.0 - np[1] / 2.0 + np[4] / 42.0 + np[5] / 6.0;
    P[3] = np[0] / 7.0 + np[2] / 2.0 + np[4] / 42.0 - np[5] / 12.0 + np[6] / 4.0;
    P[4] = np[0] / 7.0 - np[2] / 2.0 + np[4] / 42.0 - np[5] / 12.0 + np[6] / 4.0;
    P[5] = np[0] / 7.0 + np[3] / 2.0 + np[4] / 42.0 - np[5] / 12.0 - np[6] / 4.0;
    P[6] = np[0] / 7.0 - np[3] / 2.0 + np[4] / 42.0 - np[5] / 12.0 - np[6] / 4.0;
}

// Reference 2D `collision` (2DRB/2DRB.F90:686-759); returns the stored y force.
__device__ __forceinline__ double ref_collide_d2q9(const double *F, double *P, double rho, double u,
                                                   double v, double T, const KParams &k)
{
    double m[9], meq[9], mp[9], s[9], src[9];
    m[0] = F[0] + F[1] + F[2] + F[3] + F[4] + F[5] + F[6] + F[7] + F[8];
    m[1] = -4.0 * F[0] - F[1] - F[2] - F[3] - F[4] + 2.0 * (F[5] + F[6] + F[7] + F[8]);
    m[2] = 4.0 * F[0] - 2.0 * (F[1] + F[2] + F[3] + F[4]) + F[5] + F[6] + F[7] + F[8];
    m[3] = F[1] - F[3] + F[5] - F[6] - F[7] + F[8];
    m[4] = -2.0 * F[1] + 2.0 * F[3] + F[5] - F[6] - F[7] + F[8];
    m[5] = F[2] - F[4] + F[5] + F[6] - F[7] - F[8];
    m[6] = -2.0 * F[2] + 2.0 * F[4] + F[5] + F[6] - F[7] - F[8];
    m[7] = F[1] - F[2] + F[3] - F[4];
    m[8] = F[5] - F[6] + F[7] - F[8];
    meq[0] = rho;
    meq[1] = rho * (-2.0 + 3.0 * (u * u + v * v));
    meq[2] = rho * (1.0 - 3.0 * (u * u + v * v));
    meq[3] = rho * u;
    meq[4] = -rho * u;
    meq[5] = rho * v;
    meq[6] = -rho * v;
    meq[7] = rho * (u * u - v * v);
    meq[8] = rho * (u * v);
    const double Snu = k.sv, Sq = k.sq;
    s[0] = 0.0; s[1] = Snu; s[2] = Snu; s[3] = 0.0; s[4] = Sq; s[5] = 0.0; s[6] = Sq; s[7] = Snu;
    s[8] = Snu;
    const double Fx = 0.0;
    const double Fy = rho * k.gbeta * (T - k.tref);
    src[0] = 0.0;
    src[1] = (6.0 - 3.0 * s[1]) * (u * Fx + v * Fy);
    src[2] = -(6.0 - 3.0 * s[2]) * (u * Fx + v * Fy);
    src[3] = (1.0 - 0.5 * s[3]) * Fx;
    src[4] = -(1.0 - 0.5 * s[4]) * Fx;
    src[5] = (1.0 - 0.5 * s[5]) * Fy;
    src[6] = -(1.0 - 0.5 * s[6]) * Fy;
    src[7] = (2.0 - s[7]) * (u * Fx - v * Fy);
    src[8] = (1.0 - 0.5 * s[8]) * (u * Fy + v * Fx);
#pragma unroll
    for (int a = 0; a < 9; ++a) mp[a] = m[a] - s[a] * (m[a] - meq[a]) + src[a];
    P[0] = mp[0] / 9.0 - mp[1] / 9.0 + mp[2] / 9.0;
    P[1] = mp[0] / 9.0 - mp[1] / 36.0 - mp[2] / 18.0 + mp[3] / 6.0 - mp[4] / 6.0 + mp[7] / 4.0;
    P[2] = mp[0] / 9.0 - mp[1] / 36.0 - mp[2] / 18.0 + mp[5] / 6.0 - mp[6] / 6.0 - mp[7] / 4.0;
    P[3] = mp[0] / 9.0 - mp[1] / 36.0 - mp[2] / 18.0 - mp[3] / 6.0 + mp[4] / 6.0 + mp[7] / 4.0;
    P[4] = mp[0] / 9.0 - mp[1] / 36.0 - mp[2] / 18.0 - mp[5] / 6.0 + mp[6] / 6.0 - mp[7] / 4.0;
    P[5] = mp[0] / 9.0 + mp[1] / 18.0 + mp[2] / 36.0 + mp[3] / 6.0 + mp[4] / 12.0 + mp[5] / 6.0 +
           mp[6] / 12.0 + mp[8] / 4.0;
    P[6] = mp[0] / 9.0 + mp[1] / 18.0 + mp[2] / 36.0 - mp[3] / 6.0 - mp[4] / 12.0 + mp[5] / 6.0 +
           mp[6] / 12.0 - mp[8] / 4.0;
    P[7] = mp[0] / 9.0 + mp[1] / 18.0 + mp[2] / 36.0 - mp[3] / 6.0 - mp[4] / 12.0 - mp[5] / 6.0 -
           mp[6] / 12.0 + mp[8] / 4.0;
    P[8] = mp[0] / 9.0 + mp[1] / 18.0 + mp[2] / 36.0 + mp[3] / 6.0 + mp[4] / 12.0 - mp[5] / 6.0 -
           mp[6] / 12.0 - mp[8] / 4.0;
    return Fy;
}

// Reference 2D `collisionT` (2DRB/2DRB.F90:917-975).
__device__ __forceinline__ void ref_collide_d2q5(const double *G, double *P, double Tl, double u,
                                                 double v, double *bp, const KParams &k)
{
    double nm[5], neq[5], q[5], np[5];
    const double Bx = u * Tl, By = v * Tl;
    double dBx = 0.0, dBy = 0.0;
    if (k.use_g) {
        dBx = Bx - bp[0];
        dBy = By - bp[1];
        bp[0] = Bx;
        bp[1] = By;
    }
    nm[0] = G[0] + G[1] + G[2] + G[3] + G[4];
    nm[1] = G[1] - G[3];
    nm[2] = G[2] - G[4];
    nm[3] = -4.0 * G[0] + G[1] + G[2] + G[3] + G[4];
    nm[4] = G[1] - G[2] + G[3] - G[4];
    neq[0] = Tl;
    neq[1] = Tl * u;
    neq[2] = Tl * v;
    neq[3] = Tl * k.neqc;   // T (-2/3) (legacy scheme: T paraA, 2DRB.F90:949-953)
    neq[4] = 0.0;
    q[0] = 0.0; q[1] = k.qk; q[2] = k.qk; q[3] = k.qnu; q[4] = k.qnu;
    np[0] = nm[0] - q[0] * (nm[0] - neq[0]);
    np[1] = nm[1] - q[1] * (nm[1] - neq[1]) + k.sg * dBx;
    np[2] = nm[2] - q[2] * (nm[2] - neq[2]) + k.sg * dBy;
    np[3] = nm[3] - q[3] * (nm[3] - neq[3]);
    np[4] = nm[4] - q[4] * (nm[4] - neq[4]);
    P[0] = 0.2 * np[0] - 0.2 * np[3];
    P[1] = 0.2 * np[0] + 0.5 * np[1] + 0.05 * np[3] + 0.25 * np[4];
    P[2] = 0.2 * np[0] + 0.5 * np[2] + 0.05 * np[3] - 0.25 * np[4];
    P[3] = 0.2 * np[0] - 0.5 * np[1] + 0.05 * np[3] + 0.25 * np[4];
    P[4] = 0.2 * np[0] - 0.5 * np[2] + 0.05 * np[3] - 0.25 * np[4];
}

} // namespace rblbm
