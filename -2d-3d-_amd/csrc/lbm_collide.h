// lbm_collide.h -- per-node MRT collision operators (FP64), device inline.
//
// D3Q19 flow: moment transform, relaxation and Guo buoyancy source of `collision`
// (2DRB/3DRB.F90:689-949).  The inverse transform uses the orthogonality of the moment basis:
// M^-1 = M^T diag(1/|row|^2), so f_post = M^T (m_post / |row|^2), evaluated with the row
// structure shared by each population family (rest, axis pairs, diagonal pairs per plane)
// instead of the reference's 19 explicit rows of divisions.
// D3Q7 thermal: `collisionT` (2DRB/3DRB.F90:1106-1198), non-legacy branch, M1G source.
// D2Q9 / D2Q5: 2DRB/2DRB.F90:675-766 and :906-980.
#pragma once

#include "lbm_lattice.h"

namespace rblbm {

// f[19] in, p[19] out (may alias).  rho, u: macros of the previous `macro` sweep; T: current
// temperature; buoyancy F = (0, rho gbeta (T - Tref), 0).
__device__ __forceinline__ void collide_d3q19(const double *f, double *p, double rho, double u,
                                              double v, double w, double T, const KParams &k)
{
    const double f0 = f[0];
    const double sxx = f[1] + f[2], syy = f[3] + f[4], szz = f[5] + f[6];
    const double dxy = f[7] + f[8] + f[9] + f[10];
    const double dxz = f[11] + f[12] + f[13] + f[14];
    const double dyz = f[15] + f[16] + f[17] + f[18];
    const double s1 = sxx + syy + szz, s2 = dxy + dxz + dyz;

    const double xax = f[1] - f[2], yax = f[3] - f[4], zax = f[5] - f[6];
    const double a78 = f[7] - f[8], a910 = f[9] - f[10];        // x-odd, xy plane
    const double b78 = f[7] + f[8], b910 = f[9] + f[10];        // y parity, xy plane
    const double a1112 = f[11] - f[12], a1314 = f[13] - f[14];  // x-odd, xz plane
    const double b1112 = f[11] + f[12], b1314 = f[13] + f[14];
    const double a1516 = f[15] - f[16], a1718 = f[17] - f[18];  // y-odd, yz plane
    const double b1516 = f[15] + f[16], b1718 = f[17] + f[18];

    const double xd = a78 + a910 + a1112 + a1314;
    const double yd = (b78 - b910) + (a1516 + a1718);
    const double zd = (b1112 - b1314) + (b1516 - b1718);

    double m[19];
    m[0] = f0 + s1 + s2;
    m[1] = -30.0 * f0 - 11.0 * s1 + 8.0 * s2;
    m[2] = 12.0 * f0 - 4.0 * s1 + s2;
    m[3] = xax + xd;
    m[4] = -4.0 * xax + xd;
    m[5] = yax + yd;
    m[6] = -4.0 * yax + yd;
    m[7] = zax + zd;
    m[8] = -4.0 * zax + zd;
    m[9] = 2.0 * sxx - syy - szz + dxy + dxz - 2.0 * dyz;
    m[10] = -4.0 * sxx + 2.0 * syy + 2.0 * szz + dxy + dxz - 2.0 * dyz;
    m[11] = syy - szz + dxy - dxz;
    m[12] = -2.0 * syy + 2.0 * szz + dxy - dxz;
    m[13] = a78 - a910;
    m[14] = a1516 - a1718;
    m[15] = a1112 - a1314;
    m[16] = (a78 + a910) - (a1112 + a1314);
    m[17] = -(b78 - b910) + (a1516 + a1718);
    m[18] = (b1112 - b1314) - (b1516 - b1718);

    const double u2 = u * u + v * v + w * w;
    const double sv = k.sv, sq = k.sq;
    const double cv = 1.0 - 0.5 * sv, cq = 1.0 - 0.5 * sq;
    const double Fy = rho * k.gbeta * (T - k.tref);
    const double vF = v * Fy;

    // m_post = m - s (m - m_eq) + (1 - s/2) (d m_eq / d j) . F, scaled by 1/|row|^2
    double h[19];
    h[0] = m[0] * (1.0 / 19.0);
    h[1] = (m[1] - sv * (m[1] - rho * (-11.0 + 19.0 * u2)) + cv * 38.0 * vF) * (1.0 / 2394.0);
    h[2] = (m[2] - sv * (m[2] - rho * (3.0 - 5.5 * u2)) + cv * (-11.0) * vF) * (1.0 / 252.0);
    h[3] = m[3] * 0.1;
    h[4] = (m[4] - sq * (m[4] + (2.0 / 3.0) * rho * u)) * (1.0 / 40.0);
    h[5] = (m[5] + Fy) * 0.1;
    h[6] = (m[6] - sq * (m[6] + (2.0 / 3.0) * rho * v) + cq * (-2.0 / 3.0) * Fy) * (1.0 / 40.0);
    h[7] = m[7] * 0.1;
    h[8] = (m[8] - sq * (m[8] + (2.0 / 3.0) * rho * w)) * (1.0 / 40.0);
    const double pxx = rho * (2.0 * u * u - v * v - w * w);
    const double pww = rho * (v * v - w * w);
    h[9] = (m[9] - sv * (m[9] - pxx) + cv * (-2.0 * vF)) * (1.0 / 36.0);
    h[10] = (m[10] - sv * (m[10] + 0.5 * pxx) + cv * vF) * (1.0 / 72.0);
    h[11] = (m[11] - sv * (m[11] - pww) + cv * (2.0 * vF)) * (1.0 / 12.0);
    h[12] = (m[12] - sv * (m[12] + 0.5 * pww) + cv * (-vF)) * (1.0 / 24.0);
    h[13] = (m[13] - sv * (m[13] - rho * u * v) + cv * (u * Fy)) * 0.25;
    h[14] = (m[14] - sv * (m[14] - rho * v * w) + cv * (w * Fy)) * 0.25;
    h[15] = (m[15] - sv * (m[15] - rho * u * w)) * 0.25;
    h[16] = (m[16] - sq * m[16]) * 0.125;
    h[17] = (m[17] - sq * m[17]) * 0.125;
    h[18] = (m[18] - sq * m[18]) * 0.125;

    // f_post = M^T h
    p[0] = h[0] - 30.0 * h[1] + 12.0 * h[2];
    const double ba = h[0] - 11.0 * h[1] - 4.0 * h[2];
    const double qx = h[3] - 4.0 * h[4], qy = h[5] - 4.0 * h[6], qz = h[7] - 4.0 * h[8];
    const double pxa = 2.0 * h[9] - 4.0 * h[10];
    const double pya = -h[9] + 2.0 * h[10];
    const double pwa = h[11] - 2.0 * h[12];
    p[1] = ba + qx + pxa;
    p[2] = ba - qx + pxa;
    p[3] = ba + qy + pya + pwa;
    p[4] = ba - qy + pya + pwa;
    p[5] = ba + qz + pya - pwa;
    p[6] = ba - qz + pya - pwa;
    const double bd = h[0] + 8.0 * h[1] + h[2];
    const double X = h[3] + h[4], Y = h[5] + h[6], Z = h[7] + h[8];
    const double pxy = bd + (h[9] + h[10]) + (h[11] + h[12]);
    const double pxz = bd + (h[9] + h[10]) - (h[11] + h[12]);
    const double pyz = bd - 2.0 * (h[9] + h[10]);
    p[7] = pxy + X + Y + h[13] + (h[16] - h[17]);
    p[8] = pxy - X + Y - h[13] - (h[16] + h[17]);
    p[9] = pxy + X - Y - h[13] + (h[16] + h[17]);
    p[10] = pxy - X - Y + h[13] - (h[16] - h[17]);
    p[11] = pxz + X + Z + h[15] - (h[16] - h[18]);
    p[12] = pxz - X + Z - h[15] + (h[16] + h[18]);
    p[13] = pxz + X - Z - h[15] - (h[16] + h[18]);
    p[14] = pxz - X - Z + h[15] + (h[16] - h[18]);
    p[15] = pyz + Y + Z + h[14] + (h[17] - h[18]);
    p[16] = pyz - Y + Z - h[14] - (h[17] + h[18]);
    p[17] = pyz + Y - Z - h[14] + (h[17] + h[18]);
    p[18] = pyz - Y - Z + h[14] - (h[17] - h[18]);
}

// g[7] in, p[7] out.  T: temperature of the previous macroT (equals sum g except on the first
// step, where the reference uses its initial T array); B = u T with the new u; bp: B_prev,
// updated in place when use_g.
__device__ __forceinline__ void collide_d3q7(const double *g, double *p, double T, double u,
                                             double v, double w, double *bp, const KParams &k)
{
    const double Bx = u * T, By = v * T, Bz = w * T;
    double dBx = 0.0, dBy = 0.0, dBz = 0.0;
    if (k.use_g) {
        dBx = Bx - bp[0];
        dBy = By - bp[1];
        dBz = Bz - bp[2];
        bp[0] = Bx;
        bp[1] = By;
        bp[2] = Bz;
    }
    const double n0 = g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6];
    const double n1 = g[1] - g[2], n2 = g[3] - g[4], n3 = g[5] - g[6];
    const double s16 = g[1] + g[2] + g[3] + g[4] + g[5] + g[6];
    const double n4 = -6.0 * g[0] + s16;
    const double n5 = 2.0 * (g[1] + g[2]) - (g[3] + g[4] + g[5] + g[6]);
    const double n6 = (g[3] + g[4]) - (g[5] + g[6]);
    const double qk = k.qk, qn = k.qnu, sg = k.sg;
    const double p0 = n0;
    const double p1 = n1 - qk * (n1 - Bx) + sg * dBx;
    const double p2 = n2 - qk * (n2 - By) + sg * dBy;
    const double p3 = n3 - qk * (n3 - Bz) + sg * dBz;
    const double p4 = n4 - qn * (n4 + 0.75 * T);
    const double p5 = n5 - qn * n5;
    const double p6 = n6 - qn * n6;
    const double c0 = p0 * (1.0 / 7.0);
    const double c4 = p4 * (1.0 / 42.0);
    p[0] = c0 - p4 * (1.0 / 7.0);
    const double ax = c0 + c4 + p5 * (1.0 / 6.0);
    const double ay = c0 + c4 - p5 * (1.0 / 12.0);
    p[1] = ax + 0.5 * p1;
    p[2] = ax - 0.5 * p1;
    p[3] = ay + 0.5 * p2 + 0.25 * p6;
    p[4] = ay - 0.5 * p2 + 0.25 * p6;
    p[5] = ay + 0.5 * p3 - 0.25 * p6;
    p[6] = ay - 0.5 * p3 - 0.25 * p6;
}

// 2D flow collision, 2DRB.F90:675-766.  Returns the stored y force (old rho, old T).
__device__ __forceinline__ double collide_d2q9(const double *f, double *p, double rho, double u,
                                               double v, double T, const KParams &k)
{
    double m[9];
    m[0] = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    m[1] = -4.0 * f[0] - f[1] - f[2] - f[3] - f[4] + 2.0 * (f[5] + f[6] + f[7] + f[8]);
    m[2] = 4.0 * f[0] - 2.0 * (f[1] + f[2] + f[3] + f[4]) + f[5] + f[6] + f[7] + f[8];
    m[3] = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
    m[4] = -2.0 * f[1] + 2.0 * f[3] + f[5] - f[6] - f[7] + f[8];
    m[5] = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    m[6] = -2.0 * f[2] + 2.0 * f[4] + f[5] + f[6] - f[7] - f[8];
    m[7] = f[1] - f[2] + f[3] - f[4];
    m[8] = f[5] - f[6] + f[7] - f[8];
    const double uu = u * u + v * v;
    const double meq[9] = {rho, rho * (-2.0 + 3.0 * uu), rho * (1.0 - 3.0 * uu), rho * u, -rho * u,
                           rho * v, -rho * v, rho * (u * u - v * v), rho * (u * v)};
    const double sv = k.sv, sq = k.sq;
    const double s[9] = {0.0, sv, sv, 0.0, sq, 0.0, sq, sv, sv};
    const double Fy = rho * k.gbeta * (T - k.tref);
    const double Fx = 0.0;
    double src[9];
    src[0] = 0.0;
    src[1] = (6.0 - 3.0 * s[1]) * (u * Fx + v * Fy);
    src[2] = -(6.0 - 3.0 * s[2]) * (u * Fx + v * Fy);
    src[3] = (1.0 - 0.5 * s[3]) * Fx;
    src[4] = -(1.0 - 0.5 * s[4]) * Fx;
    src[5] = (1.0 - 0.5 * s[5]) * Fy;
    src[6] = -(1.0 - 0.5 * s[6]) * Fy;
    src[7] = (2.0 - s[7]) * (u * Fx - v * Fy);
    src[8] = (1.0 - 0.5 * s[8]) * (u * Fy + v * Fx);
    double mp[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) mp[a] = m[a] - s[a] * (m[a] - meq[a]) + src[a];
    const double r9 = 1.0 / 9.0, r36 = 1.0 / 36.0, r18 = 1.0 / 18.0, r6 = 1.0 / 6.0, r12 = 1.0 / 12.0;
    p[0] = mp[0] * r9 - mp[1] * r9 + mp[2] * r9;
    p[1] = mp[0] * r9 - mp[1] * r36 - mp[2] * r18 + mp[3] * r6 - mp[4] * r6 + mp[7] * 0.25;
    p[2] = mp[0] * r9 - mp[1] * r36 - mp[2] * r18 + mp[5] * r6 - mp[6] * r6 - mp[7] * 0.25;
    p[3] = mp[0] * r9 - mp[1] * r36 - mp[2] * r18 - mp[3] * r6 + mp[4] * r6 + mp[7] * 0.25;
    p[4] = mp[0] * r9 - mp[1] * r36 - mp[2] * r18 - mp[5] * r6 + mp[6] * r6 - mp[7] * 0.25;
    p[5] = mp[0] * r9 + mp[1] * r18 + mp[2] * r36 + mp[3] * r6 + mp[4] * r12 + mp[5] * r6 +
           mp[6] * r12 + mp[8] * 0.25;
    p[6] = mp[0] * r9 + mp[1] * r18 + mp[2] * r36 - mp[3] * r6 - mp[4] * r12 + mp[5] * r6 +
           mp[6] * r12 - mp[8] * 0.25;
    p[7] = mp[0] * r9 + mp[1] * r18 + mp[2] * r36 - mp[3] * r6 - mp[4] * r12 - mp[5] * r6 -
           mp[6] * r12 + mp[8] * 0.25;
    p[8] = mp[0] * r9 + mp[1] * r18 + mp[2] * r36 + mp[3] * r6 + mp[4] * r12 - mp[5] * r6 -
           mp[6] * r12 - mp[8] * 0.25;
    return Fy;
}

// 2D thermal collision, 2DRB.F90:906-980.  bp: B_prev (2 values), updated when use_g.
__device__ __forceinline__ void collide_d2q5(const double *g, double *p, double T, double u,
                                             double v, double *bp, const KParams &k)
{
    const double Bx = u * T, By = v * T;
    double dBx = 0.0, dBy = 0.0;
    if (k.use_g) {
        dBx = Bx - bp[0];
        dBy = By - bp[1];
        bp[0] = Bx;
        bp[1] = By;
    }
    const double n0 = g[0] + g[1] + g[2] + g[3] + g[4];
    const double n1 = g[1] - g[3], n2 = g[2] - g[4];
    const double n3 = -4.0 * g[0] + g[1] + g[2] + g[3] + g[4];
    const double n4 = g[1] - g[2] + g[3] - g[4];
    const double qk = k.qk, qn = k.qnu;
    const double p0 = n0;
    const double p1 = n1 - qk * (n1 - T * u) + k.sg * dBx;
    const double p2 = n2 - qk * (n2 - T * v) + k.sg * dBy;
    const double p3 = n3 - qn * (n3 - T * (-2.0 / 3.0));
    const double p4 = n4 - qn * n4;
    p[0] = 0.2 * p0 - 0.2 * p3;
    p[1] = 0.2 * p0 + 0.5 * p1 + 0.05 * p3 + 0.25 * p4;
    p[2] = 0.2 * p0 + 0.5 * p2 + 0.05 * p3 - 0.25 * p4;
    p[3] = 0.2 * p0 - 0.5 * p1 + 0.05 * p3 + 0.25 * p4;
    p[4] = 0.2 * p0 - 0.5 * p2 + 0.05 * p3 - 0.25 * p4;
}

} // namespace rblbm
