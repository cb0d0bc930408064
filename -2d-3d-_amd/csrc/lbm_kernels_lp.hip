// lbm_kernels_lp.hip -- reduced-precision storage forms of the 3D step (FP32 arithmetic).
//
// Populations are stored as deviations from the rest equilibrium, which keeps their magnitude
// near the signal and away from the constant weight:
//   h_a = f_a - w_a rho0 (rho0 = 1)            k_a = g_a - wT_a T0(x)
// T0(x) is the conduction profile between the two isothermal walls (linear along the wall
// normal; a constant when no axis has two isothermal walls).  Near a wall the temperature then
// differs from T0 by the thermal boundary layer's deviation only, so the few-ulp rounding of a
// stored population is measured against that deviation instead of against T itself: the wall
// heat flux (a difference of nearly equal temperatures) keeps its precision.  T0 is a fixed
// linear function whose slope and offset carry few mantissa bits, so T0 and every difference
// of T0 between neighbours are exact in FP32; a pulled deviation is corrected by
// wT_a (T0(x - e_a) - T0(x)) = -wT_a (e_a . n) s, exactly (wT_a is a power of two).
// LBM_FP32_SHIFT stores h, k as float; LBM_FP16_SHIFT stores h_a / w_a and k_a / wT_a as
// _Float16 (per-direction scaling puts every direction in the same numeric range).  All
// collision arithmetic is FP32 (FMA contraction on); the history B_prev and the stored previous
// temperature are FP32.  Global sums stay FP64 (lbm_kernels.hip reductions run on FP64 macros
// produced here).
//
// The moment-space collision is the reference's (`collision`, 2DRB/3DRB.F90:689-949;
// `collisionT`, :1106-1198) rewritten for deviations: with M w = (1, -11, 3, 0, ...) and
// N wT = (1, 0, 0, 0, -3/4, 0, 0), m' = M h, and the equilibria of the constant part cancel,
// leaving m'_eq(0) = drho, m'_eq(1) = -11 drho + 19 rho u^2, m'_eq(2) = 3 drho - 11/2 rho u^2,
// n'_eq(4) = -3/4 dT; every other equilibrium moment is unchanged.  Walls act on deviations
// the same way (bounce-back and periodic copies are linear and w is symmetric); anti-bounce-back
// becomes k_in = -k_out(opp) + 2 wT (T_wall - T0).
#include <utility>

#include "lbm_internal.h"
#include "lbm_pull.h"

namespace rblbm {

// Storage policy: load/store one population of direction A, converting to/from FP32 deviations.
template <typename S> struct Pop;
template <> struct Pop<float> {
    template <int Q, int A> static __device__ __forceinline__ float cvt(float v) { return v; }
    template <int Q, int A> static __device__ __forceinline__ float ld(const float *p) { return *p; }
    template <int Q, int A> static __device__ __forceinline__ void st(float *p, float v) { *p = v; }
};
template <int Q> __host__ __device__ constexpr double wq(int a)
{
    return Q == 19 ? w19(a) : (Q == 7 ? w7(a) : (Q == 9 ? w9(a) : w5(a)));
}
template <> struct Pop<_Float16> {
    // stored value = deviation / w_a
    template <int Q, int A> static __device__ __forceinline__ float cvt(_Float16 v)
    {
        return (float)v * (float)wq<Q>(A);
    }
    template <int Q, int A> static __device__ __forceinline__ float ld(const _Float16 *p)
    {
        return cvt<Q, A>(*p);
    }
    template <int Q, int A> static __device__ __forceinline__ void st(_Float16 *p, float v)
    {
        *p = (_Float16)(v * (float)(1.0 / wq<Q>(A)));
    }
};

struct LPParams {
    float gbeta, sv, sq, qk, qnu, sg;
    float thot, tcold;          // wall temperatures
    float t0a, t0s;             // storage shift T0 = t0a + t0s * (node index along t0axis)
    int t0axis;                 // 0: T0 = t0a everywhere; 1: along x; 2: along y
    double tref;                // reference temperature of the buoyancy force
};

// Storage shift T0 at node (i, j) (exact in FP32, see above).
__device__ __forceinline__ float t0_of(const LPParams &lp, int i, int j)
{
    return lp.t0axis == 1 ? lp.t0a + lp.t0s * (float)i
                          : (lp.t0axis == 2 ? lp.t0a + lp.t0s * (float)j : lp.t0a);
}
// Correction of a pulled thermal deviation of direction A: wT_A (T0(x - e_A) - T0(x)).
template <int A> __device__ __forceinline__ float t0_pull(const LPParams &lp)
{
    const int e = lp.t0axis == 1 ? D3Q7::ex[A] : (lp.t0axis == 2 ? D3Q7::ey[A] : 0);
    return -(float)w7(A) * ((float)e * lp.t0s);
}

// -------------------------------------------------------------------------------------------
// pulls (walls as in lbm_kernels.hip: last outside axis decides for the flow lattice)
// -------------------------------------------------------------------------------------------
// Incoming thermal deviation of direction A from the value v read at heat_src (t0 = T0(i, j)).
template <class L, int A>
__device__ __forceinline__ float heat_fix(float v, unsigned mode, const LPParams &lp, float t0)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    if (mode == PULL_T0) return v + t0_pull<A>(lp);
    if (mode == PULL_BACK) return v;
    constexpr bool low_wall = (ex + ey + ez) > 0;
    const float tw = (low_wall ? lp.thot : lp.tcold) - t0;
    return -v + 2.0f * (float)wq<L::Q>(A) * tw;
}

// Pulls (the weights of A and of its opposite are equal, so one FP16 scale serves both).
template <class L, typename S, int A>
__device__ __forceinline__ float flow_in(const S *__restrict__ P, uint64_t N, int i, int j, int k,
                                         uint64_t c, bool border, const KParams &kp)
{
    return Pop<S>::template ld<L::Q, A>(flow_src<L, S, A>(P, N, i, j, k, c, border, kp));
}

template <class L, typename S, int A>
__device__ __forceinline__ float heat_in(const S *__restrict__ P, uint64_t N, int i, int j, int k,
                                         uint64_t c, bool border, const KParams &kp,
                                         const LPParams &lp)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    if (!border)
        return Pop<S>::template ld<L::Q, A>(P + A * N + c - (ex + ey * kp.nx + ez * kp.nx * kp.ny)) +
               t0_pull<A>(lp);
    unsigned mode;
    const S *src = heat_src<L, S, A>(P, N, i, j, k, c, true, kp, mode);
    return heat_fix<L, A>(Pop<S>::template ld<L::Q, A>(src), mode, lp, t0_of(lp, i, j));
}

// Staged pulls of the one-pass kernel (4-byte storage): every active lane loads the element its
// pull reads straight into LDS, direction A to slot dst[A * pitch + lane] (dst: the wave's first
// slot; the LDS address is the wave's, the lane offset is implicit).
template <class L, typename S, int... A>
__device__ __forceinline__ void stage_flow(const S *P, uint64_t N, int i, int j, int k, uint64_t c,
                                           bool border, const KParams &kp, S *dst, int pitch,
                                           std::integer_sequence<int, A...>)
{
    static_assert(sizeof(S) == 4, "staged loads move 4-byte elements");
    (__builtin_amdgcn_global_load_lds(flow_src<L, S, A>(P, N, i, j, k, c, border, kp),
                                      dst + A * pitch, 4, 0, 0), ...);
}
template <class L, typename S, int... A>
__device__ __forceinline__ unsigned stage_heat(const S *P, uint64_t N, int i, int j, int k,
                                               uint64_t c, bool border, const KParams &kp, S *dst,
                                               int pitch, std::integer_sequence<int, A...>)
{
    static_assert(sizeof(S) == 4, "staged loads move 4-byte elements");
    unsigned modes = 0, m;
    ((__builtin_amdgcn_global_load_lds(heat_src<L, S, A>(P, N, i, j, k, c, border, kp, m),
                                       dst + A * pitch, 4, 0, 0),
      modes |= m << (2 * A)), ...);
    return modes;
}
template <class L, int... A>
__device__ __forceinline__ void unstage_heat(const float *slot, int pitch, unsigned modes,
                                             const LPParams &lp, float t0, float *g,
                                             std::integer_sequence<int, A...>)
{
    ((g[A] = heat_fix<L, A>(slot[A * pitch], (modes >> (2 * A)) & 3u, lp, t0)), ...);
}

// Pulls split at the load (the one-pass kernel's prefetching into registers): the stored
// elements at the pulls' source addresses, and their conversion to the incoming FP32 values
// once they have arrived.  raw + cook give what flow_in / heat_in give.
template <class L, typename S, int... A>
__device__ __forceinline__ void raw_flow(const S *P, uint64_t N, int i, int j, int k, uint64_t c,
                                         bool border, const KParams &kp, S *f,
                                         std::integer_sequence<int, A...>)
{
    ((f[A] = P[flow_elem<L, A>(N, i, j, k, c, kp)]), ...);
}
template <class L, typename S, int... A>
__device__ __forceinline__ unsigned raw_heat(const S *P, uint64_t N, int i, int j, int k,
                                             uint64_t c, bool border, const KParams &kp, S *g,
                                             std::integer_sequence<int, A...>)
{
    unsigned modes = 0, m;
    ((g[A] = P[heat_elem<L, A>(N, i, j, k, c, kp, m)], modes |= m << (2 * A)), ...);
    return modes;
}
template <class L, typename S, int... A>
__device__ __forceinline__ void cook_flow(const S *raw, float *f, std::integer_sequence<int, A...>)
{
    ((f[A] = Pop<S>::template cvt<L::Q, A>(raw[A])), ...);
}
template <class L, typename S, int... A>
__device__ __forceinline__ void cook_heat(const S *raw, unsigned modes, const LPParams &lp, float t0,
                                          float *g, std::integer_sequence<int, A...>)
{
    ((g[A] = heat_fix<L, A>(Pop<S>::template cvt<L::Q, A>(raw[A]), (modes >> (2 * A)) & 3u, lp, t0)),
     ...);
}

// In-place (A-A) layout (LAY_ODD / LAY_EVEN, lbm_lattice.h): the same incoming values as the
// pulls above, with every face a wall.
__device__ __forceinline__ bool inside3_lp(int i, int j, int k, const KParams &kp)
{
    return (unsigned)i < (unsigned)kp.nx && (unsigned)j < (unsigned)kp.ny && (unsigned)k < (unsigned)kp.nz;
}

template <class L, typename S, int LAY, int A>
__device__ __forceinline__ float flow_in_aa(const S *P, uint64_t N, int i, int j, int k, uint64_t c,
                                            bool border, const KParams &kp)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    if (LAY == LAY_ODD) return Pop<S>::template ld<L::Q, A>(P + A * N + c);
    const int64_t off = ex + ey * kp.nx + ez * (int64_t)kp.nx * kp.ny;
    const bool in = !border || inside3_lp(i - ex, j - ey, k - ez, kp);
    return Pop<S>::template ld<L::Q, A>(P + (in ? L::opp[A] * N + c - off : A * N + c));
}

template <class L, typename S, int LAY, int A>
__device__ __forceinline__ float heat_in_aa(const S *P, uint64_t N, int i, int j, int k, uint64_t c,
                                            bool border, const KParams &kp, const LPParams &lp)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    int kind = -1;
    if (border) {
        if (ex != 0) {
            if ((unsigned)(i - ex) >= (unsigned)kp.nx) kind = kp.temp_x;
        } else if (ey != 0) {
            if ((unsigned)(j - ey) >= (unsigned)kp.ny) kind = kp.temp_y;
        } else if (ez != 0) {
            if ((unsigned)(k - ez) >= (unsigned)kp.nz) kind = kp.temp_z;
        }
    }
    if (kind < 0) {
        const int64_t off = ex + ey * kp.nx + ez * (int64_t)kp.nx * kp.ny;
        const float v = LAY == LAY_ODD ? Pop<S>::template ld<L::Q, A>(P + A * N + c)
                                       : Pop<S>::template ld<L::Q, A>(P + L::opp[A] * N + c - off);
        return v + t0_pull<A>(lp);
    }
    // g_post of the opposite direction, in both states
    const float back = Pop<S>::template ld<L::Q, L::opp[A]>(P + A * N + c);
    if (kind == T_ADIABATIC) return back;
    constexpr bool low_wall = (ex + ey + ez) > 0;
    const float tw = (low_wall ? lp.thot : lp.tcold) - t0_of(lp, i, j);
    return -back + 2.0f * (float)wq<L::Q>(A) * tw;
}

template <class L, typename S, int LAY, int A>
__device__ __forceinline__ void store_aa(S *P, uint64_t N, int i, int j, int k, uint64_t c,
                                         bool border, const KParams &kp, float v)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    if (LAY == LAY_ODD) {
        Pop<S>::template st<L::Q, A>(P + L::opp[A] * N + c, v);
        return;
    }
    const int64_t off = ex + ey * kp.nx + ez * (int64_t)kp.nx * kp.ny;
    const bool in = !border || inside3_lp(i + ex, j + ey, k + ez, kp);
    Pop<S>::template st<L::Q, A>(P + (in ? A * N + c + off : L::opp[A] * N + c), v);
}

// Incoming populations in layout LAY and the store of the post-collision ones.
template <class L, typename S, int LAY, int... A>
__device__ __forceinline__ void in_flow_lp(const S *P, uint64_t N, int i, int j, int k, uint64_t c,
                                           bool border, const KParams &kp, float *f,
                                           std::integer_sequence<int, A...>)
{
    if constexpr (LAY == LAY_PULL)
        ((f[A] = flow_in<L, S, A>(P, N, i, j, k, c, border, kp)), ...);
    else
        ((f[A] = flow_in_aa<L, S, LAY, A>(P, N, i, j, k, c, border, kp)), ...);
}
template <class L, typename S, int LAY, int... A>
__device__ __forceinline__ void in_heat_lp(const S *P, uint64_t N, int i, int j, int k, uint64_t c,
                                           bool border, const KParams &kp, const LPParams &lp,
                                           float *g, std::integer_sequence<int, A...>)
{
    if constexpr (LAY == LAY_PULL)
        ((g[A] = heat_in<L, S, A>(P, N, i, j, k, c, border, kp, lp)), ...);
    else
        ((g[A] = heat_in_aa<L, S, LAY, A>(P, N, i, j, k, c, border, kp, lp)), ...);
}
template <class L, typename S, int LAY, int... A>
__device__ __forceinline__ void out_lp(S *P, uint64_t N, int i, int j, int k, uint64_t c,
                                       bool border, const KParams &kp, const float *v,
                                       std::integer_sequence<int, A...>)
{
    if constexpr (LAY == LAY_PULL)
        (Pop<S>::template st<L::Q, A>(P + A * N + c, v[A]), ...);
    else
        (store_aa<L, S, LAY, A>(P, N, i, j, k, c, border, kp, v[A]), ...);
}
constexpr int lay_flip_lp(int lay) { return lay == LAY_PULL ? LAY_PULL : 3 - lay; }

template <class L, typename S, int... A>
__device__ __forceinline__ void pull_heat_lp(const S *__restrict__ P, uint64_t N, int i, int j, int k,
                                             uint64_t c, bool border, const KParams &kp,
                                             const LPParams &lp, float *g,
                                             std::integer_sequence<int, A...>)
{
    ((g[A] = heat_in<L, S, A>(P, N, i, j, k, c, border, kp, lp)), ...);
}
template <class L, typename S, int... A>
__device__ __forceinline__ void store_all(S *__restrict__ P, uint64_t N, uint64_t c, const float *v,
                                          std::integer_sequence<int, A...>)
{
    (Pop<S>::template st<L::Q, A>(P + A * N + c, v[A]), ...);
}
template <class L, typename S, int... A>
__device__ __forceinline__ void load_all(const S *__restrict__ P, uint64_t N, uint64_t c, float *v,
                                         std::integer_sequence<int, A...>)
{
    ((v[A] = Pop<S>::template ld<L::Q, A>(P + A * N + c)), ...);
}
#define SEQ(L) std::make_integer_sequence<int, L::Q>{}

template <typename S, int Q, int A>
__device__ __forceinline__ float ld_at(const S *P, int64_t off, uint32_t vo)
{
    return Pop<S>::template ld<Q, A>((const S *)((const char *)P + off + vo));
}
template <typename S, int Q, int A>
__device__ __forceinline__ void st_at(S *P, int64_t off, uint32_t vo, float v)
{
    Pop<S>::template st<Q, A>((S *)((char *)P + off + vo), v);
}
// Offset rows in layout LAY (two arrays: pulls; in place: the slots of state LAY).
template <typename S, int LAY, int... A>
__device__ __forceinline__ void fast_flow_in(const S *P, const PullOffsets &pl, uint32_t vo, int i, int j,
                                             int k, uint64_t c, const KParams &kp, float *f,
                                             std::integer_sequence<int, A...>)
{
    ((f[A] = ld_at<S, 19, A>(P, LAY == LAY_PULL ? pl.f[A] : (LAY == LAY_ODD ? pl.s[A] : pl.fe[A]), vo)), ...);
    // x-wall nodes redo the pulls whose source lies beyond their wall (ex = +1 at i = 0,
    // ex = -1 at i = nx - 1) with the wall rules; a wave holds one of the two walls
    auto redo = [&](auto sgn) {
        constexpr int X = decltype(sgn)::value;
        if constexpr (LAY == LAY_PULL)
            ((f[A] = D3Q19::ex[A] == X ? flow_in<D3Q19, S, A>(P, kp.P, i, j, k, c, true, kp) : f[A]), ...);
        else
            ((f[A] = D3Q19::ex[A] == X ? flow_in_aa<D3Q19, S, LAY, A>(P, kp.P, i, j, k, c, true, kp) : f[A]), ...);
    };
    if (i == 0) redo(std::integral_constant<int, 1>{});
    if (i == kp.nx - 1) redo(std::integral_constant<int, -1>{});
}
template <typename S, int LAY, int... A>
__device__ __forceinline__ void fast_heat_in(const S *P, const PullOffsets &pl, uint32_t vo, int i, int j,
                                             int k, uint64_t c, const KParams &kp,
                                             const LPParams &lp, float *g,
                                             std::integer_sequence<int, A...>)
{
    ((g[A] = ld_at<S, 7, A>(P, LAY == LAY_PULL ? pl.g[A] : (LAY == LAY_ODD ? pl.s[A] : pl.ge[A]), vo) +
             t0_pull<A>(lp)), ...);
    auto redo = [&](auto sgn) {   // as in fast_flow_in
        constexpr int X = decltype(sgn)::value;
        if constexpr (LAY == LAY_PULL)
            ((g[A] = D3Q7::ex[A] == X ? heat_in<D3Q7, S, A>(P, kp.P, i, j, k, c, true, kp, lp) : g[A]), ...);
        else
            ((g[A] = D3Q7::ex[A] == X ? heat_in_aa<D3Q7, S, LAY, A>(P, kp.P, i, j, k, c, true, kp, lp) : g[A]), ...);
    };
    if (i == 0) redo(std::integral_constant<int, 1>{});
    if (i == kp.nx - 1) redo(std::integral_constant<int, -1>{});
}
template <class L, typename S, int LAY, int... A>
__device__ __forceinline__ void fast_out(S *P, const PullOffsets &pl, uint32_t vo, int i, int j, int k,
                                         uint64_t c, const KParams &kp, const float *v,
                                         std::integer_sequence<int, A...>)
{
    if (LAY == LAY_EVEN && (i == 0 || i == kp.nx - 1)) {   // populations leaving through x walls
        (store_aa<L, S, LAY, A>(P, kp.P, i, j, k, c, true, kp, v[A]), ...);
        return;
    }
    constexpr bool F = L::Q == 19;
    (st_at<S, L::Q, A>(P, LAY == LAY_PULL ? pl.s[A] : (LAY == LAY_ODD ? pl.s[L::opp[A]] : (F ? pl.fw[A] : pl.gw[A])),
                       vo, v[A]), ...);
}

// -------------------------------------------------------------------------------------------
// deviation-form collisions (FP32)
// -------------------------------------------------------------------------------------------
__device__ __forceinline__ void collide19_dev(const float *f, float *p, float drho, float u, float v,
                                              float w, float dT, const LPParams &k)
{
    const float rho = 1.0f + drho;
    const float f0 = f[0];
    const float sxx = f[1] + f[2], syy = f[3] + f[4], szz = f[5] + f[6];
    const float dxy = (f[7] + f[8]) + (f[9] + f[10]);
    const float dxz = (f[11] + f[12]) + (f[13] + f[14]);
    const float dyz = (f[15] + f[16]) + (f[17] + f[18]);
    const float s1 = sxx + syy + szz, s2 = dxy + dxz + dyz;
    const float xax = f[1] - f[2], yax = f[3] - f[4], zax = f[5] - f[6];
    const float a78 = f[7] - f[8], a910 = f[9] - f[10], b78 = f[7] + f[8], b910 = f[9] + f[10];
    const float a1112 = f[11] - f[12], a1314 = f[13] - f[14], b1112 = f[11] + f[12], b1314 = f[13] + f[14];
    const float a1516 = f[15] - f[16], a1718 = f[17] - f[18], b1516 = f[15] + f[16], b1718 = f[17] + f[18];
    const float xd = (a78 + a910) + (a1112 + a1314);
    const float yd = (b78 - b910) + (a1516 + a1718);
    const float zd = (b1112 - b1314) + (b1516 - b1718);
    float m[19];
    m[1] = -30.0f * f0 - 11.0f * s1 + 8.0f * s2;
    m[2] = 12.0f * f0 - 4.0f * s1 + s2;
    m[3] = xax + xd;
    m[4] = -4.0f * xax + xd;
    m[5] = yax + yd;
    m[6] = -4.0f * yax + yd;
    m[7] = zax + zd;
    m[8] = -4.0f * zax + zd;
    m[9] = 2.0f * sxx - syy - szz + dxy + dxz - 2.0f * dyz;
    m[10] = -4.0f * sxx + 2.0f * syy + 2.0f * szz + dxy + dxz - 2.0f * dyz;
    m[11] = syy - szz + dxy - dxz;
    m[12] = -2.0f * syy + 2.0f * szz + dxy - dxz;
    m[13] = a78 - a910;
    m[14] = a1516 - a1718;
    m[15] = a1112 - a1314;
    m[16] = (a78 + a910) - (a1112 + a1314);
    m[17] = -(b78 - b910) + (a1516 + a1718);
    m[18] = (b1112 - b1314) - (b1516 - b1718);

    const float u2 = u * u + v * v + w * w;
    const float sv = k.sv, sq = k.sq;
    const float cv = 1.0f - 0.5f * sv, cq = 1.0f - 0.5f * sq;
    const float Fy = rho * k.gbeta * dT;
    const float vF = v * Fy;
    float h[19];
    h[0] = drho * (1.0f / 19.0f);
    h[1] = (m[1] - sv * (m[1] - (-11.0f * drho + 19.0f * rho * u2)) + cv * 38.0f * vF) * (1.0f / 2394.0f);
    h[2] = (m[2] - sv * (m[2] - (3.0f * drho - 5.5f * rho * u2)) + cv * (-11.0f) * vF) * (1.0f / 252.0f);
    h[3] = m[3] * 0.1f;
    h[4] = (m[4] - sq * (m[4] + (2.0f / 3.0f) * rho * u)) * (1.0f / 40.0f);
    h[5] = (m[5] + Fy) * 0.1f;
    h[6] = (m[6] - sq * (m[6] + (2.0f / 3.0f) * rho * v) + cq * (-2.0f / 3.0f) * Fy) * (1.0f / 40.0f);
    h[7] = m[7] * 0.1f;
    h[8] = (m[8] - sq * (m[8] + (2.0f / 3.0f) * rho * w)) * (1.0f / 40.0f);
    const float pxx = rho * (2.0f * u * u - v * v - w * w);
    const float pww = rho * (v * v - w * w);
    h[9] = (m[9] - sv * (m[9] - pxx) + cv * (-2.0f * vF)) * (1.0f / 36.0f);
    h[10] = (m[10] - sv * (m[10] + 0.5f * pxx) + cv * vF) * (1.0f / 72.0f);
    h[11] = (m[11] - sv * (m[11] - pww) + cv * (2.0f * vF)) * (1.0f / 12.0f);
    h[12] = (m[12] - sv * (m[12] + 0.5f * pww) + cv * (-vF)) * (1.0f / 24.0f);
    h[13] = (m[13] - sv * (m[13] - rho * u * v) + cv * (u * Fy)) * 0.25f;
    h[14] = (m[14] - sv * (m[14] - rho * v * w) + cv * (w * Fy)) * 0.25f;
    h[15] = (m[15] - sv * (m[15] - rho * u * w)) * 0.25f;
    h[16] = (1.0f - sq) * m[16] * 0.125f;
    h[17] = (1.0f - sq) * m[17] * 0.125f;
    h[18] = (1.0f - sq) * m[18] * 0.125f;

    p[0] = h[0] - 30.0f * h[1] + 12.0f * h[2];
    const float ba = h[0] - 11.0f * h[1] - 4.0f * h[2];
    const float qx = h[3] - 4.0f * h[4], qy = h[5] - 4.0f * h[6], qz = h[7] - 4.0f * h[8];
    const float pxa = 2.0f * h[9] - 4.0f * h[10];
    const float pya = -h[9] + 2.0f * h[10];
    const float pwa = h[11] - 2.0f * h[12];
    p[1] = ba + qx + pxa;
    p[2] = ba - qx + pxa;
    p[3] = ba + qy + pya + pwa;
    p[4] = ba - qy + pya + pwa;
    p[5] = ba + qz + pya - pwa;
    p[6] = ba - qz + pya - pwa;
    const float bd = h[0] + 8.0f * h[1] + h[2];
    const float X = h[3] + h[4], Y = h[5] + h[6], Z = h[7] + h[8];
    const float pxy = bd + (h[9] + h[10]) + (h[11] + h[12]);
    const float pxz = bd + (h[9] + h[10]) - (h[11] + h[12]);
    const float pyz = bd - 2.0f * (h[9] + h[10]);
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

// g deviations k[7] -> k_post[7]; dT the temperature deviation used by collisionT (T - T0);
// Tfull = T0 + dT in FP32; bp the M1G history.
__device__ __forceinline__ void collide7_dev(const float *g, float *p, float dT, float Tfull, float u,
                                             float v, float w, float *bp, int use_g, const LPParams &k)
{
    const float Bx = u * Tfull, By = v * Tfull, Bz = w * Tfull;
    float dBx = 0.f, dBy = 0.f, dBz = 0.f;
    if (use_g) {
        dBx = Bx - bp[0]; dBy = By - bp[1]; dBz = Bz - bp[2];
        bp[0] = Bx; bp[1] = By; bp[2] = Bz;
    }
    const float n0 = (g[0] + g[1]) + (g[2] + g[3]) + (g[4] + g[5]) + g[6];
    const float n1 = g[1] - g[2], n2 = g[3] - g[4], n3 = g[5] - g[6];
    const float s16 = (g[1] + g[2]) + (g[3] + g[4]) + (g[5] + g[6]);
    const float n4 = -6.0f * g[0] + s16;
    const float n5 = 2.0f * (g[1] + g[2]) - ((g[3] + g[4]) + (g[5] + g[6]));
    const float n6 = (g[3] + g[4]) - (g[5] + g[6]);
    const float p1 = n1 - k.qk * (n1 - Bx) + k.sg * dBx;
    const float p2 = n2 - k.qk * (n2 - By) + k.sg * dBy;
    const float p3 = n3 - k.qk * (n3 - Bz) + k.sg * dBz;
    const float p4 = n4 - k.qnu * (n4 + 0.75f * dT);
    const float p5 = (1.0f - k.qnu) * n5;
    const float p6 = (1.0f - k.qnu) * n6;
    const float c0 = n0 * (1.0f / 7.0f);
    const float c4 = p4 * (1.0f / 42.0f);
    p[0] = c0 - p4 * (1.0f / 7.0f);
    const float ax = c0 + c4 + p5 * (1.0f / 6.0f);
    const float ay = c0 + c4 - p5 * (1.0f / 12.0f);
    p[1] = ax + 0.5f * p1;
    p[2] = ax - 0.5f * p1;
    p[3] = ay + 0.5f * p2 + 0.25f * p6;
    p[4] = ay - 0.5f * p2 + 0.25f * p6;
    p[5] = ay + 0.5f * p3 - 0.25f * p6;
    p[6] = ay - 0.5f * p3 - 0.25f * p6;
}

__device__ __forceinline__ bool on_border_lp(int i, int j, int k, const KParams &kp)
{
    return i == 0 || j == 0 || k == 0 || i == kp.nx - 1 || j == kp.ny - 1 || k == kp.nz - 1;
}
__device__ __forceinline__ bool node_of_lp(const KParams &kp, int &i, int &j, int &k, uint64_t &c)
{
    const unsigned lb = xcd_remap(blockIdx.x, kp.nblocks);
    const unsigned xb = lb % (unsigned)kp.nxb;
    const unsigned row = lb / (unsigned)kp.nxb;
    j = (int)(row % (unsigned)kp.ny);
    k = (int)(row / (unsigned)kp.ny);
    i = (int)(xb * blockDim.x + threadIdx.x);
    c = (uint64_t)i + (uint64_t)kp.nx * ((uint64_t)j + (uint64_t)kp.ny * (uint64_t)k);
    return i < kp.nx;
}

__device__ __forceinline__ void moments19(const float *f, float &drho, float &jx, float &jy, float &jz)
{
    drho = ((f[0] + f[1]) + (f[2] + f[3])) + ((f[4] + f[5]) + (f[6] + f[7])) +
           ((f[8] + f[9]) + (f[10] + f[11])) + ((f[12] + f[13]) + (f[14] + f[15])) +
           ((f[16] + f[17]) + f[18]);
    jx = ((f[1] - f[2]) + (f[7] - f[8])) + ((f[9] - f[10]) + (f[11] - f[12])) + (f[13] - f[14]);
    jy = ((f[3] - f[4]) + (f[7] + f[8])) - ((f[9] + f[10]) - (f[15] - f[16])) + (f[17] - f[18]);
    jz = ((f[5] - f[6]) + (f[11] + f[12])) - ((f[13] + f[14]) - (f[15] + f[16])) - (f[17] + f[18]);
}

// -------------------------------------------------------------------------------------------
// kernels
// -------------------------------------------------------------------------------------------
// In place (AA): Fin is Fout; the plain state (LAY_ODD) becomes LAY_EVEN.
template <typename S, bool AA>
__global__ void __launch_bounds__(256) lp3_prologue(KParams kp, LPParams lp, const S *__restrict__ Fin,
                                                    S *__restrict__ Fout, const double *__restrict__ mac)
{
    int i, j, k;
    uint64_t c;
    if (!node_of_lp(kp, i, j, k, c)) return;
    const uint64_t N = kp.P;   // plane pitch
    float f[19];
    load_all<D3Q19, S>(AA ? Fout : Fin, N, c, f, SEQ(D3Q19));
    const float drho = (float)(mac[c] - 1.0);
    const float dT = (float)(mac[4 * N + c] - lp.tref);
    collide19_dev(f, f, drho, (float)mac[N + c], (float)mac[2 * N + c], (float)mac[3 * N + c], dT, lp);
    out_lp<D3Q19, S, AA ? LAY_ODD : LAY_PULL>(Fout, N, i, j, k, c, false, kp, f, SEQ(D3Q19));
}

// Pass A of step n (see lbm_kernels.hip).  aux holds dT^n as float.
// LAY: layout of the thermal lattice (LAY_PULL: Gin -> Gout; in place: Gout holds state LAY
// and is updated, F is read in the other state).
template <typename S, bool FIRST, int LAY>
__device__ __forceinline__ void lp3_thermal_node(const KParams &kp, const LPParams &lp, int i, int j,
                                                 int k, uint64_t c, const S *__restrict__ Fcur,
                                                 const S *__restrict__ Gin, S *__restrict__ Gout,
                                                 float *__restrict__ Bp, float *__restrict__ aux,
                                                 const double *__restrict__ mac,
                                                 const PullOffsets *pl = nullptr)
{
    const uint64_t N = kp.P;   // plane pitch
    const bool border = on_border_lp(i, j, k, kp);
    const float t0 = t0_of(lp, i, j), b0 = (float)((double)t0 - lp.tref);
    const S *G = LAY == LAY_PULL ? Gin : Gout;
    const bool fast = pl != nullptr && offset_row(*pl, j, k, kp);
    const uint32_t vo = (uint32_t)c * (uint32_t)sizeof(S);   // used when fast
    float g[7], dT;   // dT = T - T0(x)
    if (FIRST) {   // the plain state; in place it is LAY_ODD, read the same way
        load_all<D3Q7, S>(G, N, c, g, SEQ(D3Q7));
        dT = (float)(mac[4 * N + c] - (double)t0);
    } else {
        if (fast) fast_heat_in<S, LAY>(G, *pl, vo, i, j, k, c, kp, lp, g, SEQ(D3Q7));
        else in_heat_lp<D3Q7, S, LAY>(G, N, i, j, k, c, border, kp, lp, g, SEQ(D3Q7));
        dT = (g[0] + g[1]) + (g[2] + g[3]) + (g[4] + g[5]) + g[6];
    }
    float f[19];
    if (fast) fast_flow_in<S, lay_flip_lp(LAY)>(Fcur, *pl, vo, i, j, k, c, kp, f, SEQ(D3Q19));
    else in_flow_lp<D3Q19, S, lay_flip_lp(LAY)>(Fcur, N, i, j, k, c, border, kp, f, SEQ(D3Q19));
    float drho, jx, jy, jz;
    moments19(f, drho, jx, jy, jz);
    const float rho = 1.0f + drho;
    const float ir = 1.0f / rho;
    const float bT = b0 + dT;   // T - Tref
    const float Fy = rho * lp.gbeta * bT;
    const float u = jx * ir, v = (jy + 0.5f * Fy) * ir, w = jz * ir;
    float bp[3] = {Bp[c], Bp[N + c], Bp[2 * N + c]};
    float p[7];
    collide7_dev(g, p, dT, t0 + dT, u, v, w, bp, kp.use_g, lp);
    if (fast) fast_out<D3Q7, S, LAY>(Gout, *pl, vo, i, j, k, c, kp, p, SEQ(D3Q7));
    else out_lp<D3Q7, S, LAY>(Gout, N, i, j, k, c, border, kp, p, SEQ(D3Q7));
    if (kp.use_g) {
        Bp[c] = bp[0];
        Bp[N + c] = bp[1];
        Bp[2 * N + c] = bp[2];
    }
    aux[c] = bT;
}

template <typename S, bool FIRST, int LAY>
__global__ void __launch_bounds__(256) lp3_thermal(KParams kp, LPParams lp, const S *__restrict__ Fcur,
                                                   const S *__restrict__ Gin, S *__restrict__ Gout,
                                                   float *__restrict__ Bp, float *__restrict__ aux,
                                                   const double *__restrict__ mac, PullOffsets pl)
{
    int i, j, k;
    uint64_t c;
    if (!node_of_lp(kp, i, j, k, c)) return;
    lp3_thermal_node<S, FIRST, LAY>(kp, lp, i, j, k, c, Fcur, Gin, Gout, Bp, aux, mac,
                                    &pl);
}

// LAY: layout of both lattices (LAY_PULL: Fcur -> Fnext; in place: Fnext holds state LAY and
// is updated, Gpost is read in state LAY).
template <typename S, int LAY>
__device__ __forceinline__ void lp3_flow_node(const KParams &kp, const LPParams &lp, int i, int j,
                                              int k, uint64_t c, const S *__restrict__ Fcur,
                                              S *__restrict__ Fnext, const S *__restrict__ Gpost,
                                              const float *__restrict__ aux,
                                              const PullOffsets *pl = nullptr)
{
    const uint64_t N = kp.P;   // plane pitch
    const bool border = on_border_lp(i, j, k, kp);
    const bool fast = pl != nullptr && offset_row(*pl, j, k, kp);
    const uint32_t vo = (uint32_t)c * (uint32_t)sizeof(S);   // used when fast
    float g[7];
    if (fast) fast_heat_in<S, LAY>(Gpost, *pl, vo, i, j, k, c, kp, lp, g, SEQ(D3Q7));
    else in_heat_lp<D3Q7, S, LAY>(Gpost, N, i, j, k, c, border, kp, lp, g, SEQ(D3Q7));
    const float dT1 = (g[0] + g[1]) + (g[2] + g[3]) + (g[4] + g[5]) + g[6];
    const float b0 = (float)((double)t0_of(lp, i, j) - lp.tref);
    float f[19];
    if (fast) fast_flow_in<S, LAY>(LAY == LAY_PULL ? Fcur : Fnext, *pl, vo, i, j, k, c, kp, f, SEQ(D3Q19));
    else in_flow_lp<D3Q19, S, LAY>(LAY == LAY_PULL ? Fcur : Fnext, N, i, j, k, c, border, kp, f,
                                   SEQ(D3Q19));
    float drho, jx, jy, jz;
    moments19(f, drho, jx, jy, jz);
    const float rho = 1.0f + drho;
    const float ir = 1.0f / rho;
    const float Fy = rho * lp.gbeta * aux[c];
    const float u = jx * ir, v = (jy + 0.5f * Fy) * ir, w = jz * ir;
    collide19_dev(f, f, drho, u, v, w, b0 + dT1, lp);
    if (fast) fast_out<D3Q19, S, LAY>(Fnext, *pl, vo, i, j, k, c, kp, f, SEQ(D3Q19));
    else out_lp<D3Q19, S, LAY>(Fnext, N, i, j, k, c, border, kp, f, SEQ(D3Q19));
}

template <typename S, int LAY>
__global__ void __launch_bounds__(256) lp3_flow(KParams kp, LPParams lp, const S *__restrict__ Fcur,
                                                S *__restrict__ Fnext, const S *__restrict__ Gpost,
                                                const float *__restrict__ aux, PullOffsets pl)
{
    int i, j, k;
    uint64_t c;
    if (!node_of_lp(kp, i, j, k, c)) return;
    lp3_flow_node<S, LAY>(kp, lp, i, j, k, c, Fcur, Fnext, Gpost, aux,
                          &pl);
}

// Batched two-pass step over a per-domain table (see k3_thermal_batch).
template <typename S> struct TwoPassEntryLP {
    KParams kp;
    LPParams lp;
    PullOffsets po;   // offset rows, decided per node (a block may span rows)
    const S *Fcur;
    S *Fnext;
    const S *Gin;
    S *Gout;
    float *Bp, *aux;
    unsigned first_block;
};

__device__ __forceinline__ bool batch_node_lp(const KParams &kp, unsigned lb, int &i, int &j,
                                              int &k, uint64_t &c)
{
    c = (uint64_t)lb * 256u + threadIdx.x;
    if (c >= kp.N) return false;
    const uint64_t nx = (uint64_t)kp.nx, row = c / nx;
    i = (int)(c - row * nx);
    j = (int)(row % (uint64_t)kp.ny);
    k = (int)(row / (uint64_t)kp.ny);
    return true;
}

template <typename S>
__global__ void __launch_bounds__(256) lp3_thermal_batch(const TwoPassEntryLP<S> *__restrict__ tab,
                                                         const uint16_t *__restrict__ owner,
                                                         unsigned nblocks)
{
    const unsigned b = xcd_remap(blockIdx.x, nblocks);
    const TwoPassEntryLP<S> &d = tab[owner[b]];
    int i, j, k;
    uint64_t c;
    if (!batch_node_lp(d.kp, b - d.first_block, i, j, k, c)) return;
    lp3_thermal_node<S, false, LAY_PULL>(d.kp, d.lp, i, j, k, c, gptr(d.Fcur), gptr(d.Gin), gptr(d.Gout),
                                         gptr(d.Bp), gptr(d.aux), nullptr, &d.po);
}

template <typename S>
__global__ void __launch_bounds__(256) lp3_flow_batch(const TwoPassEntryLP<S> *__restrict__ tab,
                                                      const uint16_t *__restrict__ owner,
                                                      unsigned nblocks)
{
    const unsigned b = xcd_remap(blockIdx.x, nblocks);
    const TwoPassEntryLP<S> &d = tab[owner[b]];
    int i, j, k;
    uint64_t c;
    if (!batch_node_lp(d.kp, b - d.first_block, i, j, k, c)) return;
    lp3_flow_node<S, LAY_PULL>(d.kp, d.lp, i, j, k, c, gptr(d.Fcur), gptr(d.Fnext), gptr(d.Gout),
                               gptr(d.aux), &d.po);
}

// Reference state (absolute FP64) from the pipelined buffers; see k3_state.
template <typename S, int LAY>
__global__ void __launch_bounds__(256) lp3_state(KParams kp, LPParams lp, const S *__restrict__ Fprev,
                                                 const S *__restrict__ Gpost, const float *__restrict__ aux,
                                                 const S *__restrict__ Gprev,
                                                 double *__restrict__ mac, double *__restrict__ fo,
                                                 double *__restrict__ go, int k0, int nk)
{
    int i, j, k;
    uint64_t c;
    if (!node_of_lp(kp, i, j, k, c)) return;
    const uint64_t N = kp.P;   // plane pitch
    const bool border = on_border_lp(i, j, k, kp);
    float g[7], f[19];
    in_heat_lp<D3Q7, S, LAY>(Gpost, N, i, j, k, c, border, kp, lp, g, SEQ(D3Q7));
    in_flow_lp<D3Q19, S, LAY>(Fprev, N, i, j, k, c, border, kp, f, SEQ(D3Q19));
    double drho = 0.0, jx = 0.0, jy = 0.0, jz = 0.0, dT = 0.0;
#pragma unroll
    for (int a = 0; a < 19; ++a) {
        drho += (double)f[a];
        jx += (double)D3Q19::ex[a] * f[a];
        jy += (double)D3Q19::ey[a] * f[a];
        jz += (double)D3Q19::ez[a] * f[a];
    }
#pragma unroll
    for (int a = 0; a < 7; ++a) dT += (double)g[a];
    const double rho = 1.0 + drho;
    const float t0 = t0_of(lp, i, j);
    if (mac) {
        // T^n - Tref of the half force: stored by the two-pass form, else the pull-sum of
        // g_post^{n-1}
        float bTn;
        if (aux) {
            bTn = aux[c];
        } else {
            float gq[7];
            pull_heat_lp<D3Q7, S>(Gprev, N, i, j, k, c, border, kp, lp, gq, SEQ(D3Q7));
            bTn = (float)((double)t0 - lp.tref) +
                  ((gq[0] + gq[1]) + (gq[2] + gq[3]) + (gq[4] + gq[5]) + gq[6]);
        }
        const double Fy = rho * (double)lp.gbeta * (double)bTn;
        mac[c] = rho;
        mac[N + c] = jx / rho;
        mac[2 * N + c] = (jy + 0.5 * Fy) / rho;
        mac[3 * N + c] = jz / rho;
        mac[4 * N + c] = (double)t0 + dT;
    }
    if (fo && k >= k0 && k < k0 + nk) {
        const uint64_t r = c - (uint64_t)k0 * kp.nx * kp.ny;
#pragma unroll
        for (int a = 0; a < 19; ++a) fo[r * 19 + a] = w19(a) + (double)f[a];
#pragma unroll
        for (int a = 0; a < 7; ++a) go[r * 7 + a] = w7(a) * (double)t0 + (double)g[a];
    }
}

// Plain initial state: deviations of f = w (zero) and of g = wT T.
template <typename S>
__global__ void __launch_bounds__(256) lp_init(KParams kp, LPParams lp, S *__restrict__ F, S *__restrict__ G,
                                               float *__restrict__ Bp, const double *__restrict__ mac)
{
    int i, j, k;
    uint64_t c;
    if (!node_of_lp(kp, i, j, k, c)) return;
    const uint64_t N = kp.P;   // plane pitch
    float f[19] = {}, g[7];
    const double T = mac[4 * N + c], t0 = (double)t0_of(lp, i, j);
#pragma unroll
    for (int a = 0; a < 7; ++a) g[a] = (float)(w7(a) * T - w7(a) * t0);
    store_all<D3Q19, S>(F, N, c, f, SEQ(D3Q19));
    store_all<D3Q7, S>(G, N, c, g, SEQ(D3Q7));
    Bp[c] = (float)(mac[N + c] * T);
    Bp[N + c] = (float)(mac[2 * N + c] * T);
    Bp[2 * N + c] = (float)(mac[3 * N + c] * T);
}

// Plain state from absolute FP64 SoA planes (restart): deviations rounded once.
template <typename S>
__global__ void lp_from_abs(KParams kp, LPParams lp, const double *__restrict__ Fd, const double *__restrict__ Gd,
                            S *__restrict__ F, S *__restrict__ G, uint64_t c0, uint64_t n)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint64_t c = c0 + r, N = kp.P;   // N: plane pitch
    const double t0 = (double)t0_of(lp, (int)(c % (uint64_t)kp.nx),
                                    (int)((c / (uint64_t)kp.nx) % (uint64_t)kp.ny));
    float f[19], g[7];
#pragma unroll
    for (int a = 0; a < 19; ++a) f[a] = (float)(Fd[r * 19 + a] - w19(a));
#pragma unroll
    for (int a = 0; a < 7; ++a) g[a] = (float)(Gd[r * 7 + a] - w7(a) * t0);
    store_all<D3Q19, S>(F, N, c, f, SEQ(D3Q19));
    store_all<D3Q7, S>(G, N, c, g, SEQ(D3Q7));
}

// Plain state back to absolute FP64 AoS (gather before the first step).
template <typename S>
__global__ void lp_to_abs(KParams kp, LPParams lp, const S *__restrict__ F, const S *__restrict__ G,
                          double *__restrict__ Fd, double *__restrict__ Gd, uint64_t c0, uint64_t n)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint64_t c = c0 + r, N = kp.P;   // N: plane pitch
    float f[19], g[7];
    load_all<D3Q19, S>(F, N, c, f, SEQ(D3Q19));
    load_all<D3Q7, S>(G, N, c, g, SEQ(D3Q7));
    const double t0 = (double)t0_of(lp, (int)(c % (uint64_t)kp.nx),
                                    (int)((c / (uint64_t)kp.nx) % (uint64_t)kp.ny));
#pragma unroll
    for (int a = 0; a < 19; ++a) Fd[r * 19 + a] = w19(a) + (double)f[a];
#pragma unroll
    for (int a = 0; a < 7; ++a) Gd[r * 7 + a] = w7(a) * t0 + (double)g[a];
}

// reconstruct_macro_from_fg (2DRB/3DRB.F90:1392-1463) on an absolute FP64 AoS chunk.
__global__ void lp_recon_aos(KParams kp, const double *__restrict__ fo, const double *__restrict__ go,
                             double *__restrict__ mac, uint64_t c0, uint64_t n, int *bad)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    const uint64_t c = c0 + r, N = kp.P;   // N: plane pitch
    const double *F = fo + r * 19, *G = go + r * 7;
    const double T = G[0] + G[1] + G[2] + G[3] + G[4] + G[5] + G[6];
    const double rho = F[0] + F[1] + F[2] + F[3] + F[4] + F[5] + F[6] + F[7] + F[8] + F[9] + F[10] +
                       F[11] + F[12] + F[13] + F[14] + F[15] + F[16] + F[17] + F[18];
    const double mx = F[1] - F[2] + F[7] - F[8] + F[9] - F[10] + F[11] - F[12] + F[13] - F[14];
    const double my = F[3] - F[4] + F[7] + F[8] - F[9] - F[10] + F[15] - F[16] + F[17] - F[18];
    const double mz = F[5] - F[6] + F[11] + F[12] - F[13] - F[14] + F[15] + F[16] - F[17] - F[18];
    double u = 0.0, v = 0.0, w = 0.0;
    if (rho > 0.0) {
        const double Fy = rho * kp.gbeta * (T - kp.tref);
        u = mx / rho;
        v = (my + 0.5 * Fy) / rho;
        w = mz / rho;
    } else {
        atomicOr(bad, 1);
    }
    mac[c] = rho;
    mac[N + c] = u;
    mac[2 * N + c] = v;
    mac[3 * N + c] = w;
    mac[4 * N + c] = T;
}

__global__ void lp_bp_from_mac(KParams kp, const double *__restrict__ mac, float *__restrict__ Bp)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t N = kp.P;   // plane pitch
    if (c >= kp.N) return;
    const double T = mac[4 * N + c];
    Bp[c] = kp.use_g ? (float)(mac[N + c] * T) : 0.0f;
    Bp[N + c] = kp.use_g ? (float)(mac[2 * N + c] * T) : 0.0f;
    Bp[2 * N + c] = kp.use_g ? (float)(mac[3 * N + c] * T) : 0.0f;
}

// B_prev as FP64 planes (gather).
__global__ void lp_bprev_to_f64(const float *__restrict__ Bp, double *__restrict__ out, uint64_t n)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r < n) out[r] = (double)Bp[r];
}

// -------------------------------------------------------------------------------------------
// One-pass form (lbm_fused.h) for the deviation storage.
// -------------------------------------------------------------------------------------------
template <typename SS> struct FormLP {
    using R = float;
    using S = SS;
    using B = float;
    using P = LPParams;
    // per-column constants: the storage shift T0 and T0 - Tref
    struct Node {
        float t0, b0;
    };
    static __device__ __forceinline__ Node node(const KParams &, const P &lp, int i, int j)
    {
        const float t0 = t0_of(lp, i, j);
        return Node{t0, (float)((double)t0 - lp.tref)};
    }
    // post-collision thermal deviations as the storage holds them (FP16: rounded; the weights
    // of D3Q7 are powers of two, so a second rounding leaves them unchanged): the values the
    // one-pass kernel exchanges are then the ones the two-pass form reads back
    static __device__ __forceinline__ void round_g(float *g)
    {
        if constexpr (sizeof(S) < sizeof(float)) {
#pragma unroll
            for (int a = 0; a < 7; ++a) {
                const float w = (float)w7(a);
                g[a] = (float)(S)(g[a] * (1.0f / w)) * w;
            }
        }
    }
    // a neighbour's post-collision deviation of direction a pulled into this node
    static __device__ __forceinline__ float pulled(int a, float v, const P &lp)
    {
        const int e = lp.t0axis == 1 ? D3Q7::ex[a] : (lp.t0axis == 2 ? D3Q7::ey[a] : 0);
        return v + -(float)w7(a) * ((float)e * lp.t0s);
    }
    // temperature deviation dT = T - T0(x)
    static __device__ __forceinline__ float tsum(const float *g)
    {
        return (g[0] + g[1]) + (g[2] + g[3]) + (g[4] + g[5]) + g[6];
    }
    // returns drho (density deviation) in the rho slot
    static __device__ __forceinline__ void macro(const float *f, float dTn, const KParams &,
                                                 const P &lp, const Node &nd, float &drho, float &u,
                                                 float &v, float &w)
    {
        float jx, jy, jz;
        moments19(f, drho, jx, jy, jz);
        const float rho = 1.0f + drho;
        const float ir = 1.0f / rho;
        const float Fy = rho * lp.gbeta * (nd.b0 + dTn);
        u = jx * ir;
        v = (jy + 0.5f * Fy) * ir;
        w = jz * ir;
    }
    static __device__ __forceinline__ void thermal(const float *g, float *p, float dT, float u,
                                                   float v, float w, float *bp, const KParams &kp,
                                                   const P &lp, const Node &nd)
    {
        collide7_dev(g, p, dT, nd.t0 + dT, u, v, w, bp, kp.use_g, lp);
    }
    static __device__ __forceinline__ void flow(const float *f, float *out, float drho, float u,
                                                float v, float w, float dT1, const KParams &,
                                                const P &lp, const Node &nd)
    {
        collide19_dev(f, out, drho, u, v, w, nd.b0 + dT1, lp);
    }
    static __device__ __forceinline__ void store_g(S *G, uint64_t N, uint64_t c, const float *g)
    {
        store_all<D3Q7, S>(G, N, c, g, SEQ(D3Q7));
    }
    static __device__ __forceinline__ void store_f(S *F, uint64_t N, uint64_t c, const float *f)
    {
        store_all<D3Q19, S>(F, N, c, f, SEQ(D3Q19));
    }
    // as store_g / store_f at plane element pk and in-plane byte offset vo (FP32: buffer
    // descriptors; FP16 keeps the converting stores at c = pk + vo / 2)
    template <int... A>
    static __device__ __forceinline__ void store_inner_seq(S *P, uint64_t N, int64_t pk, uint32_t vo,
                                                           const float *v, std::integer_sequence<int, A...>)
    {
        (buf_st(P + ((int64_t)A * (int64_t)N + pk), vo, v[A]), ...);
    }
    static __device__ __forceinline__ void store_g_inner(S *G, uint64_t N, int64_t pk, uint32_t vo,
                                                         const float *g)
    {
        if constexpr (sizeof(S) == 4) store_inner_seq(G, N, pk, vo, g, SEQ(D3Q7));
        else store_g(G, N, (uint64_t)pk + vo / sizeof(S), g);
    }
    static __device__ __forceinline__ void store_f_inner(S *F, uint64_t N, int64_t pk, uint32_t vo,
                                                         const float *f)
    {
        if constexpr (sizeof(S) == 4) store_inner_seq(F, N, pk, vo, f, SEQ(D3Q19));
        else store_f(F, N, (uint64_t)pk + vo / sizeof(S), f);
    }
    // prefetching into registers (lbm_fused.h, PF == 1 / 3): raw loads at the pulls' source
    // addresses (returning the thermal wall modes, two bits each), converted after arrival
    static __device__ __forceinline__ void raw_f(const S *F, uint64_t N, int i, int j, int k,
                                                 uint64_t c, bool border, const KParams &kp, S *f)
    {
        raw_flow<D3Q19, S>(F, N, i, j, k, c, border, kp, f, SEQ(D3Q19));
    }
    static __device__ __forceinline__ unsigned raw_g(const S *G, uint64_t N, int i, int j, int k,
                                                     uint64_t c, bool border, const KParams &kp, S *g)
    {
        return raw_heat<D3Q7, S>(G, N, i, j, k, c, border, kp, g, SEQ(D3Q7));
    }
    // interior waves (every lane an interior node of plane kq): global loads with the
    // population's upstream element for vo = 0 as the scalar base and the lane's in-plane byte
    // offset vo as the only vector operand; thermal modes all PULL_T0
    static __device__ __forceinline__ void raw_f_inner(const S *F, uint64_t N, int kq,
                                                       uint64_t plane, uint32_t vo,
                                                       const KParams &kp, S *f)
    {
        const int64_t sy = kp.nx, sz = (int64_t)kp.nx * kp.ny, pk = (int64_t)kq * (int64_t)plane;
#pragma unroll
        for (int a = 0; a < 19; ++a) {
            const int64_t sh = D3Q19::ex[a] + D3Q19::ey[a] * sy + D3Q19::ez[a] * sz;
            f[a] = ld_inner(F, (int64_t)a * (int64_t)N + pk - sh, vo);
        }
    }
    static __device__ __forceinline__ void raw_g_inner(const S *G, uint64_t N, int kq,
                                                       uint64_t plane, uint32_t vo,
                                                       const KParams &kp, S *g)
    {
        const int64_t sy = kp.nx, sz = (int64_t)kp.nx * kp.ny, pk = (int64_t)kq * (int64_t)plane;
#pragma unroll
        for (int a = 0; a < 7; ++a) {
            const int64_t sh = D3Q7::ex[a] + D3Q7::ey[a] * sy + D3Q7::ez[a] * sz;
            g[a] = ld_inner(G, (int64_t)a * (int64_t)N + pk - sh, vo);
        }
    }
    template <typename T>
    static __device__ __forceinline__ T ld_inner(const T *P, int64_t elem, uint32_t vo)
    {
        return buf_ld(P + elem, vo);
    }
    static __device__ __forceinline__ void cook_f(const S *raw, float *f)
    {
        cook_flow<D3Q19, S>(raw, f, SEQ(D3Q19));
    }
    static __device__ __forceinline__ void cook_g(const S *raw, unsigned modes, const P &lp,
                                                  const Node &nd, float *g)
    {
        cook_heat<D3Q7, S>(raw, modes, lp, nd.t0, g, SEQ(D3Q7));
    }
    // one-pass thread order (lbm_fused.h): interior rows first, ring nodes in waves of their own
    static constexpr bool ring_last = true;
    // staged loads of the one-pass kernel (lbm_fused.h, PF == 2): 4-byte storage only
    static constexpr bool can_stage = sizeof(S) == 4;
    static __device__ __forceinline__ void stage_f(const S *F, uint64_t N, int i, int j, int k,
                                                   uint64_t c, bool border, const KParams &kp,
                                                   S *dst, int pitch)
    {
        if constexpr (can_stage) stage_flow<D3Q19, S>(F, N, i, j, k, c, border, kp, dst, pitch, SEQ(D3Q19));
    }
    // returns the wall modes of the seven thermal pulls, two bits each
    static __device__ __forceinline__ unsigned stage_g(const S *G, uint64_t N, int i, int j, int k,
                                                       uint64_t c, bool border, const KParams &kp,
                                                       S *dst, int pitch)
    {
        if constexpr (can_stage) return stage_heat<D3Q7, S>(G, N, i, j, k, c, border, kp, dst, pitch, SEQ(D3Q7));
        return 0;
    }
    static __device__ __forceinline__ void unstage_f(const S *slot, int pitch, float *f)
    {
        if constexpr (can_stage) {
#pragma unroll
            for (int a = 0; a < 19; ++a) f[a] = slot[a * pitch];
        }
    }
    static __device__ __forceinline__ void unstage_g(const S *slot, int pitch, unsigned modes,
                                                     const P &lp, const Node &nd, float *g)
    {
        if constexpr (can_stage) unstage_heat<D3Q7>(slot, pitch, modes, lp, nd.t0, g, SEQ(D3Q7));
    }
    static __device__ __forceinline__ float wall_g(int a, const float *gp, int kind, const P &lp,
                                                   const Node &nd)
    {
        const float back = gp[D3Q7::opp[a]];
        if (kind == T_ADIABATIC) return back;
        const float tw = ((a & 1) ? lp.thot : lp.tcold) - nd.t0;
        return -back + 2.0f * (float)w7(a) * tw;
    }
};

} // namespace rblbm

#include "lbm_fused.h"

namespace rblbm {

// -------------------------------------------------------------------------------------------
// launchers
// -------------------------------------------------------------------------------------------
static LPParams lp_params(const KParams &kp)
{
    LPParams lp;
    lp.gbeta = (float)kp.gbeta;
    lp.sv = (float)kp.sv;
    lp.sq = (float)kp.sq;
    lp.qk = (float)kp.qk;
    lp.qnu = (float)kp.qnu;
    lp.sg = (float)kp.sg;
    lp.tref = kp.tref;
    lp.thot = (float)kp.thot;
    lp.tcold = (float)kp.tcold;
    // storage shift: the conduction profile between two isothermal walls, y before x
    lp.t0axis = kp.temp_y == T_CONST ? 2 : (kp.temp_x == T_CONST ? 1 : 0);
    if (lp.t0axis == 0) {
        lp.t0a = (float)kp.tref;
        lp.t0s = 0.0f;
    } else {
        const int n = lp.t0axis == 2 ? kp.ny : kp.nx;
        const double s = (kp.tcold - kp.thot) / n;   // per node, from the low (hot) wall
        if (s == 0.0) {
            lp.t0a = (float)kp.thot;
            lp.t0s = 0.0f;
        } else {
            // 10 significant bits for the slope and the offset on the slope's last bit: every
            // T0 = t0a + t0s * idx (idx < 2^13) and its differences are exact in FP32
            int e;
            frexp(s, &e);
            const double q = ldexp(1.0, e - 10);
            lp.t0s = (float)(nearbyint(s / q) * q);
            lp.t0a = (float)(nearbyint((kp.thot + 0.5 * (double)lp.t0s) / q) * q);
        }
    }
    return lp;
}

template <typename S>
static hipError_t lp_dispatch(int op, const KParams &kp, int bx, hipStream_t s, const LPArgs &a)
{
    const LPParams lp = lp_params(kp);
    const PullOffsets pl = pull_offsets(kp, (int64_t)sizeof(S));
    const dim3 grid(kp.nblocks);
    const S *Fc = (const S *)a.Fc;
    switch (op) {
    case LP_PROLOGUE:
        if (a.lay == LAY_PULL) lp3_prologue<S, false><<<grid, bx, 0, s>>>(kp, lp, Fc, (S *)a.Fn, a.mac);
        else lp3_prologue<S, true><<<grid, bx, 0, s>>>(kp, lp, nullptr, (S *)a.Fn, a.mac);
        break;
    case LP_STEP:
        if (a.first)
            lp3_thermal<S, true, LAY_PULL><<<grid, bx, 0, s>>>(kp, lp, Fc, (const S *)a.Gi, (S *)a.Go, a.Bp, a.aux, a.mac, pl);
        else
            lp3_thermal<S, false, LAY_PULL><<<grid, bx, 0, s>>>(kp, lp, Fc, (const S *)a.Gi, (S *)a.Go, a.Bp, a.aux, a.mac, pl);
        lp3_flow<S, LAY_PULL><<<grid, bx, 0, s>>>(kp, lp, Fc, (S *)a.Fn, (const S *)a.Go, a.aux, pl);
        break;
    case LP_FLOW_AA:
        if (a.lay == LAY_ODD)
            lp3_flow<S, LAY_ODD><<<grid, bx, 0, s>>>(kp, lp, nullptr, (S *)a.Fn, (const S *)a.Gi, a.aux, pl);
        else if (a.lay == LAY_EVEN)
            lp3_flow<S, LAY_EVEN><<<grid, bx, 0, s>>>(kp, lp, nullptr, (S *)a.Fn, (const S *)a.Gi, a.aux, pl);
        else
            return hipErrorInvalidValue;
        break;
    case LP_THERMAL_AA:
        if (a.first && a.lay == LAY_ODD)
            lp3_thermal<S, true, LAY_ODD><<<grid, bx, 0, s>>>(kp, lp, Fc, nullptr, (S *)a.Go, a.Bp, a.aux, a.mac, pl);
        else if (a.first)
            return hipErrorInvalidValue;   // the plain state is LAY_ODD
        else if (a.lay == LAY_ODD)
            lp3_thermal<S, false, LAY_ODD><<<grid, bx, 0, s>>>(kp, lp, Fc, nullptr, (S *)a.Go, a.Bp, a.aux, a.mac, pl);
        else if (a.lay == LAY_EVEN)
            lp3_thermal<S, false, LAY_EVEN><<<grid, bx, 0, s>>>(kp, lp, Fc, nullptr, (S *)a.Go, a.Bp, a.aux, a.mac, pl);
        else
            return hipErrorInvalidValue;
        break;
    case LP_STATE:
        if (a.lay == LAY_ODD)
            lp3_state<S, LAY_ODD><<<grid, bx, 0, s>>>(kp, lp, Fc, (const S *)a.Gi, a.aux, nullptr,
                                                      a.mac, a.fo, a.go, a.k0, a.nk);
        else if (a.lay == LAY_EVEN)
            lp3_state<S, LAY_EVEN><<<grid, bx, 0, s>>>(kp, lp, Fc, (const S *)a.Gi, a.aux, nullptr,
                                                       a.mac, a.fo, a.go, a.k0, a.nk);
        else
            lp3_state<S, LAY_PULL><<<grid, bx, 0, s>>>(kp, lp, Fc, (const S *)a.Gi, a.aux,
                                                       (const S *)a.Gprev, a.mac, a.fo, a.go, a.k0, a.nk);
        break;
    case LP_INIT:
        lp_init<S><<<grid, bx, 0, s>>>(kp, lp, (S *)a.Fn, (S *)a.Go, a.Bp, a.mac);
        break;
    case LP_FROM_ABS: {
        const unsigned b = (unsigned)((a.n + 255) / 256);
        if (b) lp_from_abs<S><<<b, 256, 0, s>>>(kp, lp, a.fo, a.go, (S *)a.Fn, (S *)a.Go, a.c0, a.n);
        break;
    }
    case LP_TO_ABS: {
        const unsigned b = (unsigned)((a.n + 255) / 256);
        if (b) lp_to_abs<S><<<b, 256, 0, s>>>(kp, lp, Fc, (const S *)a.Gi, a.fo, a.go, a.c0, a.n);
        break;
    }
    case LP_RECON_AOS: {
        const unsigned b = (unsigned)((a.n + 255) / 256);
        if (b) lp_recon_aos<<<b, 256, 0, s>>>(kp, a.fo, a.go, a.mac, a.c0, a.n, a.flag);
        break;
    }
    case LP_BP_FROM_MAC: {
        const unsigned b = (unsigned)((kp.N + 255) / 256);
        lp_bp_from_mac<<<b, 256, 0, s>>>(kp, a.mac, a.Bp);
        break;
    }
    case LP_BPREV: {
        const unsigned b = (unsigned)((a.n + 255) / 256);
        if (b) lp_bprev_to_f64<<<b, 256, 0, s>>>(a.Bp, a.fo, a.n);
        break;
    }
    default:
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

template <typename S>
static void fill_lp(void *dst, const KParams &kp, const FusedGeom &fg, const FusedPtrs &p,
                    unsigned first_block)
{
    FusedInstance<FormLP<S>> d;
    d.kp = kp;
    d.fp = lp_params(kp);
    d.fg = fg;
    d.Fcur = (const S *)p.Fcur;
    d.Fnext = (S *)p.Fnext;
    d.Gin = (const S *)p.Gin;
    d.Gout = (S *)p.Gout;
    d.Bp = (const float *)p.Bp;
    d.Bpout = (float *)p.Bpout;
    d.first_block = first_block;
    memcpy(dst, &d, sizeof(d));
}

template <typename S>
static void fill_twopass_lp(void *dst, const KParams &kp, const FusedPtrs &p, float *aux,
                            unsigned first_block)
{
    TwoPassEntryLP<S> d;
    d.kp = kp;
    d.lp = lp_params(kp);
    d.po = pull_offsets(kp, (int64_t)sizeof(S));
    d.Fcur = (const S *)p.Fcur;
    d.Fnext = (S *)p.Fnext;
    d.Gin = (const S *)p.Gin;
    d.Gout = (S *)p.Gout;
    d.Bp = (float *)p.Bp;
    d.aux = aux;
    d.first_block = first_block;
    memcpy(dst, &d, sizeof(d));
}

size_t twopass_entry_bytes_lp(int precision)
{
    return precision == 1 ? sizeof(TwoPassEntryLP<float>) : sizeof(TwoPassEntryLP<_Float16>);
}

void twopass_entry_fill_lp(int precision, void *dst, const KParams &kp, const FusedPtrs &p,
                           float *aux, unsigned first_block)
{
    if (precision == 1) fill_twopass_lp<float>(dst, kp, p, aux, first_block);
    else fill_twopass_lp<_Float16>(dst, kp, p, aux, first_block);
}

template <typename S>
static hipError_t twopass_batch(const void *tab, const uint16_t *owner, unsigned nblocks,
                                hipStream_t s)
{
    const TwoPassEntryLP<S> *d = (const TwoPassEntryLP<S> *)tab;
    lp3_thermal_batch<S><<<nblocks, 256, 0, s>>>(d, owner, nblocks);
    lp3_flow_batch<S><<<nblocks, 256, 0, s>>>(d, owner, nblocks);
    return hipGetLastError();
}

hipError_t launch_twopass_batch_lp(int precision, const void *tab, const uint16_t *owner,
                                   unsigned nblocks, hipStream_t s)
{
    if (precision == 1) return twopass_batch<float>(tab, owner, nblocks, s);
    return twopass_batch<_Float16>(tab, owner, nblocks, s);
}

size_t fused_instance_bytes_lp(int precision)
{
    return precision == 1 ? sizeof(FusedInstance<FormLP<float>>)
                          : sizeof(FusedInstance<FormLP<_Float16>>);
}

void fused_instance_fill_lp(int precision, void *dst, const KParams &kp, const FusedGeom &fg,
                            const FusedPtrs &p, unsigned first_block)
{
    if (precision == 1) fill_lp<float>(dst, kp, fg, p, first_block);
    else fill_lp<_Float16>(dst, kp, fg, p, first_block);
}

hipError_t launch_fused_batch_lp(int precision, const FusedGeom &fg0, const void *inst,
                                 const uint16_t *owner, unsigned nblocks, hipStream_t s)
{
    if (precision == 1) return launch_fused_table<FormLP<float>>(fg0, inst, owner, nblocks, s);
    return launch_fused_table<FormLP<_Float16>>(fg0, inst, owner, nblocks, s);
}

hipError_t launch_lp(int precision, int op, const KParams &kp, int bx, hipStream_t s, const LPArgs &a)
{
    if (precision == 1) return lp_dispatch<float>(op, kp, bx, s, a);
    if (precision == 2) return lp_dispatch<_Float16>(op, kp, bx, s, a);
    return hipErrorInvalidValue;
}

} // namespace rblbm
