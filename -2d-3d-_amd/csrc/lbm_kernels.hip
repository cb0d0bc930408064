// lbm_kernels.hip -- gfx950 kernels of the thermal MRT-LBM Rayleigh-Benard step (FP64 form).
//
// Data layout in HBM (structure of arrays, x fastest):
//   F[2]  flow populations,   F[b][a*N + c]    19 (2D: 9) planes of N nodes, ping-pong
//   G[2]  thermal populations, G[b][a*N + c]   7 (2D: 5) planes, ping-pong
//   Bp    M1G history B_prev = u T, 3 planes
//   aux   3D: T of the previous step (half force of `macro`); 2D: two planes holding the
//         stored y force of the two latest flow collisions (2DRB/2DRB.F90:718-719, :887-888)
//   mac   rho, u, v, w, T planes (initial state, then diagnostics output)
// with c = i + nx (j + ny k).  Storage holds POST-collision populations; every kernel pulls
// f(a, x) = f_post(a, x - e_a) and applies the walls inline (the reference's streaming +
// bounceback sweeps, 2DRB/3DRB.F90:957-1058 and :1209-1351).
//
// Arithmetic: this translation unit is compiled with -ffp-contract=off and the collisions follow
// the reference's expression order (lbm_collide_ref.h), so the FP64 form is bitwise identical
// to the reference's CPU path on the same inputs.
//
// One reference time step n (collision .. macroT, 2DRB/3DRB.F90:227-243) is two launches:
//   A(n): pull f^{n+1} from F (f_post^n) -> rho, u^{n+1} with the half force of T^n;
//         pull g^n from G (g_post^{n-1}) -> T^n; thermal collision -> g_post^n; B_prev.
//   B(n): pull f^{n+1} again; pull-sum g_post^n -> T^{n+1}; flow collision of step n+1
//         -> f_post^{n+1}.
// The flow collision of step n+1 needs T^{n+1} at the node, which needs the thermal
// collision at its six neighbours, which needs u^{n+1} there: two neighbour hops per step,
// hence two passes.  A "prologue" launch performs the flow collision of step 1 from the plain
// initial state, and the first A launch reads g^1 and T^1 unpulled from that state.
#include <utility>

#include "lbm_collide_ref.h"
#include "lbm_internal.h"
#include "lbm_pull.h"

namespace rblbm {

// ---------------------------------------------------------------------------------------
// Pull with walls.
// ---------------------------------------------------------------------------------------
// Flow lattice: the wall sweeps of `bounceback` overwrite in the order x, y, z, so the value
// of a population whose source lies outside is decided by the LAST outside axis:
//   z outside: periodic -> f_post(a, i, j, wrap k) (same i, j); no-slip -> f_post(opp a, x)
//   y outside: no-slip  -> f_post(opp a, x)
//   x outside: periodic -> f_post(a, wrap i, j, k) (same j, k); no-slip -> f_post(opp a, x)
template <class L>
__device__ __forceinline__ void pull_flow(const double *__restrict__ P, uint64_t N, int i, int j,
                                          int k, uint64_t c, bool border, const KParams &kp,
                                          double *f)
{
    const int sy = kp.nx, sz = kp.nx * kp.ny;
    if (!border) {
#pragma unroll
        for (int a = 0; a < L::Q; ++a)
            f[a] = P[a * N + c - (L::ex[a] + L::ey[a] * sy + L::ez[a] * sz)];
        return;
    }
#pragma unroll
    for (int a = 0; a < L::Q; ++a) {
        const int si = i - L::ex[a], sj = j - L::ey[a], sk = k - L::ez[a];
        const bool ox = (unsigned)si >= (unsigned)kp.nx;
        const bool oy = (unsigned)sj >= (unsigned)kp.ny;
        const bool oz = (unsigned)sk >= (unsigned)kp.nz;
        uint64_t src;
        int pa = a;
        if (oz) {
            if (kp.vel_z == VEL_PERIODIC) src = c + (uint64_t)(L::ez[a] == 1 ? kp.nz - 1 - k : -k) * sz;
            else { src = c; pa = L::opp[a]; }
        } else if (oy) {
            src = c;
            pa = L::opp[a];
        } else if (ox) {
            if (kp.vel_x == VEL_PERIODIC) src = c + (uint64_t)(int64_t)(L::ex[a] == 1 ? kp.nx - 1 - i : -i);
            else { src = c; pa = L::opp[a]; }
        } else {
            src = c - (int64_t)(L::ex[a] + L::ey[a] * sy + L::ez[a] * sz);
        }
        f[a] = P[pa * N + src];
    }
}

// Thermal lattice (one non-zero component per direction, so only one wall applies):
//   periodic -> g_post(a, wrap); constant T -> -g_post(opp a, x) + 2 wT_a T_wall (hot wall at
//   the low index); adiabatic -> g_post(opp a, x).
// The direction is a template argument so that its axis and sign are compile-time constants.
template <class L, int A>
__device__ __forceinline__ double heat_in(const double *__restrict__ P, uint64_t N, int i, int j,
                                          int k, uint64_t c, const KParams &kp)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    const int sy = kp.nx, sz = kp.nx * kp.ny;
    int kind = -1;
    int64_t wrap = 0;
    if (ex != 0) {
        if ((unsigned)(i - ex) >= (unsigned)kp.nx) {
            kind = kp.temp_x;
            wrap = ex == 1 ? kp.nx - 1 - i : -i;
        }
    } else if (ey != 0) {
        if ((unsigned)(j - ey) >= (unsigned)kp.ny) {
            kind = kp.temp_y;
            wrap = (int64_t)(ey == 1 ? kp.ny - 1 - j : -j) * sy;
        }
    } else if (ez != 0) {
        if ((unsigned)(k - ez) >= (unsigned)kp.nz) {
            kind = kp.temp_z;
            wrap = (int64_t)(ez == 1 ? kp.nz - 1 - k : -k) * sz;
        }
    }
    if (kind < 0) return P[A * N + c - (int64_t)(ex + ey * sy + ez * sz)];
    if (kind == T_PERIODIC) return P[A * N + c + wrap];
    const double back = P[L::opp[A] * N + c];
    if (kind == T_ADIABATIC) return back;
    // kp.abb = 2 wT_A (every moving direction has the same weight); legacy scheme
    // (6 + paraA) / 21 (3D) or (4 + paraA) / 10 (2D)
    constexpr bool low_wall = (ex + ey + ez) > 0;   // entered through the low-index wall
    const double tw = low_wall ? kp.thot : kp.tcold;
    return -back + kp.abb * tw;
}

template <class L, int... A>
__device__ __forceinline__ void heat_in_all(const double *__restrict__ P, uint64_t N, int i, int j,
                                            int k, uint64_t c, const KParams &kp, double *g,
                                            std::integer_sequence<int, A...>)
{
    ((g[A] = heat_in<L, A>(P, N, i, j, k, c, kp)), ...);
}

template <class L>
__device__ __forceinline__ void pull_heat(const double *__restrict__ P, uint64_t N, int i, int j,
                                          int k, uint64_t c, bool border, const KParams &kp,
                                          double *g)
{
    if (!border) {
        const int sy = kp.nx, sz = kp.nx * kp.ny;
#pragma unroll
        for (int a = 0; a < L::Q; ++a)
            g[a] = P[a * N + c - (L::ex[a] + L::ey[a] * sy + L::ez[a] * sz)];
        return;
    }
    heat_in_all<L>(P, N, i, j, k, c, kp, g, std::make_integer_sequence<int, L::Q>{});
}

// In-place (A-A) layout (see LAY_ODD / LAY_EVEN): the incoming populations, their wall values
// and the store of the post-collision populations.  Walls only, as in pull_flow / heat_in
// without their periodic branches.
__device__ __forceinline__ bool inside3(int i, int j, int k, const KParams &kp)
{
    return (unsigned)i < (unsigned)kp.nx && (unsigned)j < (unsigned)kp.ny && (unsigned)k < (unsigned)kp.nz;
}

template <class L, int LAY>
__device__ __forceinline__ void in_flow_aa(const double *P, uint64_t N, int i, int j, int k,
                                           uint64_t c, bool border, const KParams &kp, double *f)
{
    const int sy = kp.nx, sz = kp.nx * kp.ny;
#pragma unroll
    for (int a = 0; a < L::Q; ++a) {
        if (LAY == LAY_ODD) {
            f[a] = P[a * N + c];
        } else {
            const int64_t off = L::ex[a] + L::ey[a] * sy + L::ez[a] * sz;
            const bool in = !border || inside3(i - L::ex[a], j - L::ey[a], k - L::ez[a], kp);
            f[a] = P[in ? L::opp[a] * N + c - off : a * N + c];
        }
    }
}

template <class L, int LAY, int A>
__device__ __forceinline__ double heat_in_aa(const double *P, uint64_t N, int i, int j, int k,
                                             uint64_t c, bool border, const KParams &kp)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    const int sy = kp.nx, sz = kp.nx * kp.ny;
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
    if (kind < 0)
        return LAY == LAY_ODD ? P[A * N + c]
                              : P[L::opp[A] * N + c - (int64_t)(ex + ey * sy + ez * sz)];
    const double back = P[A * N + c];   // g_post of the opposite direction, in both states
    if (kind == T_ADIABATIC) return back;
    constexpr bool low_wall = (ex + ey + ez) > 0;
    const double tw = low_wall ? kp.thot : kp.tcold;
    return -back + kp.abb * tw;
}

template <class L, int LAY, int... A>
__device__ __forceinline__ void in_heat_aa(const double *P, uint64_t N, int i, int j, int k,
                                           uint64_t c, bool border, const KParams &kp, double *g,
                                           std::integer_sequence<int, A...>)
{
    ((g[A] = heat_in_aa<L, LAY, A>(P, N, i, j, k, c, border, kp)), ...);
}

template <class L, int LAY>
__device__ __forceinline__ void out_aa(double *P, uint64_t N, int i, int j, int k, uint64_t c,
                                       bool border, const KParams &kp, const double *v)
{
    const int sy = kp.nx, sz = kp.nx * kp.ny;
#pragma unroll
    for (int a = 0; a < L::Q; ++a) {
        if (LAY == LAY_ODD) {
            P[L::opp[a] * N + c] = v[a];
        } else {
            const int64_t off = L::ex[a] + L::ey[a] * sy + L::ez[a] * sz;
            const bool in = !border || inside3(i + L::ex[a], j + L::ey[a], k + L::ez[a], kp);
            P[in ? a * N + c + off : L::opp[a] * N + c] = v[a];
        }
    }
}

// Incoming populations of a node in layout LAY (pull from the other array, or A-A in place).
template <int LAY>
__device__ __forceinline__ void in_flow(const double *P, uint64_t N, int i, int j, int k,
                                        uint64_t c, bool border, const KParams &kp, double *f)
{
    if (LAY == LAY_PULL) pull_flow<D3Q19>(P, N, i, j, k, c, border, kp, f);
    else in_flow_aa<D3Q19, LAY>(P, N, i, j, k, c, border, kp, f);
}
template <int LAY>
__device__ __forceinline__ void in_heat(const double *P, uint64_t N, int i, int j, int k,
                                        uint64_t c, bool border, const KParams &kp, double *g)
{
    if (LAY == LAY_PULL) pull_heat<D3Q7>(P, N, i, j, k, c, border, kp, g);
    else in_heat_aa<D3Q7, LAY>(P, N, i, j, k, c, border, kp, g, std::make_integer_sequence<int, 7>{});
}
// the state an in-place pass leaves behind, and the layout the flow lattice is read with while
// the thermal lattice is in state LAY (pass A runs between the two flow states)
constexpr int lay_flip(int lay) { return lay == LAY_PULL ? LAY_PULL : 3 - lay; }

// Offset rows (PullOffsets, lbm_lattice.h): each population is read through a buffer descriptor
// whose base is the array plus the population's byte offset (built in scalar registers), with
// the node's 32-bit byte offset as the only vector operand.  Plain pointer arithmetic here was
// compiled into one 64-bit vector add per load.  Word 3 of the descriptor is the raw-buffer
// format of the gfx9 family; the record count covers 4 GiB (the offsets fit 32 bits, so no
// access is clamped).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t row_rsrc(const double *P, int64_t off)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)((const char *)P + off), (short)0, -1, 0x00020000);
}
__device__ __forceinline__ double ld_off(const double *P, int64_t off, uint32_t vo)
{
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(row_rsrc(P, off), vo, 0, 0));
}
template <int X, int... A>
__device__ __forceinline__ void redo_x_flow(const double *P, uint64_t N, int i, int j, int k, uint64_t c,
                                            const KParams &kp, double *f, std::integer_sequence<int, A...>)
{
    ((f[A] = D3Q19::ex[A] == X ? P[flow_elem<D3Q19, A>(N, i, j, k, c, kp)] : f[A]), ...);
}
template <int X, int... A>
__device__ __forceinline__ void redo_x_heat(const double *P, uint64_t N, int i, int j, int k, uint64_t c,
                                            const KParams &kp, double *g, std::integer_sequence<int, A...>)
{
    ((g[A] = D3Q7::ex[A] == X ? heat_in<D3Q7, A>(P, N, i, j, k, c, kp) : g[A]), ...);
}
__device__ __forceinline__ void offset_flow_in(const double *P, const PullOffsets &po, uint32_t vo,
                                               int i, int j, int k, uint64_t c, const KParams &kp,
                                               double *f)
{
#pragma unroll
    for (int a = 0; a < 19; ++a) f[a] = ld_off(P, po.f[a], vo);
    // x-wall nodes redo the pulls whose source lies beyond their wall (ex = +1 at i = 0,
    // ex = -1 at i = nx - 1) with the wall rules (flow_elem: the rules of pull_flow); a wave
    // holds one of the two walls
    if (i == 0) redo_x_flow<1>(P, kp.P, i, j, k, c, kp, f, std::make_integer_sequence<int, 19>{});
    if (i == kp.nx - 1) redo_x_flow<-1>(P, kp.P, i, j, k, c, kp, f, std::make_integer_sequence<int, 19>{});
}
__device__ __forceinline__ void offset_heat_in(const double *P, const PullOffsets &po, uint32_t vo,
                                               int i, int j, int k, uint64_t c, const KParams &kp,
                                               double *g)
{
#pragma unroll
    for (int a = 0; a < 7; ++a) g[a] = ld_off(P, po.g[a], vo);
    if (i == 0) redo_x_heat<1>(P, kp.P, i, j, k, c, kp, g, std::make_integer_sequence<int, 7>{});
    if (i == kp.nx - 1) redo_x_heat<-1>(P, kp.P, i, j, k, c, kp, g, std::make_integer_sequence<int, 7>{});
}
template <int Q>
__device__ __forceinline__ void offset_out(double *P, const PullOffsets &po, uint32_t vo,
                                           const double *v)
{
#pragma unroll
    for (int a = 0; a < Q; ++a) {
        const auto bits = __builtin_bit_cast(unsigned long long, v[a]);
        __builtin_amdgcn_raw_buffer_store_b64(
            __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(row_rsrc(P, 0), 0, 0, 0)), bits),
            row_rsrc(P, po.s[a]), vo, 0, 0);
    }
}

__device__ __forceinline__ bool on_border(int i, int j, int k, const KParams &kp)
{
    return i == 0 || j == 0 || k == 0 || i == kp.nx - 1 || j == kp.ny - 1 || k == kp.nz - 1;
}

// Node coordinates of this thread; false if the thread is past the end of its row.
__device__ __forceinline__ bool node_of(const KParams &kp, int &i, int &j, int &k, uint64_t &c)
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

__device__ __forceinline__ double sum19(const double *f)
{
    return f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8] + f[9] + f[10] + f[11] +
           f[12] + f[13] + f[14] + f[15] + f[16] + f[17] + f[18];
}
// Momentum sums in the reference's term order (macro, 2DRB/3DRB.F90:1079-1090).
__device__ __forceinline__ double jx19(const double *f)
{
    return f[1] - f[2] + f[7] - f[8] + f[9] - f[10] + f[11] - f[12] + f[13] - f[14];
}
__device__ __forceinline__ double jy19(const double *f)
{
    return f[3] - f[4] + f[7] + f[8] - f[9] - f[10] + f[15] - f[16] + f[17] - f[18];
}
__device__ __forceinline__ double jz19(const double *f)
{
    return f[5] - f[6] + f[11] + f[12] - f[13] - f[14] + f[15] + f[16] - f[17] - f[18];
}

// ---------------------------------------------------------------------------------------
// 3D step kernels
// ---------------------------------------------------------------------------------------
// Flow collision of step 1 from the plain initial state (f^1 unpulled in Fin, macros in mac).
// In place (AA): Fin is Fout, the plain state is LAY_ODD and the result LAY_EVEN.
template <bool AA>
__global__ void __launch_bounds__(256) k3_prologue(KParams kp, const double *__restrict__ Fin,
                                                   double *__restrict__ Fout,
                                                   const double *__restrict__ mac)
{
    int i, j, k;
    uint64_t c;
    if (!node_of(kp, i, j, k, c)) return;
    const uint64_t N = kp.P;   // plane pitch
    const double *src = AA ? Fout : Fin;
    double f[19];
#pragma unroll
    for (int a = 0; a < 19; ++a) f[a] = src[a * N + c];
    ref_collide_d3q19(f, f, mac[c], mac[N + c], mac[2 * N + c], mac[3 * N + c], mac[4 * N + c], kp);
#pragma unroll
    for (int a = 0; a < 19; ++a) Fout[(AA ? D3Q19::opp[a] : a) * N + c] = f[a];
}

// Pass A of step n.  FIRST: g^n and T^n come unpulled from the initial state.  LAY: layout of
// the thermal lattice (LAY_PULL: Gin -> Gout; in place: Gout holds state LAY and is updated).
template <bool FIRST, int LAY>
__device__ __forceinline__ void thermal3_node(const KParams &kp, int i, int j, int k, uint64_t c,
                                              const double *__restrict__ Fcur,
                                              const double *__restrict__ Gin,
                                              double *__restrict__ Gout, double *__restrict__ Bp,
                                              double *__restrict__ aux,
                                              const double *__restrict__ mac,
                                              const PullOffsets *po = nullptr)
{
    const uint64_t N = kp.P;   // plane pitch
    const bool border = on_border(i, j, k, kp);
    const bool fast = po != nullptr && offset_row(*po, j, k, kp);
    const uint32_t vo = (uint32_t)c * 8u;   // used when fast
    const double *G = LAY == LAY_PULL ? Gin : Gout;
    double g[7], Tn;
    if (FIRST) {   // the plain state; in place it is LAY_ODD, read the same way
#pragma unroll
        for (int a = 0; a < 7; ++a) g[a] = G[a * N + c];
        Tn = mac[4 * N + c];
    } else {
        if (fast) offset_heat_in(G, *po, vo, i, j, k, c, kp, g);
        else in_heat<LAY>(G, N, i, j, k, c, border, kp, g);
        Tn = g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6];
    }
    double f[19];
    if (fast) offset_flow_in(Fcur, *po, vo, i, j, k, c, kp, f);
    else in_flow<lay_flip(LAY)>(Fcur, N, i, j, k, c, border, kp, f);
    const double rho = sum19(f);
    const double Fy = rho * kp.gbeta * (Tn - kp.tref);
    const double u = jx19(f) / rho;
    const double v = (jy19(f) + 0.5 * Fy) / rho;
    const double w = jz19(f) / rho;
    double bp[3] = {Bp[c], Bp[N + c], Bp[2 * N + c]};
    double p[7];
    ref_collide_d3q7(g, p, Tn, u, v, w, bp, kp);
    if (fast) {
        offset_out<7>(Gout, *po, vo, p);
    } else if (LAY == LAY_PULL) {
#pragma unroll
        for (int a = 0; a < 7; ++a) Gout[a * N + c] = p[a];
    } else {
        out_aa<D3Q7, LAY>(Gout, N, i, j, k, c, border, kp, p);
    }
    if (kp.use_g) {
        Bp[c] = bp[0];
        Bp[N + c] = bp[1];
        Bp[2 * N + c] = bp[2];
    }
    aux[c] = Tn;
}

template <bool FIRST, int LAY>
__global__ void __launch_bounds__(256) k3_thermal(KParams kp, const double *__restrict__ Fcur,
                                                  const double *__restrict__ Gin,
                                                  double *__restrict__ Gout,
                                                  double *__restrict__ Bp,
                                                  double *__restrict__ aux,
                                                  const double *__restrict__ mac, PullOffsets po)
{
    int i, j, k;
    uint64_t c;
    if (!node_of(kp, i, j, k, c)) return;
    thermal3_node<FIRST, LAY>(kp, i, j, k, c, Fcur, Gin, Gout, Bp, aux, mac,
                              LAY == LAY_PULL && !FIRST ? &po : nullptr);
}

// Pass B of step n: flow collision of step n+1.  LAY: layout of both lattices (LAY_PULL:
// Fcur -> Fnext; in place: Fnext holds state LAY and is updated, Gpost is read in state LAY).
template <int LAY>
__device__ __forceinline__ void flow3_node(const KParams &kp, int i, int j, int k, uint64_t c,
                                           const double *__restrict__ Fcur,
                                           double *__restrict__ Fnext,
                                           const double *__restrict__ Gpost,
                                           const double *__restrict__ aux,
                                           const PullOffsets *po = nullptr)
{
    const uint64_t N = kp.P;   // plane pitch
    const bool border = on_border(i, j, k, kp);
    const bool fast = po != nullptr && offset_row(*po, j, k, kp);
    const uint32_t vo = (uint32_t)c * 8u;   // used when fast
    double g[7];
    if (fast) offset_heat_in(Gpost, *po, vo, i, j, k, c, kp, g);
    else in_heat<LAY>(Gpost, N, i, j, k, c, border, kp, g);
    const double T1 = g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6];
    double f[19];
    if (fast) offset_flow_in(Fcur, *po, vo, i, j, k, c, kp, f);
    else in_flow<LAY>(LAY == LAY_PULL ? Fcur : Fnext, N, i, j, k, c, border, kp, f);
    const double rho = sum19(f);
    const double Fy = rho * kp.gbeta * (aux[c] - kp.tref);
    const double u = jx19(f) / rho;
    const double v = (jy19(f) + 0.5 * Fy) / rho;
    const double w = jz19(f) / rho;
    ref_collide_d3q19(f, f, rho, u, v, w, T1, kp);
    if (fast) {
        offset_out<19>(Fnext, *po, vo, f);
    } else if (LAY == LAY_PULL) {
#pragma unroll
        for (int a = 0; a < 19; ++a) Fnext[a * N + c] = f[a];
    } else {
        out_aa<D3Q19, LAY>(Fnext, N, i, j, k, c, border, kp, f);
    }
}

template <int LAY>
__global__ void __launch_bounds__(256) k3_flow(KParams kp, const double *__restrict__ Fcur,
                                               double *__restrict__ Fnext,
                                               const double *__restrict__ Gpost,
                                               const double *__restrict__ aux, PullOffsets po)
{
    int i, j, k;
    uint64_t c;
    if (!node_of(kp, i, j, k, c)) return;
    flow3_node<LAY>(kp, i, j, k, c, Fcur, Fnext, Gpost, aux, LAY == LAY_PULL ? &po : nullptr);
}

// Batched two-pass step: many domains per launch.  Block b (after the XCD remap) belongs to
// domain owner[b] and covers its nodes [(b - first_block) * 256, + 256) in linear order.
struct TwoPassEntryF64 {
    KParams kp;
    PullOffsets po;   // offset rows, decided per node (a block may span rows)
    const double *Fcur;
    double *Fnext;
    const double *Gin;
    double *Gout;
    double *Bp, *aux;
    unsigned first_block;
};

__device__ __forceinline__ bool batch_node(const KParams &kp, unsigned lb, int &i, int &j, int &k,
                                           uint64_t &c)
{
    c = (uint64_t)lb * 256u + threadIdx.x;
    if (c >= kp.N) return false;
    const uint64_t nx = (uint64_t)kp.nx, row = c / nx;
    i = (int)(c - row * nx);
    j = (int)(row % (uint64_t)kp.ny);
    k = (int)(row / (uint64_t)kp.ny);
    return true;
}

__global__ void __launch_bounds__(256) k3_thermal_batch(const TwoPassEntryF64 *__restrict__ tab,
                                                        const uint16_t *__restrict__ owner,
                                                        unsigned nblocks)
{
    const unsigned b = xcd_remap(blockIdx.x, nblocks);
    const TwoPassEntryF64 &d = tab[owner[b]];
    int i, j, k;
    uint64_t c;
    if (!batch_node(d.kp, b - d.first_block, i, j, k, c)) return;
    thermal3_node<false, LAY_PULL>(d.kp, i, j, k, c, gptr(d.Fcur), gptr(d.Gin), gptr(d.Gout),
                                   gptr(d.Bp), gptr(d.aux), nullptr, &d.po);
}

__global__ void __launch_bounds__(256) k3_flow_batch(const TwoPassEntryF64 *__restrict__ tab,
                                                     const uint16_t *__restrict__ owner,
                                                     unsigned nblocks)
{
    const unsigned b = xcd_remap(blockIdx.x, nblocks);
    const TwoPassEntryF64 &d = tab[owner[b]];
    int i, j, k;
    uint64_t c;
    if (!batch_node(d.kp, b - d.first_block, i, j, k, c)) return;
    flow3_node<LAY_PULL>(d.kp, i, j, k, c, gptr(d.Fcur), gptr(d.Fnext), gptr(d.Gout), gptr(d.aux),
                         &d.po);
}

size_t twopass_entry_bytes_f64() { return sizeof(TwoPassEntryF64); }

void twopass_entry_fill_f64(void *dst, const KParams &kp, const FusedPtrs &p, double *aux,
                            unsigned first_block)
{
    TwoPassEntryF64 d;
    d.kp = kp;
    d.po = pull_offsets(kp, 8);
    d.Fcur = (const double *)p.Fcur;
    d.Fnext = (double *)p.Fnext;
    d.Gin = (const double *)p.Gin;
    d.Gout = (double *)p.Gout;
    d.Bp = (double *)p.Bp;
    d.aux = aux;
    d.first_block = first_block;
    memcpy(dst, &d, sizeof(d));
}

hipError_t launch_twopass_batch_f64(const void *tab, const uint16_t *owner, unsigned nblocks,
                                    hipStream_t s)
{
    const TwoPassEntryF64 *d = (const TwoPassEntryF64 *)tab;
    k3_thermal_batch<<<nblocks, 256, 0, s>>>(d, owner, nblocks);
    k3_flow_batch<<<nblocks, 256, 0, s>>>(d, owner, nblocks);
    return hipGetLastError();
}

// Reference state after the last step (f^{n+1}, g^{n+1}, macros) from the pipelined buffers.
// Writes macros to mac (5 planes) and, if fo/go are non-null, AoS populations for the nodes
// of planes [k0, k0 + nk) into fo/go (node index relative to plane k0).
template <int LAY>
__global__ void __launch_bounds__(256) k3_state(KParams kp, const double *__restrict__ Fprev,
                                                const double *__restrict__ Gpost,
                                                const double *__restrict__ aux,
                                                const double *__restrict__ Gprev,
                                                double *__restrict__ mac, double *__restrict__ fo,
                                                double *__restrict__ go, int k0, int nk)
{
    int i, j, k;
    uint64_t c;
    if (!node_of(kp, i, j, k, c)) return;
    const uint64_t N = kp.P;   // plane pitch
    const bool border = on_border(i, j, k, kp);
    double g[7], f[19];
    in_heat<LAY>(Gpost, N, i, j, k, c, border, kp, g);
    in_flow<LAY>(Fprev, N, i, j, k, c, border, kp, f);
    const double T = g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6];
    // T^n of the half force: stored by the two-pass form, else the pull-sum of g_post^{n-1}
    double Tn;
    if (aux) {
        Tn = aux[c];
    } else {
        double gq[7];
        pull_heat<D3Q7>(Gprev, N, i, j, k, c, border, kp, gq);
        Tn = gq[0] + gq[1] + gq[2] + gq[3] + gq[4] + gq[5] + gq[6];
    }
    const double rho = sum19(f);
    const double Fy = rho * kp.gbeta * (Tn - kp.tref);
    if (mac) {
        mac[c] = rho;
        mac[N + c] = jx19(f) / rho;
        mac[2 * N + c] = (jy19(f) + 0.5 * Fy) / rho;
        mac[3 * N + c] = jz19(f) / rho;
        mac[4 * N + c] = T;
    }
    if (fo && k >= k0 && k < k0 + nk) {
        const uint64_t r = c - (uint64_t)k0 * kp.nx * kp.ny;
#pragma unroll
        for (int a = 0; a < 19; ++a) fo[r * 19 + a] = f[a];
#pragma unroll
        for (int a = 0; a < 7; ++a) go[r * 7 + a] = g[a];
    }
}

// ---------------------------------------------------------------------------------------
// 2D step kernels (nz = 1).  Same two-pass structure; the half force of `macro` is the force
// stored by the preceding flow collision (old rho, old T), kept in aux planes.
// ---------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k2_prologue(KParams kp, const double *__restrict__ Fin,
                                                   double *__restrict__ Fout,
                                                   const double *__restrict__ mac,
                                                   double *__restrict__ fy_out)
{
    int i, j, k;
    uint64_t c;
    if (!node_of(kp, i, j, k, c)) return;
    const uint64_t N = kp.P;   // plane pitch
    double f[9];
#pragma unroll
    for (int a = 0; a < 9; ++a) f[a] = Fin[a * N + c];
    fy_out[c] = ref_collide_d2q9(f, f, mac[c], mac[N + c], mac[2 * N + c], mac[4 * N + c], kp);
#pragma unroll
    for (int a = 0; a < 9; ++a) Fout[a * N + c] = f[a];
}

template <bool FIRST>
__global__ void __launch_bounds__(256) k2_thermal(KParams kp, const double *__restrict__ Fcur,
                                                  const double *__restrict__ Gin,
                                                  double *__restrict__ Gout,
                                                  double *__restrict__ Bp,
                                                  const double *__restrict__ fy_cur,
                                                  const double *__restrict__ mac)
{
    int i, j, k;
    uint64_t c;
    if (!node_of(kp, i, j, k, c)) return;
    const uint64_t N = kp.P;   // plane pitch
    const bool border = on_border(i, j, k, kp);
    double g[5], Tn;
    if (FIRST) {
#pragma unroll
        for (int a = 0; a < 5; ++a) g[a] = Gin[a * N + c];
        Tn = mac[4 * N + c];
    } else {
        pull_heat<D2Q5>(Gin, N, i, j, k, c, border, kp, g);
        Tn = g[0] + g[1] + g[2] + g[3] + g[4];
    }
    double f[9];
    pull_flow<D2Q9>(Fcur, N, i, j, k, c, border, kp, f);
    const double rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    const double u = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8] + 0.5 * 0.0) / rho;
    const double v = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8] + 0.5 * fy_cur[c]) / rho;
    double bp[2] = {Bp[c], Bp[N + c]};
    double p[5];
    ref_collide_d2q5(g, p, Tn, u, v, bp, kp);
#pragma unroll
    for (int a = 0; a < 5; ++a) Gout[a * N + c] = p[a];
    if (kp.use_g) {
        Bp[c] = bp[0];
        Bp[N + c] = bp[1];
    }
}

__global__ void __launch_bounds__(256) k2_flow(KParams kp, const double *__restrict__ Fcur,
                                               double *__restrict__ Fnext,
                                               const double *__restrict__ Gpost,
                                               const double *__restrict__ fy_cur,
                                               double *__restrict__ fy_next)
{
    int i, j, k;
    uint64_t c;
    if (!node_of(kp, i, j, k, c)) return;
    const uint64_t N = kp.P;   // plane pitch
    const bool border = on_border(i, j, k, kp);
    double g[5];
    pull_heat<D2Q5>(Gpost, N, i, j, k, c, border, kp, g);
    const double T1 = g[0] + g[1] + g[2] + g[3] + g[4];
    double f[9];
    pull_flow<D2Q9>(Fcur, N, i, j, k, c, border, kp, f);
    const double rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    const double u = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8] + 0.5 * 0.0) / rho;
    const double v = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8] + 0.5 * fy_cur[c]) / rho;
    fy_next[c] = ref_collide_d2q9(f, f, rho, u, v, T1, kp);
#pragma unroll
    for (int a = 0; a < 9; ++a) Fnext[a * N + c] = f[a];
}

__global__ void __launch_bounds__(256) k2_state(KParams kp, const double *__restrict__ Fprev,
                                                const double *__restrict__ Gpost,
                                                const double *__restrict__ fy_prev,
                                                double *__restrict__ mac, double *__restrict__ fo,
                                                double *__restrict__ go, int k0, int nk)
{
    int i, j, k;
    uint64_t c;
    if (!node_of(kp, i, j, k, c)) return;
    const uint64_t N = kp.P;   // plane pitch
    const bool border = on_border(i, j, k, kp);
    double g[5], f[9];
    pull_heat<D2Q5>(Gpost, N, i, j, k, c, border, kp, g);
    pull_flow<D2Q9>(Fprev, N, i, j, k, c, border, kp, f);
    const double rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
    if (mac) {
        mac[c] = rho;
        mac[N + c] = (f[1] - f[3] + f[5] - f[6] - f[7] + f[8] + 0.5 * 0.0) / rho;
        mac[2 * N + c] = (f[2] - f[4] + f[5] + f[6] - f[7] - f[8] + 0.5 * fy_prev[c]) / rho;
        mac[3 * N + c] = 0.0;
        mac[4 * N + c] = g[0] + g[1] + g[2] + g[3] + g[4];
    }
    if (fo && k >= k0 && k < k0 + nk) {
        const uint64_t r = c - (uint64_t)k0 * kp.nx * kp.ny;
#pragma unroll
        for (int a = 0; a < 9; ++a) fo[r * 9 + a] = f[a];
#pragma unroll
        for (int a = 0; a < 5; ++a) go[r * 5 + a] = g[a];
    }
}

// ---------------------------------------------------------------------------------------
// Initial state, restart and plain-state transposes.
// ---------------------------------------------------------------------------------------
__device__ __forceinline__ uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

// Exact branch of `initial` (2DRB/3DRB.F90:517-600): T from the reference profile (or from
// mac[4N..] when use_T0), optional seeded noise, then f = w, g = wT T, rho = 1, u = 0,
// B_prev = 0.  T is written back to mac, the plain f/g to Fout/Gout.
__global__ void __launch_bounds__(256) k_init(KParams kp, InitParams ip, double *__restrict__ Fout,
                                              double *__restrict__ Gout, double *__restrict__ Bp,
                                              double *__restrict__ mac, int qf, int qt)
{
    int i, j, k;
    uint64_t c;
    if (!node_of(kp, i, j, k, c)) return;
    const uint64_t N = kp.P;   // plane pitch
    double T;
    if (ip.use_T0) {
        T = mac[4 * N + c];
    } else {
        T = 0.0;
        // coordinates x_i = (i - 1/2)/L with x_0 = 0 and x_{n+1} = n/L (2DRB/3DRB.F90:451-470)
        const double L = ip.L;
        const double xi = ((double)(i + 1) - 0.5) / L, xn = (double)kp.nx / L;
        const double yj = ((double)(j + 1) - 0.5) / L, yn = (double)kp.ny / L;
        if (kp.temp_x == T_CONST) T = kp.thot + (xi - 0.0) / (xn - 0.0) * (kp.tcold - kp.thot);
        if (kp.temp_y == T_CONST) {
            T = kp.thot + (yj - 0.0) / (yn - 0.0) * (kp.tcold - kp.thot);
            if (ip.rb_perturb == 3)
                T = T + ip.perturb_amp * cos(2.0 * ip.pi * xi / xn) * sin(ip.pi * yj / yn);
            else if (ip.rb_perturb == 2)
                T = T + ip.perturb_amp * sin(2.0 * ip.pi * xi / xn) * sin(ip.pi * yj / yn);
        }
    }
    if (ip.noise_scale != 0.0) {
        const uint64_t key = ip.seed + 0x9E3779B97F4A7C15ULL * (c + 1ULL + (ip.instance << 40));
        const double r = (double)(mix64(key) >> 11) * 0x1.0p-53;
        const double noise = ip.noise_scale * (2.0 * r - 1.0);
        T = T + noise;
    }
    if (Fout) {   // FP64 form; the reduced-precision forms build their state from mac
        for (int a = 0; a < qf; ++a)
            Fout[a * N + c] = qf == 19 ? w19(a) : w9(a);
        for (int a = 0; a < qt; ++a)
            Gout[a * N + c] = (a == 0 ? kp.wt0 : kp.wt1) * T;
        Bp[c] = 0.0;
        Bp[N + c] = 0.0;
        Bp[2 * N + c] = 0.0;
    }
    mac[c] = 1.0;
    mac[N + c] = 0.0;
    mac[2 * N + c] = 0.0;
    mac[3 * N + c] = 0.0;
    mac[4 * N + c] = T;
}

// AoS chunk (reference order) <-> plain SoA planes.  Nodes [c0, c0 + n).
__global__ void k_aos_to_soa(const double *__restrict__ in, double *__restrict__ out, int q,
                             uint64_t N, uint64_t c0, uint64_t n)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int a = 0; a < q; ++a) out[a * N + c0 + r] = in[r * q + a];
}
__global__ void k_soa_to_aos(const double *__restrict__ in, double *__restrict__ out, int q,
                             uint64_t N, uint64_t c0, uint64_t n)
{
    const uint64_t r = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r >= n) return;
    for (int a = 0; a < q; ++a) out[r * q + a] = in[a * N + c0 + r];
}

// reconstruct_macro_from_fg (2DRB/3DRB.F90:1392-1463): macroT, rho, momentum, three fixed-point
// passes of the half-force velocity, B_prev = u T.  *bad is set if any rho <= 0.
__global__ void k_reconstruct(KParams kp, const double *__restrict__ F, const double *__restrict__ G,
                              double *__restrict__ Bp, double *__restrict__ mac, int qf, int qt,
                              int *bad)
{
    const uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const uint64_t N = kp.P;   // plane pitch
    if (c >= kp.N) return;
    double T = 0.0;
    for (int a = 0; a < qt; ++a) T += G[a * N + c];
    double f[19];
    for (int a = 0; a < qf; ++a) f[a] = F[a * N + c];
    double rho, mx, my, mz = 0.0;
    if (qf == 19) {
        rho = sum19(f);
        mx = jx19(f);
        my = jy19(f);
        mz = jz19(f);
    } else {
        rho = f[0] + f[1] + f[2] + f[3] + f[4] + f[5] + f[6] + f[7] + f[8];
        mx = f[1] - f[3] + f[5] - f[6] - f[7] + f[8];
        my = f[2] - f[4] + f[5] + f[6] - f[7] - f[8];
    }
    double u = 0.0, v = 0.0, w = 0.0;
    if (rho > 0.0) {
        for (int it = 0; it < 3; ++it) {
            const double Fy = rho * kp.gbeta * (T - kp.tref);
            u = (mx + 0.5 * 0.0) / rho;
            v = (my + 0.5 * Fy) / rho;
            w = (mz + 0.5 * 0.0) / rho;
        }
    } else {
        atomicOr(bad, 1);
    }
    mac[c] = rho;
    mac[N + c] = u;
    mac[2 * N + c] = v;
    mac[3 * N + c] = w;
    mac[4 * N + c] = T;
    if (kp.use_g) {
        Bp[c] = u * T;
        Bp[N + c] = v * T;
        Bp[2 * N + c] = w * T;
    } else {
        Bp[c] = Bp[N + c] = Bp[2 * N + c] = 0.0;
    }
}

// ---------------------------------------------------------------------------------------
// Reductions: Neumaier-compensated FP64 per thread, pairwise two-sum across the wavefront
// and the block, one (sum, carry) pair per block and quantity.  The host adds the block pairs
// in a fixed order, so results do not depend on scheduling.
// ---------------------------------------------------------------------------------------
struct KS {
    double s, c;
};
__device__ __forceinline__ void ks_add(KS &a, double x)
{
    const double t = a.s + x;
    if (fabs(a.s) >= fabs(x)) a.c += (a.s - t) + x;
    else a.c += (x - t) + a.s;
    a.s = t;
}
__device__ __forceinline__ KS ks_merge(KS a, KS b)
{
    // two-sum of the leading parts, carries added
    const double s = a.s + b.s;
    const double bb = s - a.s;
    const double e = (a.s - (s - bb)) + (b.s - bb);
    return KS{s, a.c + b.c + e};
}

template <int NQ>
__device__ __forceinline__ void block_reduce_store(KS (&v)[NQ], double *out)
{
    __shared__ double sh_s[NQ][4], sh_c[NQ][4];  // up to 4 wavefronts of 64 lanes
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
        KS a = v[q];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            KS b{__shfl_down(a.s, off, 64), __shfl_down(a.c, off, 64)};
            a = ks_merge(a, b);
        }
        if (lane == 0) {
            sh_s[q][wid] = a.s;
            sh_c[q][wid] = a.c;
        }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        const int nw = (blockDim.x + 63) >> 6;
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            KS a{sh_s[q][0], sh_c[q][0]};
            for (int w = 1; w < nw; ++w) a = ks_merge(a, KS{sh_s[q][w], sh_c[q][w]});
            out[(blockIdx.x * NQ + q) * 2] = a.s;
            out[(blockIdx.x * NQ + q) * 2 + 1] = a.c;
        }
    }
}

// Volume sums over the macro planes: calNuRe, RBcalc_Nu_global and the conservation sums.
//   0: sum v T   1: sum |u|^2   2: sum qy (global Nu integrand)   3: sum rho   4: sum T
//   5: sum 0.5 rho |u|^2
// side = 1: the side-heated cell's formulas (hot wall x = 0): Nu_vol from u T
// (2DRB/3DRB.F90:1732-1741) and SideHeatedcalc_Nu_global (2DRB/3DRB.F90:1789-1828).
__global__ void __launch_bounds__(256) k_reduce_volume(KParams kp, const double *__restrict__ mac,
                                                       double L_over_kappa, double dy, int side,
                                                       double *__restrict__ out)
{
    const uint64_t N = kp.P;   // plane pitch
    const int nx = kp.nx, ny = kp.ny;
    KS acc[6] = {};
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < kp.N;
         c += (uint64_t)gridDim.x * blockDim.x) {
        const double rho = mac[c], u = mac[N + c], v = mac[2 * N + c], w = mac[3 * N + c];
        const double T = mac[4 * N + c];
        const double *Tc = mac + 4 * N + c;
        ks_add(acc[0], (side ? u : v) * T);
        const double q2 = u * u + v * v + w * w;
        ks_add(acc[1], q2);
        if (side) {
            const int i = (int)(c % (uint64_t)nx);
            double dT;
            if (i == 0) dT = (-3.0 * T - Tc[1] + 4.0 * kp.thot) / (3.0 * dy);
            else if (i == nx - 1) dT = (-4.0 * kp.tcold + 3.0 * T + Tc[-1]) / (3.0 * dy);
            else dT = (Tc[-1] - Tc[1]) / (2.0 * dy);
            ks_add(acc[2], L_over_kappa * u * (T - kp.tref) + dT);
        } else {
            const int j = (int)((c / (uint64_t)nx) % (uint64_t)ny);
            double dTdy;
            if (j == 0) dTdy = (3.0 * T + Tc[nx] - 4.0 * kp.thot) / (3.0 * dy);
            else if (j == ny - 1) dTdy = (4.0 * kp.tcold - 3.0 * T - Tc[-nx]) / (3.0 * dy);
            else dTdy = (Tc[nx] - Tc[-nx]) / (2.0 * dy);
            ks_add(acc[2], L_over_kappa * v * (T - kp.tref) - dTdy);
        }
        ks_add(acc[3], rho);
        ks_add(acc[4], T);
        ks_add(acc[5], 0.5 * rho * q2);
    }
    block_reduce_store<6>(acc, out);
}

// Wall sums of RBcalc_Nu_wall_avg: one block per x index i, reducing over z.
//   per-i outputs: T(i,1,:) average and hot-wall Nu average (plain arrays, length nx)
//   per-block pairs: 0 sum cold-wall Nu, 1 sum mid-plane Nu
// side = 1: SideHeatedcalc_Nu_wall_avg (2DRB/3DRB.F90:1889-2038), one block per y index j:
// hot wall x = 0, cold wall x = Lx, mid-plane x = Lx / 2 with u.
__global__ void __launch_bounds__(256) k_reduce_walls(KParams kp, const double *__restrict__ mac,
                                                      double L_over_kappa, double dy, int side,
                                                      double *__restrict__ tbot,
                                                      double *__restrict__ nubot,
                                                      double *__restrict__ out)
{
    const int i = blockIdx.x;
    const uint64_t N = kp.P;   // plane pitch
    const int nx = kp.nx, ny = kp.ny, nz = kp.nz;
    const double *T = mac + 4 * N, *v = mac + 2 * N, *u = mac + N;
    const double dT = kp.thot - kp.tcold;
    KS acc[4] = {};
    if (side) {
        const int j = blockIdx.x;
        const double dx = dy;
        for (int k = threadIdx.x; k < nz; k += blockDim.x) {
            const uint64_t b = (uint64_t)nx * ((uint64_t)j + (uint64_t)ny * k);  // (0, j, k)
            ks_add(acc[0], T[b]);
            ks_add(acc[1], ((8.0 * kp.thot - 9.0 * T[b] + T[b + 1]) / (3.0 * dx)) / dT);
            const uint64_t t = b + nx - 1;
            ks_add(acc[2], ((-8.0 * kp.tcold + 9.0 * T[t] - T[t - 1]) / (3.0 * dx)) / dT);
            double mid;
            if (nx % 2 == 1) {
                const uint64_t m = b + (nx + 1) / 2 - 1;
                mid = (L_over_kappa * u[m] * (T[m] - kp.tref) + (T[m - 1] - T[m + 1]) / (2.0 * dx)) / dT;
            } else {
                const uint64_t ml = b + nx / 2 - 1, mr = ml + 1;
                mid = (L_over_kappa * 0.5 * (u[ml] * (T[ml] - kp.tref) + u[mr] * (T[mr] - kp.tref)) +
                       (T[ml] - T[mr]) / dx) / dT;
            }
            ks_add(acc[3], mid);
        }
    }
    for (int k = side ? nz : (int)threadIdx.x; k < nz; k += blockDim.x) {
        const uint64_t b = (uint64_t)i + (uint64_t)nx * (uint64_t)ny * k;  // (i, 0, k)
        ks_add(acc[0], T[b]);
        ks_add(acc[1], ((8.0 * kp.thot - 9.0 * T[b] + T[b + nx]) / (3.0 * dy)) / dT);
        const uint64_t t = b + (uint64_t)nx * (ny - 1);
        ks_add(acc[2], ((-8.0 * kp.tcold + 9.0 * T[t] - T[t - nx]) / (3.0 * dy)) / dT);
        double mid;
        if (ny % 2 == 1) {
            const uint64_t m = b + (uint64_t)nx * ((ny + 1) / 2 - 1);
            mid = (L_over_kappa * v[m] * (T[m] - kp.tref) - (T[m + nx] - T[m - nx]) / (2.0 * dy)) / dT;
        } else {
            const uint64_t mb = b + (uint64_t)nx * (ny / 2 - 1), mt = mb + nx;
            mid = (L_over_kappa * 0.5 * (v[mb] * (T[mb] - kp.tref) + v[mt] * (T[mt] - kp.tref)) +
                   (T[mb] - T[mt]) / dy) / dT;
        }
        ks_add(acc[3], mid);
    }
    block_reduce_store<4>(acc, out);
    __syncthreads();
    if (threadIdx.x == 0) {
        const double s0 = out[(blockIdx.x * 4 + 0) * 2] + out[(blockIdx.x * 4 + 0) * 2 + 1];
        const double s1 = out[(blockIdx.x * 4 + 1) * 2] + out[(blockIdx.x * 4 + 1) * 2 + 1];
        tbot[i] = s0 / (double)nz;
        nubot[i] = s1 / (double)nz;
    }
}

// One plane of a macro field across `axis` (0: x, 1: y, 2: z) at the 0-based index pair
// (l, r): mode 0 interpolates (1 - w) f(l) + w f(r), or takes f(l) when l == r
// (interp_scalar_x/y/z, 3DRB.F90:2550-2600); mode 1 averages 0.5 (f(l) + f(r)) (the 2D
// mid-line averages, 2DRB.F90:2097-2108).  Output: the two in-plane axes, the lower one fastest.
__global__ void __launch_bounds__(256) k_plane(KParams kp, const double *__restrict__ f, int axis,
                                               int l, int r, double w, int mode,
                                               double *__restrict__ out)
{
    const uint64_t nx = kp.nx, ny = kp.ny, nz = kp.nz;
    const uint64_t np = axis == 0 ? ny * nz : (axis == 1 ? nx * nz : nx * ny);
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < np;
         t += (uint64_t)gridDim.x * blockDim.x) {
        uint64_t cl, cr;
        if (axis == 0) {
            const uint64_t j = t % ny, k = t / ny;
            cl = l + nx * (j + ny * k);
            cr = r + nx * (j + ny * k);
        } else if (axis == 1) {
            const uint64_t i = t % nx, k = t / nx;
            cl = i + nx * (l + ny * k);
            cr = i + nx * (r + ny * k);
        } else {
            const uint64_t i = t % nx, j = t / nx;
            cl = i + nx * (j + ny * l);
            cr = i + nx * (j + ny * r);
        }
        const double fl = f[cl], fr = f[cr];
        out[t] = mode ? 0.5 * (fl + fr) : (l == r ? fl : (1.0 - w) * fl + w * fr);
    }
}

// ---- fine-grid |psi| maximum of the 2D stream function (calc_psi_absmax_fine_spline,
// 2DRB/2DRB.F90:2873-2989).  Natural cubic splines (spline_natural / splint, :3000-3074) through
// [0, values, 0] on the node coordinates extended by the two walls; the same operations in the
// same order as the host restatement, so the maximum and its position agree bit for bit.
__device__ __forceinline__ double splint_dev(int n, const double *xa, const double *ya,
                                             const double *y2a, double x)
{
    int klo = 0, khi = n - 1;
    while (khi - klo > 1) {
        const int k = (khi + klo) / 2;
        if (xa[k] > x) khi = k;
        else klo = k;
    }
    const double h = xa[khi] - xa[klo];
    if (h == 0.0) return ya[klo];
    const double a = (xa[khi] - x) / h, b = (x - xa[klo]) / h;
    return a * ya[klo] + b * ya[khi] + ((a * a * a - a) * y2a[klo] + (b * b * b - b) * y2a[khi]) * (h * h) / 6.0;
}

// A: every row spline (ny rows of nxe values) at the nfx fine x positions -> pxf[k * nye + 1 + j]
// (the fine columns, wall zeros at 0 and nye - 1)
__global__ void __launch_bounds__(256) k_spline_rows(int nxe, int ny, int nfx, double xlen,
                                                     const double *__restrict__ xe,
                                                     const double *__restrict__ rows,
                                                     const double *__restrict__ y2rows,
                                                     double *__restrict__ cols)
{
    const int nye = ny + 2;
    const uint64_t total = (uint64_t)nfx * ny;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int j = (int)(t % ny), k = (int)(t / ny);
        const double xq = xlen * (double)k / (double)(nfx - 1);
        cols[(uint64_t)k * nye + 1 + j] =
            splint_dev(nxe, xe, rows + (uint64_t)j * nxe, y2rows + (uint64_t)j * nxe, xq);
        if (j == 0) {
            cols[(uint64_t)k * nye] = 0.0;
            cols[(uint64_t)k * nye + nye - 1] = 0.0;
        }
    }
}

// B: natural spline second derivatives of every fine column (one thread per column)
__global__ void __launch_bounds__(256) k_spline_cols(int nye, int nfx, const double *__restrict__ ye,
                                                     const double *__restrict__ cols,
                                                     double *__restrict__ y2, double *__restrict__ wk)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nfx) return;
    const double *y = cols + (uint64_t)k * nye;
    double *d = y2 + (uint64_t)k * nye, *u = wk + (uint64_t)k * nye;
    d[0] = 0.0;
    u[0] = 0.0;
    for (int i = 1; i < nye - 1; ++i) {
        const double sig = (ye[i] - ye[i - 1]) / (ye[i + 1] - ye[i - 1]);
        const double p = sig * d[i - 1] + 2.0;
        d[i] = (sig - 1.0) / p;
        u[i] = (6.0 * ((y[i + 1] - y[i]) / (ye[i + 1] - ye[i]) - (y[i] - y[i - 1]) / (ye[i] - ye[i - 1])) /
                    (ye[i + 1] - ye[i - 1]) - sig * u[i - 1]) / p;
    }
    d[nye - 1] = 0.0;
    for (int i = nye - 2; i >= 0; --i) d[i] = d[i] * d[i + 1] + u[i];
}

// C: |psi| on the fine grid (column k outer, y index l inner); per block the largest value and
// the first (k, l) where it occurs
__global__ void __launch_bounds__(256) k_spline_absmax(int nye, int nfx, int nfy, double ylen,
                                                       const double *__restrict__ ye,
                                                       const double *__restrict__ cols,
                                                       const double *__restrict__ y2,
                                                       double *__restrict__ bval,
                                                       unsigned long long *__restrict__ bidx)
{
    __shared__ double sv[256];
    __shared__ unsigned long long si[256];
    double best = -1.0;
    unsigned long long bi = ~0ULL;
    const uint64_t total = (uint64_t)nfx * nfy;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < total;
         t += (uint64_t)gridDim.x * blockDim.x) {
        const int l = (int)(t % nfy), k = (int)(t / nfy);
        const double yq = ylen * (double)l / (double)(nfy - 1);
        const double v = fabs(splint_dev(nye, ye, cols + (uint64_t)k * nye, y2 + (uint64_t)k * nye, yq));
        if (v > best || (v == best && t < bi)) {
            best = v;
            bi = t;
        }
    }
    sv[threadIdx.x] = best;
    si[threadIdx.x] = bi;
    __syncthreads();
    for (int s = blockDim.x / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            const double o = sv[threadIdx.x + s];
            const unsigned long long oi = si[threadIdx.x + s];
            if (o > sv[threadIdx.x] || (o == sv[threadIdx.x] && oi < si[threadIdx.x])) {
                sv[threadIdx.x] = o;
                si[threadIdx.x] = oi;
            }
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        bval[blockIdx.x] = sv[0];
        bidx[blockIdx.x] = si[0];
    }
}

hipError_t launch_psi_absmax(hipStream_t s, int nx, int ny, int nfx, int nfy, double xlen, double ylen,
                             const double *xe, const double *ye, const double *rows,
                             const double *y2rows, double *cols, double *y2cols, double *wk,
                             double *bval, unsigned long long *bidx, unsigned blocks)
{
    const int nxe = nx + 2, nye = ny + 2;
    const uint64_t ta = (uint64_t)nfx * ny;
    k_spline_rows<<<(unsigned)((ta + 255) / 256 < 65536 ? (ta + 255) / 256 : 65536), 256, 0, s>>>(
        nxe, ny, nfx, xlen, xe, rows, y2rows, cols);
    k_spline_cols<<<(nfx + 255) / 256, 256, 0, s>>>(nye, nfx, ye, cols, y2cols, wk);
    k_spline_absmax<<<blocks, 256, 0, s>>>(nye, nfx, nfy, ylen, ye, cols, y2cols, bval, bidx);
    return hipGetLastError();
}

// `check` sums: 0 sum |u - u_p|^2, 1 sum |u|^2, 2 sum |T - T_p|, 3 sum |T|; then u_p = u.
__global__ void __launch_bounds__(256) k_reduce_check(KParams kp, const double *__restrict__ mac,
                                                      double *__restrict__ prev,
                                                      double *__restrict__ out)
{
    const uint64_t N = kp.P;   // plane pitch
    KS acc[4] = {};
    for (uint64_t c = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; c < kp.N;
         c += (uint64_t)gridDim.x * blockDim.x) {
        const double u = mac[N + c], v = mac[2 * N + c], w = mac[3 * N + c], T = mac[4 * N + c];
        const double du = u - prev[c], dv = v - prev[N + c], dw = w - prev[2 * N + c];
        ks_add(acc[0], du * du + dv * dv + dw * dw);
        ks_add(acc[1], u * u + v * v + w * w);
        ks_add(acc[2], fabs(T - prev[3 * N + c]));
        ks_add(acc[3], fabs(T));
        prev[c] = u;
        prev[N + c] = v;
        prev[2 * N + c] = w;
        prev[3 * N + c] = T;
    }
    block_reduce_store<4>(acc, out);
}

// ---------------------------------------------------------------------------------------
// One-pass form (lbm_fused.h), FP64 exact arithmetic.
// ---------------------------------------------------------------------------------------
struct FormF64 {
    using R = double;
    using S = double;
    using B = double;
    static constexpr bool can_stage = false;   // staged loads move 4-byte elements
    static constexpr bool ring_last = false;   // one-pass thread order: row-major tile
    struct P {
        double thot, tcold, abb;
    };
    struct Node {};
    static __device__ __forceinline__ Node node(const KParams &, const P &, int, int) { return Node{}; }
    // a neighbour's post-collision population of direction a pulled into this node: unchanged
    static __device__ __forceinline__ double pulled(int, double v, const P &) { return v; }
    static __device__ __forceinline__ void round_g(double *) {}
    // prefetching into registers (lbm_fused.h, PF == 1 / 3): one load per pull at its source
    // address (the same rules as pull_flow / heat_in), the thermal wall arithmetic applied after
    // arrival from the returned modes (two bits per direction)
    template <int... A>
    static __device__ __forceinline__ void raw_f_seq(const double *F, uint64_t N, int i, int j,
                                                     int k, uint64_t c, bool border,
                                                     const KParams &kp, double *f,
                                                     std::integer_sequence<int, A...>)
    {
        ((f[A] = F[flow_elem<D3Q19, A>(N, i, j, k, c, kp)]), ...);
    }
    static __device__ __forceinline__ void raw_f(const double *F, uint64_t N, int i, int j, int k,
                                                 uint64_t c, bool border, const KParams &kp, double *f)
    {
        raw_f_seq(F, N, i, j, k, c, border, kp, f, std::make_integer_sequence<int, 19>{});
    }
    template <int... A>
    static __device__ __forceinline__ unsigned raw_g_seq(const double *G, uint64_t N, int i, int j,
                                                         int k, uint64_t c, bool border,
                                                         const KParams &kp, double *g,
                                                         std::integer_sequence<int, A...>)
    {
        unsigned modes = 0, m;
        ((g[A] = G[heat_elem<D3Q7, A>(N, i, j, k, c, kp, m)], modes |= m << (2 * A)), ...);
        return modes;
    }
    static __device__ __forceinline__ unsigned raw_g(const double *G, uint64_t N, int i, int j,
                                                     int k, uint64_t c, bool border,
                                                     const KParams &kp, double *g)
    {
        return raw_g_seq(G, N, i, j, k, c, border, kp, g, std::make_integer_sequence<int, 7>{});
    }
    // Interior waves (every lane an interior node of plane kq): population a of the node at
    // in-plane byte offset vo, loaded through a buffer descriptor based at the population's
    // upstream element for vo = 0 (buf_ld).  Thermal modes are all PULL_T0.
    static __device__ __forceinline__ void raw_f_inner(const double *F, uint64_t N, int kq,
                                                       uint64_t plane, uint32_t vo,
                                                       const KParams &kp, double *f)
    {
        const int64_t sy = kp.nx, sz = (int64_t)kp.nx * kp.ny, pk = (int64_t)kq * (int64_t)plane;
#pragma unroll
        for (int a = 0; a < 19; ++a) {
            const int64_t sh = D3Q19::ex[a] + D3Q19::ey[a] * sy + D3Q19::ez[a] * sz;
            f[a] = buf_ld(F + ((int64_t)a * (int64_t)N + pk - sh), vo);
        }
    }
    static __device__ __forceinline__ void raw_g_inner(const double *G, uint64_t N, int kq,
                                                       uint64_t plane, uint32_t vo,
                                                       const KParams &kp, double *g)
    {
        const int64_t sy = kp.nx, sz = (int64_t)kp.nx * kp.ny, pk = (int64_t)kq * (int64_t)plane;
#pragma unroll
        for (int a = 0; a < 7; ++a) {
            const int64_t sh = D3Q7::ex[a] + D3Q7::ey[a] * sy + D3Q7::ez[a] * sz;
            g[a] = buf_ld(G + ((int64_t)a * (int64_t)N + pk - sh), vo);
        }
    }
    static __device__ __forceinline__ double ld_inner(const double *P, int64_t elem, uint32_t vo)
    {
        return buf_ld(P + elem, vo);
    }
    static __device__ __forceinline__ void cook_f(const double *raw, double *f)
    {
#pragma unroll
        for (int a = 0; a < 19; ++a) f[a] = raw[a];
    }
    // as heat_in: adiabatic walls return the opposite population, isothermal walls
    // -g_post(opp a) + abb T_wall (the hot wall at the low index)
    static __device__ __forceinline__ void cook_g(const double *raw, unsigned modes, const P &p,
                                                  const Node &, double *g)
    {
#pragma unroll
        for (int a = 0; a < 7; ++a) {
            const bool low_wall = (D3Q7::ex[a] + D3Q7::ey[a] + D3Q7::ez[a]) > 0;
            const double tw = low_wall ? p.thot : p.tcold;
            g[a] = ((modes >> (2 * a)) & 3u) == PULL_ABB ? -raw[a] + p.abb * tw : raw[a];
        }
    }
    static __device__ __forceinline__ double tsum(const double *g)
    {
        return g[0] + g[1] + g[2] + g[3] + g[4] + g[5] + g[6];
    }
    static __device__ __forceinline__ void macro(const double *f, double Tn, const KParams &kp,
                                                 const P &, const Node &, double &rho, double &u,
                                                 double &v, double &w)
    {
        rho = sum19(f);
        const double Fy = rho * kp.gbeta * (Tn - kp.tref);
        u = jx19(f) / rho;
        v = (jy19(f) + 0.5 * Fy) / rho;
        w = jz19(f) / rho;
    }
    static __device__ __forceinline__ void thermal(const double *g, double *p, double T, double u,
                                                   double v, double w, double *bp,
                                                   const KParams &kp, const P &, const Node &)
    {
        ref_collide_d3q7(g, p, T, u, v, w, bp, kp);
    }
    static __device__ __forceinline__ void flow(const double *f, double *out, double rho, double u,
                                                double v, double w, double T1, const KParams &kp,
                                                const P &, const Node &)
    {
        ref_collide_d3q19(f, out, rho, u, v, w, T1, kp);
    }
    static __device__ __forceinline__ void store_g(double *G, uint64_t N, uint64_t c, const double *g)
    {
#pragma unroll
        for (int a = 0; a < 7; ++a) G[a * N + c] = g[a];
    }
    static __device__ __forceinline__ void store_f(double *F, uint64_t N, uint64_t c, const double *f)
    {
#pragma unroll
        for (int a = 0; a < 19; ++a) F[a * N + c] = f[a];
    }
    // stores of plane element pk at the lane's in-plane byte offset vo (buffer descriptors: no
    // 64-bit address arithmetic per store)
    static __device__ __forceinline__ void store_g_inner(double *G, uint64_t N, int64_t pk, uint32_t vo,
                                                         const double *g)
    {
#pragma unroll
        for (int a = 0; a < 7; ++a) buf_st(G + ((int64_t)a * (int64_t)N + pk), vo, g[a]);
    }
    static __device__ __forceinline__ void store_f_inner(double *F, uint64_t N, int64_t pk, uint32_t vo,
                                                         const double *f)
    {
#pragma unroll
        for (int a = 0; a < 19; ++a) buf_st(F + ((int64_t)a * (int64_t)N + pk), vo, f[a]);
    }
    // Wall value of thermal direction a at the node itself (adiabatic or constant T).
    static __device__ __forceinline__ double wall_g(int a, const double *gp, int kind, const P &p,
                                                    const Node &)
    {
        const double back = gp[D3Q7::opp[a]];
        if (kind == T_ADIABATIC) return back;
        const double tw = (a & 1) ? p.thot : p.tcold;
        return -back + p.abb * tw;
    }
};

} // namespace rblbm

#include "lbm_fused.h"

namespace rblbm {

size_t fused_instance_bytes_f64() { return sizeof(FusedInstance<FormF64>); }

void fused_instance_fill_f64(void *dst, const KParams &kp, const FusedGeom &fg, const FusedPtrs &p,
                             unsigned first_block)
{
    FusedInstance<FormF64> d;
    d.kp = kp;
    d.fp = FormF64::P{kp.thot, kp.tcold, kp.abb};
    d.fg = fg;
    d.Fcur = (const double *)p.Fcur;
    d.Fnext = (double *)p.Fnext;
    d.Gin = (const double *)p.Gin;
    d.Gout = (double *)p.Gout;
    d.Bp = (const double *)p.Bp;
    d.Bpout = (double *)p.Bpout;
    d.first_block = first_block;
    memcpy(dst, &d, sizeof(d));
}

hipError_t launch_fused_batch_f64(const FusedGeom &fg0, const void *inst, const uint16_t *owner,
                                  unsigned nblocks, hipStream_t s)
{
    return launch_fused_table<FormF64>(fg0, inst, owner, nblocks, s);
}

// ---------------------------------------------------------------------------------------
// Host-side launchers
// ---------------------------------------------------------------------------------------
static dim3 step_grid(const KParams &kp) { return dim3(kp.nblocks); }

hipError_t launch_prologue(const KParams &kp, int bx, hipStream_t s, const double *Fin,
                           double *Fout, const double *mac, double *fy_out, bool three_d, bool aa)
{
    if (three_d && aa) k3_prologue<true><<<step_grid(kp), bx, 0, s>>>(kp, nullptr, Fout, mac);
    else if (three_d) k3_prologue<false><<<step_grid(kp), bx, 0, s>>>(kp, Fin, Fout, mac);
    else k2_prologue<<<step_grid(kp), bx, 0, s>>>(kp, Fin, Fout, mac, fy_out);
    return hipGetLastError();
}

hipError_t launch_flow3_aa(const KParams &kp, int bx, hipStream_t s, int lay, double *F,
                           const double *G, const double *aux)
{
    const PullOffsets po = pull_offsets(kp, 8);
    if (lay == LAY_ODD) k3_flow<LAY_ODD><<<step_grid(kp), bx, 0, s>>>(kp, nullptr, F, G, aux, po);
    else if (lay == LAY_EVEN) k3_flow<LAY_EVEN><<<step_grid(kp), bx, 0, s>>>(kp, nullptr, F, G, aux, po);
    else return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_thermal3_aa(const KParams &kp, int bx, hipStream_t s, int lay, bool first,
                              const double *F, double *G, double *Bp, double *aux,
                              const double *mac)
{
    const PullOffsets po = pull_offsets(kp, 8);
    if (first && lay == LAY_ODD)
        k3_thermal<true, LAY_ODD><<<step_grid(kp), bx, 0, s>>>(kp, F, nullptr, G, Bp, aux, mac, po);
    else if (first)
        return hipErrorInvalidValue;   // the plain state is LAY_ODD
    else if (lay == LAY_ODD)
        k3_thermal<false, LAY_ODD><<<step_grid(kp), bx, 0, s>>>(kp, F, nullptr, G, Bp, aux, mac, po);
    else if (lay == LAY_EVEN)
        k3_thermal<false, LAY_EVEN><<<step_grid(kp), bx, 0, s>>>(kp, F, nullptr, G, Bp, aux, mac, po);
    else
        return hipErrorInvalidValue;
    return hipGetLastError();
}

hipError_t launch_step3(const KParams &kp, int bx, hipStream_t s, bool first, const double *Fcur,
                        double *Fnext, const double *Gin, double *Gout, double *Bp, double *aux,
                        const double *mac)
{
    const PullOffsets po = pull_offsets(kp, 8);
    if (first) k3_thermal<true, LAY_PULL><<<step_grid(kp), bx, 0, s>>>(kp, Fcur, Gin, Gout, Bp, aux, mac, po);
    else k3_thermal<false, LAY_PULL><<<step_grid(kp), bx, 0, s>>>(kp, Fcur, Gin, Gout, Bp, aux, mac, po);
    k3_flow<LAY_PULL><<<step_grid(kp), bx, 0, s>>>(kp, Fcur, Fnext, Gout, aux, po);
    return hipGetLastError();
}

hipError_t launch_step2(const KParams &kp, int bx, hipStream_t s, bool first, const double *Fcur,
                        double *Fnext, const double *Gin, double *Gout, double *Bp,
                        const double *fy_cur, double *fy_next, const double *mac)
{
    if (first) k2_thermal<true><<<step_grid(kp), bx, 0, s>>>(kp, Fcur, Gin, Gout, Bp, fy_cur, mac);
    else k2_thermal<false><<<step_grid(kp), bx, 0, s>>>(kp, Fcur, Gin, Gout, Bp, fy_cur, mac);
    k2_flow<<<step_grid(kp), bx, 0, s>>>(kp, Fcur, Fnext, Gout, fy_cur, fy_next);
    return hipGetLastError();
}

hipError_t launch_state(const KParams &kp, int bx, hipStream_t s, bool three_d, const double *Fprev,
                        const double *Gpost, const double *aux_or_fy, const double *Gprev,
                        double *mac, double *fo, double *go, int k0, int nk, int lay)
{
    if (three_d && lay == LAY_ODD)
        k3_state<LAY_ODD><<<step_grid(kp), bx, 0, s>>>(kp, Fprev, Gpost, aux_or_fy, nullptr, mac, fo, go, k0, nk);
    else if (three_d && lay == LAY_EVEN)
        k3_state<LAY_EVEN><<<step_grid(kp), bx, 0, s>>>(kp, Fprev, Gpost, aux_or_fy, nullptr, mac, fo, go, k0, nk);
    else if (three_d)
        k3_state<LAY_PULL><<<step_grid(kp), bx, 0, s>>>(kp, Fprev, Gpost, aux_or_fy, Gprev, mac, fo, go, k0, nk);
    else k2_state<<<step_grid(kp), bx, 0, s>>>(kp, Fprev, Gpost, aux_or_fy, mac, fo, go, k0, nk);
    return hipGetLastError();
}

hipError_t launch_init(const KParams &kp, int bx, hipStream_t s, const InitParams &ip, double *F,
                       double *G, double *Bp, double *mac, int qf, int qt)
{
    k_init<<<step_grid(kp), bx, 0, s>>>(kp, ip, F, G, Bp, mac, qf, qt);
    return hipGetLastError();
}

hipError_t launch_transpose(bool to_soa, hipStream_t s, const double *in, double *out, int q,
                            uint64_t N, uint64_t c0, uint64_t n)
{
    const unsigned blocks = (unsigned)((n + 255) / 256);
    if (blocks == 0) return hipSuccess;
    if (to_soa) k_aos_to_soa<<<blocks, 256, 0, s>>>(in, out, q, N, c0, n);
    else k_soa_to_aos<<<blocks, 256, 0, s>>>(in, out, q, N, c0, n);
    return hipGetLastError();
}

hipError_t launch_reconstruct(const KParams &kp, hipStream_t s, const double *F, const double *G,
                              double *Bp, double *mac, int qf, int qt, int *bad)
{
    const unsigned blocks = (unsigned)((kp.N + 255) / 256);
    k_reconstruct<<<blocks, 256, 0, s>>>(kp, F, G, Bp, mac, qf, qt, bad);
    return hipGetLastError();
}

hipError_t launch_reduce_volume(const KParams &kp, hipStream_t s, const double *mac,
                                double L_over_kappa, double dy, int side, double *out,
                                unsigned blocks)
{
    k_reduce_volume<<<blocks, 256, 0, s>>>(kp, mac, L_over_kappa, dy, side, out);
    return hipGetLastError();
}

hipError_t launch_reduce_walls(const KParams &kp, hipStream_t s, const double *mac,
                               double L_over_kappa, double dy, int side, double *tbot,
                               double *nubot, double *out)
{
    k_reduce_walls<<<side ? kp.ny : kp.nx, 256, 0, s>>>(kp, mac, L_over_kappa, dy, side, tbot,
                                                       nubot, out);
    return hipGetLastError();
}

hipError_t launch_plane(const KParams &kp, hipStream_t s, const double *f, int axis, int l, int r,
                        double w, int mode, double *out)
{
    const uint64_t np = axis == 0 ? (uint64_t)kp.ny * kp.nz
                                  : (axis == 1 ? (uint64_t)kp.nx * kp.nz : (uint64_t)kp.nx * kp.ny);
    const unsigned blocks = (unsigned)((np + 255) / 256 < 4096 ? (np + 255) / 256 : 4096);
    k_plane<<<blocks, 256, 0, s>>>(kp, f, axis, l, r, w, mode, out);
    return hipGetLastError();
}

__global__ void k_series_append(const double *__restrict__ a, size_t n1, const double *__restrict__ b,
                                size_t n2, double *__restrict__ series, size_t stride, int cap,
                                const int *__restrict__ count)
{
    const int slot = *count;
    if (slot < 0 || slot >= cap) return;
    double *dst = series + (size_t)slot * stride;
    for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n1 + n2;
         t += (size_t)gridDim.x * blockDim.x)
        dst[t] = t < n1 ? a[t] : b[t - n1];
}

__global__ void k_series_advance(int *count)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) *count += 1;
}

__global__ void k_series_reset(int *count, int n)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) count[t] = 0;
}

hipError_t launch_series_append(hipStream_t s, const double *a, size_t n1, const double *b,
                                size_t n2, double *series, size_t stride, int cap, int *count)
{
    const size_t n = n1 + n2;
    const unsigned blocks = (unsigned)((n + 255) / 256 < 1024 ? (n + 255) / 256 : 1024);
    k_series_append<<<blocks ? blocks : 1, 256, 0, s>>>(a, n1, b, n2, series, stride, cap, count);
    k_series_advance<<<1, 64, 0, s>>>(count);
    return hipGetLastError();
}

hipError_t launch_series_reset(hipStream_t s, int *count, int n)
{
    k_series_reset<<<(n + 255) / 256, 256, 0, s>>>(count, n);
    return hipGetLastError();
}

hipError_t launch_reduce_check(const KParams &kp, hipStream_t s, const double *mac, double *prev,
                               double *out, unsigned blocks)
{
    k_reduce_check<<<blocks, 256, 0, s>>>(kp, mac, prev, out);
    return hipGetLastError();
}

} // namespace rblbm
