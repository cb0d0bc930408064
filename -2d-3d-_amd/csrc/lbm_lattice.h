// lbm_lattice.h -- lattice tables and shared device helpers for the gfx950 kernels.
//
// Direction order, opposite table and weights follow the reference so that the C-ABI gather
// returns populations in its order: D3Q19 / D3Q7 in 2DRB/3DRB.F90:656-678, D2Q9 / D2Q5 in
// 2DRB/2DRB.F90:188-190 and :506-524.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace rblbm {

struct D3Q19 {
    static constexpr int Q = 19;
    static constexpr int ex[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
    static constexpr int ey[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
    static constexpr int ez[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
    static constexpr int opp[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
};
struct D3Q7 {
    static constexpr int Q = 7;
    static constexpr int ex[7] = {0, 1, -1, 0, 0, 0, 0};
    static constexpr int ey[7] = {0, 0, 0, 1, -1, 0, 0};
    static constexpr int ez[7] = {0, 0, 0, 0, 0, 1, -1};
    static constexpr int opp[7] = {0, 2, 1, 4, 3, 6, 5};
};
struct D2Q9 {
    static constexpr int Q = 9;
    static constexpr int ex[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static constexpr int ey[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    static constexpr int ez[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr int opp[9] = {0, 3, 4, 1, 2, 7, 8, 5, 6};
};
struct D2Q5 {
    static constexpr int Q = 5;
    static constexpr int ex[5] = {0, 1, 0, -1, 0};
    static constexpr int ey[5] = {0, 0, 1, 0, -1};
    static constexpr int ez[5] = {0, 0, 0, 0, 0};
    static constexpr int opp[5] = {0, 3, 4, 1, 2};
};

// Weights of the flow lattices (rest, axis, diagonal) and of the thermal lattices (rest, axis).
__host__ __device__ constexpr double w19(int a) { return a == 0 ? 1.0 / 3.0 : (a <= 6 ? 1.0 / 18.0 : 1.0 / 36.0); }
__host__ __device__ constexpr double w7(int a) { return a == 0 ? 0.25 : 0.125; }
__host__ __device__ constexpr double w9(int a) { return a == 0 ? 4.0 / 9.0 : (a <= 4 ? 1.0 / 9.0 : 1.0 / 36.0); }
__host__ __device__ constexpr double w5(int a) { return a == 0 ? 1.0 / 3.0 : 1.0 / 6.0; }

// Wall codes (mirror LBM_VEL_* / LBM_T_* of lbm_rb.h).
enum { VEL_NOSLIP = 0, VEL_PERIODIC = 1 };
enum { T_CONST = 0, T_ADIABATIC = 1, T_PERIODIC = 2 };

// Population layouts of the two-pass step.
//   LAY_PULL: two arrays per lattice; a node stores its post-collision populations in its own
//     slots of the other array, and the next pass pulls population a of node x from slot
//     (a, x - e_a) (or, at a wall, from its own slot of the opposite direction).
//   In-place (A-A) layout: one array per lattice, every face a wall (no periodic axis).  In
//     state LAY_ODD the incoming population a of node x (what the pull returns) sits in slot
//     (a, x).  In state LAY_EVEN it sits in slot (opp a, x - e_a), or in (a, x) when x - e_a
//     is outside (the value the wall sends back).  A node that stores its post-collision
//     populations into the slots it read turns one state into the other:
//       from LAY_ODD:  f*_a -> (opp a, x)
//       from LAY_EVEN: f*_a -> (a, x + e_a), or (opp a, x) when x + e_a is outside.
//     Each slot is read and written by one node only, so a pass updates the array in place.
enum { LAY_PULL = 0, LAY_ODD = 1, LAY_EVEN = 2 };

// Everything a step kernel needs, passed by value (lives in SGPRs / kernarg segment).
struct KParams {
    int nx, ny, nz;
    int vel_x, vel_z;
    int temp_x, temp_y, temp_z;
    int use_g;
    int nxb;          // x blocks per row
    unsigned nblocks; // total blocks of the launch (for the XCD remap)
    uint64_t N;       // nodes
    uint64_t P;       // plane pitch (elements) of every plane-structured array: F, G, B_prev,
                      // aux, macros; N rounded up so that planes do not start at the same
                      // offset modulo the memory interleave
    double gbeta, tref, thot, tcold;
    double sv, sq;    // 1/tau_f (s_e = s_eps = s_nu = s_pi) and s_q = s_m
    double qk, qnu, sg;
    // thermal scheme constants (default or EnableLegacyThermalScheme): the energy-moment
    // equilibrium factor (neq = neqc T), the anti-bounce-back wall factor (g_in = -g_out(opp) +
    // abb Tw) and the rest / moving thermal weights (initial state)
    double neqc, abb, wt0, wt1;
};

// Rows with no y or z wall, two arrays: every pull is one load at a fixed byte offset from the
// array (population plane minus streaming shift, computed on the host) plus the node's own
// 32-bit byte offset.  The offsets stay in scalar registers and the loads take the node offset
// as their only vector operand, so a wave spends no per-population address arithmetic.  Nodes
// on an x wall then redo their pulls with the generic wall rules; the addresses loaded are
// those of the generic pulls, so the results are identical.
// The in-place layout uses the same scheme: state LAY_ODD reads slot (a, x) and writes
// (opp a, x), both at s[]; state LAY_EVEN reads (opp a, x - e_a) and writes (a, x + e_a).
struct PullOffsets {
    int64_t f[19];    // bytes: (a P - (ex + ey nx + ez nx ny)) es, flow lattice
    int64_t g[7];     // the same for the thermal lattice
    int64_t s[19];    // bytes: a P es, the node's own slot of direction a
    int64_t fe[19];   // bytes: (opp(a) P - (ex + ey nx + ez nx ny)) es, read in LAY_EVEN
    int64_t ge[7];
    int64_t fw[19];   // bytes: (a P + (ex + ey nx + ez nx ny)) es, written from LAY_EVEN
    int64_t gw[7];
    int fast;         // 1: P es < 2^32, so node byte offsets fit 32 bits
};

inline PullOffsets pull_offsets(const KParams &kp, int64_t es)
{
    PullOffsets po;
    const int64_t P = (int64_t)kp.P, sy = kp.nx, sz = (int64_t)kp.nx * kp.ny;
    for (int a = 0; a < 19; ++a) {
        const int64_t sh = D3Q19::ex[a] + D3Q19::ey[a] * sy + D3Q19::ez[a] * sz;
        po.f[a] = (a * P - sh) * es;
        po.s[a] = a * P * es;
        po.fe[a] = (D3Q19::opp[a] * P - sh) * es;
        po.fw[a] = (a * P + sh) * es;
    }
    for (int a = 0; a < 7; ++a) {
        const int64_t sh = D3Q7::ex[a] + D3Q7::ey[a] * sy + D3Q7::ez[a] * sz;
        po.g[a] = (a * P - sh) * es;
        po.ge[a] = (D3Q7::opp[a] * P - sh) * es;
        po.gw[a] = (a * P + sh) * es;
    }
    po.fast = (uint64_t)P * (uint64_t)es < ((uint64_t)1 << 32);
    return po;
}

// A row takes the offset path when it has no y or z wall (a block of a two-pass launch covers
// part of one row, so the choice is the same for the whole block).
__device__ __forceinline__ bool offset_row(const PullOffsets &po, int j, int k, const KParams &kp)
{
    return po.fast && j > 0 && k > 0 && j < kp.ny - 1 && k < kp.nz - 1;
}

// Launch geometry of the one-pass kernel (lbm_fused.h): ntx x nty tiles of (TX-2) x (TY-2)
// nodes, z split into chunks of kchunk planes; tile x fastest, then tile y, then chunk.
struct FusedGeom {
    unsigned ntx, nty, nzc, nblocks;
    int kchunk;
    int tile;     // tile shape: 0 = 32 x 16 threads, 1 = 16 x 16, 2 = 32 x 8, 3 = 32 x 32,
                  // 4 = 34 x 15, 5 = 66 x 15
    int pf;       // next plane's loads issued before the current flow collision: 1 into
                  // registers, 2 into LDS (staged)
};

// Blocks are dealt to the 8 XCDs round-robin by hardware.  Renumber them so that each XCD
// owns one contiguous run of logical blocks (a z-slab of the domain): the neighbour rows a
// pull reads then sit in the same XCD's L2 instead of in another XCD's.
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned nblocks)
{
    constexpr unsigned XCD = 8;
    unsigned per = nblocks / XCD, rem = nblocks % XCD;
    unsigned x = b % XCD, r = b / XCD;
    // XCD x holds per (+1 for the first rem) consecutive logical ids
    unsigned base = x * per + (x < rem ? x : rem);
    unsigned cnt = per + (x < rem ? 1u : 0u);
    return r < cnt ? base + r : b;  // b beyond the even split keeps identity (not reached)
}

// A pointer read from device memory (a parameter table) is generic to the compiler, which then
// emits flat loads and stores; flat operations share the counter of the scalar loads that
// re-read table parameters, so each such reload would wait for every load in flight.  The
// pointer is cast to the global address space and made opaque (an empty asm on a uniform
// value), so that the cast back to a generic pointer is not folded away and the compiler
// infers global operations for every access derived from it.  p must be wave-uniform.
template <typename T> __device__ __forceinline__ T *gptr(T *p)
{
    auto q = (__attribute__((address_space(1))) T *)p;
    asm("" : "+s"(q));
    return (T *)q;
}

} // namespace rblbm
