// lbm_pull.h -- source addresses of the pulls with the reference's walls, for any storage type.
//
// Every kernel pulls f(a, x) = f_post(a, x - e_a) and applies the walls inline (the reference's
// streaming + bounceback sweeps, 2DRB/3DRB.F90:957-1058 and :1209-1351).  Written as an address
// choice followed by one load, a pull issues the same load instruction for every lane of a
// wave whether or not the lane sits at a wall: no branch of the pull writes a loaded register,
// so a load can stay in flight while later code runs (the one-pass kernel's prefetching).  The
// thermal pulls also report how the loaded value becomes the incoming one, to be applied once
// the load has arrived.
#pragma once

#include "lbm_lattice.h"

namespace rblbm {

// Loads of interior waves: a raw buffer descriptor based at a wave-uniform element (built in
// scalar registers) and the lane's 32-bit byte offset, the load's only vector operand.  Word 3
// is the raw-buffer format of the gfx9 family; the record count covers 4 GiB, so no offset
// below 2^32 is clamped.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t rsrc_at(const void *base)
{
    return __builtin_amdgcn_make_buffer_rsrc((void *)base, (short)0, -1, 0x00020000);
}
template <typename T> __device__ __forceinline__ T buf_ld(const T *base, uint32_t vo)
{
    if constexpr (sizeof(T) == 8)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(rsrc_at(base), vo, 0, 0));
    else if constexpr (sizeof(T) == 4)
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(rsrc_at(base), vo, 0, 0));
    else
        return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(rsrc_at(base), vo, 0, 0));
}
// Stores of interior nodes, addressed as buf_ld's loads.
template <typename T> __device__ __forceinline__ void buf_st(T *base, uint32_t vo, T v)
{
    const auto r = rsrc_at(base);
    if constexpr (sizeof(T) == 8)
        __builtin_amdgcn_raw_buffer_store_b64(
            __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b64(r, 0, 0, 0)), v), r, vo, 0, 0);
    else if constexpr (sizeof(T) == 4)
        __builtin_amdgcn_raw_buffer_store_b32(
            __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b32(r, 0, 0, 0)), v), r, vo, 0, 0);
    else
        __builtin_amdgcn_raw_buffer_store_b16(
            __builtin_bit_cast(decltype(__builtin_amdgcn_raw_buffer_load_b16(r, 0, 0, 0)), v), r, vo, 0, 0);
}
// The stored element a pull of direction A reads: its upstream neighbour's, the periodic image,
// or (walls) the node's own population of the opposite direction.  The thermal version also
// says how the read value becomes the incoming one: PULL_T0 (plus the shift correction),
// PULL_BACK (adiabatic wall: the opposite population as it is), PULL_ABB (isothermal wall:
// anti-bounce-back).
enum : unsigned { PULL_T0 = 0, PULL_BACK = 1, PULL_ABB = 2 };

template <class L, typename S, int A>
__device__ __forceinline__ const S *flow_src(const S *__restrict__ P, uint64_t N, int i, int j,
                                             int k, uint64_t c, bool border, const KParams &kp)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    const int sy = kp.nx, sz = kp.nx * kp.ny;
    if (!border) return P + A * N + c - (ex + ey * sy + ez * sz);
    const bool ox = (unsigned)(i - ex) >= (unsigned)kp.nx;
    const bool oy = (unsigned)(j - ey) >= (unsigned)kp.ny;
    const bool oz = (unsigned)(k - ez) >= (unsigned)kp.nz;
    if (oz) {
        if (kp.vel_z == VEL_PERIODIC) return P + A * N + c + (int64_t)(ez == 1 ? kp.nz - 1 - k : -k) * sz;
        return P + L::opp[A] * N + c;
    }
    if (oy) return P + L::opp[A] * N + c;
    if (ox) {
        if (kp.vel_x == VEL_PERIODIC) return P + A * N + c + (int64_t)(ex == 1 ? kp.nx - 1 - i : -i);
        return P + L::opp[A] * N + c;
    }
    return P + A * N + c - (int64_t)(ex + ey * sy + ez * sz);
}

template <class L, typename S, int A>
__device__ __forceinline__ const S *heat_src(const S *__restrict__ P, uint64_t N, int i, int j,
                                             int k, uint64_t c, bool border, const KParams &kp,
                                             unsigned &mode)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    const int sy = kp.nx, sz = kp.nx * kp.ny;
    mode = PULL_T0;
    if (!border) return P + A * N + c - (ex + ey * sy + ez * sz);
    int kind = -1;
    int64_t wrap = 0;
    if (ex != 0) {
        if ((unsigned)(i - ex) >= (unsigned)kp.nx) { kind = kp.temp_x; wrap = ex == 1 ? kp.nx - 1 - i : -i; }
    } else if (ey != 0) {
        if ((unsigned)(j - ey) >= (unsigned)kp.ny) { kind = kp.temp_y; wrap = (int64_t)(ey == 1 ? kp.ny - 1 - j : -j) * sy; }
    } else if (ez != 0) {
        if ((unsigned)(k - ez) >= (unsigned)kp.nz) { kind = kp.temp_z; wrap = (int64_t)(ez == 1 ? kp.nz - 1 - k : -k) * sz; }
    }
    if (kind < 0) return P + A * N + c - (int64_t)(ex + ey * sy + ez * sz);
    // a periodic pair of walls is never the shift axis (it has two isothermal walls)
    if (kind == T_PERIODIC) return P + A * N + c + wrap;
    mode = kind == T_ADIABATIC ? PULL_BACK : PULL_ABB;
    return P + L::opp[A] * N + c;
}

// The same choices written as selects (every lane runs the same instructions, no branch), as
// the element index of the source relative to P: for the waves of the one-pass kernel that
// hold wall nodes, whose loads stay in flight across the plane.  The axes are applied in the
// order x, y, z, so the last outside axis decides, as in flow_src.
template <class L, int A>
__device__ __forceinline__ uint64_t flow_elem(uint64_t N, int i, int j, int k, uint64_t c,
                                              const KParams &kp)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    const int64_t sy = kp.nx, sz = (int64_t)kp.nx * kp.ny;
    int64_t d = -(ex + ey * sy + ez * sz);   // from the node, within plane A
    bool back = false;                        // the node's own slot of the opposite direction
    if constexpr (ex != 0) {
        const bool ox = (unsigned)(i - ex) >= (unsigned)kp.nx;
        const bool per = kp.vel_x == VEL_PERIODIC;
        const int64_t w = ex == 1 ? kp.nx - 1 - i : -i;
        d = ox ? (per ? w : 0) : d;
        back = ox && !per;
    }
    if constexpr (ey != 0) {
        const bool oy = (unsigned)(j - ey) >= (unsigned)kp.ny;
        d = oy ? 0 : d;
        back = oy || back;
    }
    if constexpr (ez != 0) {
        const bool oz = (unsigned)(k - ez) >= (unsigned)kp.nz;
        const bool per = kp.vel_z == VEL_PERIODIC;
        const int64_t w = (int64_t)(ez == 1 ? kp.nz - 1 - k : -k) * sz;
        d = oz ? (per ? w : 0) : d;
        back = oz ? !per : back;
    }
    return (uint64_t)(back ? L::opp[A] : A) * N + c + (uint64_t)d;
}

template <class L, int A>
__device__ __forceinline__ uint64_t heat_elem(uint64_t N, int i, int j, int k, uint64_t c,
                                              const KParams &kp, unsigned &mode)
{
    constexpr int ex = L::ex[A], ey = L::ey[A], ez = L::ez[A];
    const int64_t sy = kp.nx, sz = (int64_t)kp.nx * kp.ny;
    int64_t d = -(ex + ey * sy + ez * sz);
    mode = PULL_T0;
    if constexpr (ex != 0 || ey != 0 || ez != 0) {
        // one non-zero component: the wall of that axis
        bool out;
        int kind;
        int64_t w;
        if constexpr (ex != 0) {
            out = (unsigned)(i - ex) >= (unsigned)kp.nx;
            kind = kp.temp_x;
            w = ex == 1 ? kp.nx - 1 - i : -i;
        } else if constexpr (ey != 0) {
            out = (unsigned)(j - ey) >= (unsigned)kp.ny;
            kind = kp.temp_y;
            w = (int64_t)(ey == 1 ? kp.ny - 1 - j : -j) * sy;
        } else {
            out = (unsigned)(k - ez) >= (unsigned)kp.nz;
            kind = kp.temp_z;
            w = (int64_t)(ez == 1 ? kp.nz - 1 - k : -k) * sz;
        }
        const bool per = kind == T_PERIODIC, wall = out && !per;
        d = out ? (per ? w : 0) : d;
        mode = wall ? (kind == T_ADIABATIC ? PULL_BACK : PULL_ABB) : PULL_T0;
        return (uint64_t)(wall ? L::opp[A] : A) * N + c + (uint64_t)d;
    }
    return (uint64_t)A * N + c;
}

} // namespace rblbm
