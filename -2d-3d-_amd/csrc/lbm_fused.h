// lbm_fused.h -- one-pass 3D step: both lattices advance one time step in a single launch.
//
// One reference step n needs two neighbour hops: u^{n+1} at a node needs f_post^n of its
// neighbours, and T^{n+1} at a node needs the thermal collision (which needs u^{n+1}) of its six
// face neighbours.  The two-pass form splits the hops over two launches and reads f twice.
// This kernel instead gives each block a 2D tile of nodes plus a one-node ring around it, and
// marches the tile through a chunk of z-planes:
//
//   iteration kk:  thermal part at plane kk for tile + ring   (pull g^n, pull f^{n+1} -> u^{n+1},
//                  thermal collision -> g_post^n; tile nodes store g_post^n and B_prev)
//                  exchange the four in-plane components of g_post^n through LDS
//                  flow part at plane p = kk - 1 for the tile  (T^{n+1} = pull-sum of g_post^n:
//                  in-plane neighbours from LDS, z neighbours from the thread's own registers;
//                  flow collision of step n+1 with f^{n+1}(p) kept from the previous iteration)
//
// B_prev is read from one buffer and written to another: a ring node reads the history of a
// node that the neighbouring tile updates during the same launch.
// Ring nodes repeat the thermal part of their owner tile (their loads mostly hit L2); every
// population and history value of a tile node crosses HBM once per step: 29 loads and 29
// stores per node, the algorithmic minimum.  The planes k0 - 1 and k1 of a z-chunk are
// likewise recomputed by both neighbouring chunks.
//
// A periodic thermal wall would need g_post^n of the opposite side of the domain, which no
// tile holds; configurations with periodic thermal walls use the two-pass form.
//
// Tile shapes and prefetching are compile-time variants (FusedGeom::tile / ::pf pick one at run
// time): one 512-thread block per CU keeps every wave in step with the plane barrier, while
// 256-thread tiles let two independent blocks share a CU.  Prefetching issues the next plane's
// loads before the current flow collision, so that their latency hides behind it:
//   PF = 1  into registers (29 more live registers per lane);
//   PF = 2  staged: each lane loads the elements its pulls read straight into LDS
//           (global_load_lds_dword; a wave fills 64 consecutive slots of one population), so
//           no register holds them in flight.  4-byte storage (the FP32 form) only.  The
//           flow stores of a plane are deferred past the next plane's wait, which waits for
//           every vector memory operation of the wave.
// A tile row of 32 interior nodes spans one 128-byte line in FP32 and two in FP64; its two ring
// columns touch the neighbouring lines, so the narrow tiles fetch up to 3x the FP32 rows they
// use.  The 66x15 tile (64 interior columns, 16 waves, 128 registers per lane: reduced-precision
// forms) halves that share.
//
// The including translation unit provides `Form` with the storage type S, arithmetic type R,
// history type B, a per-column constant record Node, and the per-node operations (pulls, the
// pull of a neighbour's value exchanged through LDS, macro, collisions) of its precision form.
#pragma once

#include "lbm_lattice.h"

namespace rblbm {

// One block's work: tile number lb (tile x fastest, then tile y, then z-chunk) of one domain.
template <class Form, int TX, int TY, int PF>
__device__ __forceinline__ void
fused_tile(const KParams &kp, const typename Form::P &fp, const FusedGeom &fg, unsigned lb,
           const typename Form::S *__restrict__ Fcur, typename Form::S *__restrict__ Fnext,
           const typename Form::S *__restrict__ Gin, typename Form::S *__restrict__ Gout,
           const typename Form::B *__restrict__ Bp, typename Form::B *__restrict__ Bpout)
{
    using R = typename Form::R;
    using S = typename Form::S;
    constexpr int NT = TX * TY;
    __shared__ R lds[3][4][NT];
    // PF == 2: slot stage[a][tid] receives pulled population a (flow 0..18, thermal 19..25) of
    // the next plane; NTP rounds the slot rows up to whole waves
    constexpr int NTP = (NT + 63) / 64 * 64;
    __shared__ S stage[PF == 2 ? 26 : 1][PF == 2 ? NTP : 1];
    static_assert(PF != 2 || Form::can_stage, "staged loads need 4-byte storage");
    constexpr bool REGS = PF == 1;   // the next plane's loads go into registers

    const unsigned tile_x = lb % fg.ntx;
    const unsigned rest = lb / fg.ntx;
    const unsigned tile_y = rest % fg.nty;
    const unsigned chunk = rest / fg.nty;
    const int tid = threadIdx.x;
    // Thread order.  Form::ring_last: threads take the tile's interior row by row first, then
    // the ring (the two ring rows, then the two ring columns), so that an interior row of 32 or
    // 64 nodes lies in one wave, its loads and stores cover whole 128-byte lines, and ring nodes
    // share waves with ring nodes only.  Otherwise the tile is row-major (each wave holds parts
    // of two rows with their ring columns).  Measured at 512^3: the FP32 staged 66x15 step
    // 7.54-7.58 -> 6.92-6.97 ms with ring-last order; the FP64 34x15 register form 14.34-14.38
    // -> 14.86-14.89 ms (2.5 GB more read per step), so FP64 keeps the row-major order.
    // s is the node's position in the tile (the slot of the LDS exchange).
    constexpr int IX = TX - 2, IY = TY - 2, NI = IX * IY;
    int lx, ly;
    if constexpr (Form::ring_last) {
        if (tid < NI) {
            lx = 1 + tid % IX;
            ly = 1 + tid / IX;
        } else {
            const int r = tid - NI;   // < 2 TX + 2 IY
            lx = r < 2 * TX ? r % TX : (r < 2 * TX + IY ? 0 : TX - 1);
            ly = r < 2 * TX ? (r < TX ? 0 : TY - 1) : 1 + (r - 2 * TX) % IY;
        }
    } else {
        lx = tid % TX;
        ly = tid / TX;
    }
    const int s = ly * TX + lx;
    const int i = (int)tile_x * (TX - 2) + lx - 1;
    const int j = (int)tile_y * (TY - 2) + ly - 1;
    const int nx = kp.nx, ny = kp.ny, nz = kp.nz;
    const bool in_dom = (unsigned)i < (unsigned)nx && (unsigned)j < (unsigned)ny;
    const bool xring = lx == 0 || lx == TX - 1, yring = ly == 0 || ly == TY - 1;
    const bool active = in_dom && !(xring && yring);
    const bool tile = in_dom && !xring && !yring;
    const int k0 = (int)chunk * fg.kchunk;
    const int k1 = min(nz, k0 + fg.kchunk);
    const uint64_t N = kp.P;   // plane pitch
    const uint64_t col = (uint64_t)i + (uint64_t)nx * (uint64_t)j;   // valid when in_dom
    const uint64_t plane = (uint64_t)nx * (uint64_t)ny;
    const typename Form::Node nd = Form::node(kp, fp, i, j);          // per-column constants

    R fprev[19];
    R rho_p = 0, u_p = 0, v_p = 0, w_p = 0;
    R gprev[7] = {};
    R g5pp = 0;
    bool have_prev = false, have_pp = false;

    // Loads of the thermal part of plane kq: pulled g^n, pulled f^{n+1}, B_prev, as the stored
    // elements (one load each at the pull's source address) and the thermal pulls' wall modes.
    // Form::cook_f / cook_g turn them into the incoming values once they have arrived, so that
    // no arithmetic on a loaded value sits between a prefetch and its use.
    // A wave whose lanes are all interior nodes (their ring corners included, which load and
    // discard) takes the interior loads: a scalar base per population and the lane's in-plane
    // byte offset (which fits 32 bits when a plane of populations does).  Other waves choose
    // each lane's source address with the wall rules (Form::raw_g / raw_f).  Either way every
    // pull is one load with no branch between it and its use.
    // Lanes outside `active` load too (their results are unused): lanes outside the domain at
    // the clamped node, so that every lane writes the registers it prefetches into and no old
    // value has to be kept for it.
    using B = typename Form::B;
    const bool col_inner = in_dom && i > 0 && j > 0 && i < nx - 1 && j < ny - 1 &&
                           plane * sizeof(S) < ((uint64_t)1 << 32) && plane * sizeof(B) < ((uint64_t)1 << 32);
    const uint32_t vo_s = (uint32_t)(col * sizeof(S)), vo_b = (uint32_t)(col * sizeof(B));
    // Tile nodes store through buffer descriptors based at their plane's element, with the
    // lane's in-plane byte offset vo_s / vo_b: no 64-bit address arithmetic per store.  The host
    // runs the one-pass form only when a plane's bytes fit 32 bits (fused_plane_fits).
    auto store_f_at = [&](int q, const R *v) {
        Form::store_f_inner(Fnext, N, (int64_t)q * (int64_t)plane, vo_s, v);
    };
    const int ic = min(max(i, 0), nx - 1), jc = min(max(j, 0), ny - 1);
    const uint64_t colc = (uint64_t)ic + (uint64_t)nx * (uint64_t)jc;
    auto load = [&](int kq, S(&lg)[7], S(&lf)[19], B(&lb)[3], unsigned &modes) -> bool {
        if (kq < 0 || kq >= nz) return false;
        const bool inner = col_inner && kq > 0 && kq < nz - 1;
        if (__builtin_amdgcn_ballot_w64(!inner) == 0) {
            modes = 0;
            Form::raw_g_inner(Gin, N, kq, plane, vo_s, kp, lg);
            Form::raw_f_inner(Fcur, N, kq, plane, vo_s, kp, lf);
            const int64_t pk = (int64_t)kq * (int64_t)plane;
            lb[0] = Form::ld_inner(Bp, pk, vo_b);
            lb[1] = Form::ld_inner(Bp, (int64_t)N + pk, vo_b);
            lb[2] = Form::ld_inner(Bp, 2 * (int64_t)N + pk, vo_b);
        } else {
            const uint64_t c = colc + plane * (uint64_t)kq;
            modes = Form::raw_g(Gin, N, ic, jc, kq, c, true, kp, lg);
            Form::raw_f(Fcur, N, ic, jc, kq, c, true, kp, lf);
            lb[0] = Bp[c];
            lb[1] = Bp[N + c];
            lb[2] = Bp[2 * N + c];
        }
        return active;
    };
    // PF == 2: the same loads, staged.  The lane's pulls go to its LDS slots (the wave's slots
    // start at wbase); the wall mode of each thermal pull and the history stay in registers.
    // Their destination is LDS, so a branch around them holds no register, and they keep one
    // address per active lane: loading every lane, interior waves through buffer descriptors
    // as in `load`, read 18% more from memory (26.9 instead of 22.6 GB per 512^3 FP32 step).
    const int wbase = tid & ~63;
    auto stage_load = [&](int kq, unsigned &modes, B(&lb)[3]) -> bool {
        if constexpr (PF == 2) {
            if (!active || kq < 0 || kq >= nz) return false;
            const uint64_t c = col + plane * (uint64_t)kq;
            const bool border = i == 0 || j == 0 || kq == 0 || i == nx - 1 || j == ny - 1 || kq == nz - 1;
            Form::stage_f(Fcur, N, i, j, kq, c, border, kp, &stage[0][wbase], NTP);
            modes = Form::stage_g(Gin, N, i, j, kq, c, border, kp, &stage[19][wbase], NTP);
            lb[0] = Bp[c];
            lb[1] = Bp[N + c];
            lb[2] = Bp[2 * N + c];
            return true;
        } else {
            return false;
        }
    };
    // With PF the loads of plane kk + 1 are issued before the flow part of iteration kk, so
    // that their latency hides behind that collision.
    S qg[7] = {}, qf[19] = {};
    B qb[3] = {};
    unsigned qmodes = 0;
    bool q_have = REGS ? load(k0 - 1, qg, qf, qb, qmodes)
                       : (PF == 2 ? stage_load(k0 - 1, qmodes, qb) : false);
    // PF == 2: the flow stores of plane `fpend_p`, issued after the next wait
    R fout[PF == 2 ? 19 : 1];
    bool fpend = false;
    int fpend_p = 0;

    for (int kk = k0 - 1; kk <= k1; ++kk) {
        R lg[7], fnew[19];
        B bp[3];
        bool have;
        if constexpr (PF == 2) {
            have = q_have;
            // every vector memory operation of this wave has completed, the staged loads of
            // plane kk among them: the LDS slots hold this lane's pulls.  The builtin (vmcnt 0,
            // the other counters at their maxima) also tells the compiler, which otherwise
            // waits once more for the loads it saw issued into LDS at the next barrier.
            __builtin_amdgcn_s_waitcnt(0x0F70);
            asm volatile("" ::: "memory");
            if (fpend) {
                store_f_at(fpend_p, fout);
                fpend = false;
            }
            if (have) {
                Form::unstage_f(&stage[0][tid], NTP, fnew);
                Form::unstage_g(&stage[19][tid], NTP, qmodes, fp, nd, lg);
#pragma unroll
                for (int a = 0; a < 3; ++a) bp[a] = qb[a];
            }
        } else if constexpr (REGS) {
            have = q_have;
            Form::cook_g(qg, qmodes, fp, nd, lg);
            Form::cook_f(qf, fnew);
#pragma unroll
            for (int a = 0; a < 3; ++a) bp[a] = qb[a];
        } else {
            S rg[7] = {}, rf[19] = {};
            unsigned modes = 0;
            have = load(kk, rg, rf, bp, modes);
            Form::cook_g(rg, modes, fp, nd, lg);
            Form::cook_f(rf, fnew);
        }
        R gnew[7] = {};
        R rho_n = 0, u_n = 0, v_n = 0, w_n = 0;
        if (have) {
            const R Tn = Form::tsum(lg);
            Form::macro(fnew, Tn, kp, fp, nd, rho_n, u_n, v_n, w_n);
            Form::thermal(lg, gnew, Tn, u_n, v_n, w_n, bp, kp, fp, nd);
            Form::round_g(gnew);   // exchanged and stored as the storage rounds them
            if (tile && kk >= k0 && kk < k1) {
                const int64_t pk = (int64_t)kk * (int64_t)plane;
                Form::store_g_inner(Gout, N, pk, vo_s, gnew);
                if (kp.use_g) {
                    buf_st(Bpout + pk, vo_b, bp[0]);
                    buf_st(Bpout + ((int64_t)N + pk), vo_b, bp[1]);
                    buf_st(Bpout + (2 * (int64_t)N + pk), vo_b, bp[2]);
                }
            }
        }
        const int wb = (kk - k0 + 1) % 3;
        lds[wb][0][s] = gnew[1];
        lds[wb][1][s] = gnew[2];
        lds[wb][2][s] = gnew[3];
        lds[wb][3][s] = gnew[4];
        __syncthreads();
        if constexpr (REGS) q_have = kk < k1 && load(kk + 1, qg, qf, qb, qmodes);
        if constexpr (PF == 2) q_have = kk < k1 && stage_load(kk + 1, qmodes, qb);

        const int p = kk - 1;
        if (tile && p >= k0 && p < k1 && have_prev) {
            const int rb = (p - k0 + 1) % 3;
            // T^{n+1}(i, j, p): sum of the pulled post-wall g_post^n in direction order 0..6
            R gin[7];
            gin[0] = gprev[0];
            gin[1] = i > 0 ? Form::pulled(1, lds[rb][0][s - 1], fp)
                           : Form::wall_g(1, gprev, kp.temp_x, fp, nd);
            gin[2] = i < nx - 1 ? Form::pulled(2, lds[rb][1][s + 1], fp)
                                : Form::wall_g(2, gprev, kp.temp_x, fp, nd);
            gin[3] = j > 0 ? Form::pulled(3, lds[rb][2][s - TX], fp)
                           : Form::wall_g(3, gprev, kp.temp_y, fp, nd);
            gin[4] = j < ny - 1 ? Form::pulled(4, lds[rb][3][s + TX], fp)
                                : Form::wall_g(4, gprev, kp.temp_y, fp, nd);
            gin[5] = (p > 0 && have_pp) ? Form::pulled(5, g5pp, fp)
                                        : Form::wall_g(5, gprev, kp.temp_z, fp, nd);
            gin[6] = p < nz - 1 ? Form::pulled(6, gnew[6], fp)
                                : Form::wall_g(6, gprev, kp.temp_z, fp, nd);
            const R T1 = Form::tsum(gin);
            R out[19];
            Form::flow(fprev, out, rho_p, u_p, v_p, w_p, T1, kp, fp, nd);
            if constexpr (PF == 2) {
#pragma unroll
                for (int a = 0; a < 19; ++a) fout[a] = out[a];
                fpend = true;
                fpend_p = p;
            } else {
                store_f_at(p, out);
            }
        }
        // shift the plane pipeline
        g5pp = gprev[5];
        have_pp = have_prev;
#pragma unroll
        for (int a = 0; a < 7; ++a) gprev[a] = gnew[a];
#pragma unroll
        for (int a = 0; a < 19; ++a) fprev[a] = fnew[a];
        rho_p = rho_n;
        u_p = u_n;
        v_p = v_n;
        w_p = w_n;
        have_prev = have;
    }
    if constexpr (PF == 2)
        if (fpend) store_f_at(fpend_p, fout);
}

// The launch's domains and their parameters come from a table in device memory: block b (after
// the XCD remap) belongs to domain owner[b] (domain 0 when owner is null), whose first block is
// inst[d].first_block.  Reading the parameters through a pointer, rather than as kernel
// arguments, lets the compiler reload them with scalar loads instead of holding them in vector
// registers (FP64: about 200 instead of 256 registers, no spills).
template <class Form> struct FusedInstance {
    KParams kp;
    typename Form::P fp;
    FusedGeom fg;
    const typename Form::S *Fcur;
    typename Form::S *Fnext;
    const typename Form::S *Gin;
    typename Form::S *Gout;
    const typename Form::B *Bp;
    typename Form::B *Bpout;
    unsigned first_block;
};

// At least two waves per SIMD: with 256 registers per lane at most, a 512-thread tile fits one
// CU and 256-thread tiles two per CU.  Staged (PF == 2): four waves per SIMD, two 34x15 blocks
// per CU (staging slots and exchange buffers: 78 KB each) or one 66x15 block (154 KB).
template <class Form, int TX, int TY, int PF>
__global__ void __launch_bounds__(TX *TY, PF == 2 ? 4 : 2)
k3_fused(const FusedInstance<Form> *__restrict__ inst, const uint16_t *__restrict__ owner,
         unsigned nblocks)
{
    const unsigned b = xcd_remap(blockIdx.x, nblocks);
    const FusedInstance<Form> &d = inst[owner ? owner[b] : 0];
    fused_tile<Form, TX, TY, PF>(d.kp, d.fp, d.fg, b - d.first_block, gptr(d.Fcur), gptr(d.Fnext),
                                 gptr(d.Gin), gptr(d.Gout), gptr(d.Bp), gptr(d.Bpout));
}

// Run-time tile shapes: index -> (TX, TY).
constexpr int fused_tx(int tile) { return tile == 1 ? 16 : (tile == 4 ? 34 : (tile == 5 ? 66 : 32)); }
constexpr int fused_ty(int tile)
{
    return tile == 0 ? 16 : (tile == 1 ? 16 : (tile == 2 ? 8 : (tile == 3 ? 32 : 15)));
}

// One launch over `nblocks` blocks of the domains in table `inst` (owner null: one domain).
template <class Form>
hipError_t launch_fused_table(const FusedGeom &fg0, const void *inst, const uint16_t *owner,
                              unsigned nblocks, hipStream_t s)
{
    const FusedInstance<Form> *d = (const FusedInstance<Form> *)inst;
#define RBLBM_FUSED_B(T, PF)                                                                     \
    k3_fused<Form, fused_tx(T), fused_ty(T), PF><<<nblocks, fused_tx(T) * fused_ty(T), 0, s>>>( \
        d, owner, nblocks)
    constexpr bool lp = sizeof(typename Form::R) == 4;
    if (fg0.pf == 2) {
        // staged loads: 4-byte storage, the 34x15 and 66x15 tiles
        if constexpr (Form::can_stage) {
            if (fg0.tile == 4) RBLBM_FUSED_B(4, 2);
            else if (fg0.tile == 5) RBLBM_FUSED_B(5, 2);
            else return hipErrorInvalidValue;
            return hipGetLastError();
        } else {
            return hipErrorInvalidValue;
        }
    }
    switch (fg0.tile * 2 + (fg0.pf ? 1 : 0)) {
    case 0: RBLBM_FUSED_B(0, 0); break;
    case 1: RBLBM_FUSED_B(0, 1); break;
    case 2: RBLBM_FUSED_B(1, 0); break;
    case 3: RBLBM_FUSED_B(1, 1); break;
    case 4: RBLBM_FUSED_B(2, 0); break;
    case 5: RBLBM_FUSED_B(2, 1); break;
    case 8: RBLBM_FUSED_B(4, 0); break;
    case 9: RBLBM_FUSED_B(4, 1); break;
    default:
        // 32 x 32 = 1024 and 66 x 15 = 990 threads leave 128 registers per lane:
        // reduced-precision forms only
        if constexpr (lp) {
            switch (fg0.tile * 2 + (fg0.pf ? 1 : 0)) {
            case 6: RBLBM_FUSED_B(3, 0); break;
            case 7: RBLBM_FUSED_B(3, 1); break;
            case 10: RBLBM_FUSED_B(5, 0); break;
            case 11: RBLBM_FUSED_B(5, 1); break;
            default: return hipErrorInvalidValue;
            }
        } else {
            return hipErrorInvalidValue;
        }
        break;
    }
#undef RBLBM_FUSED_B
    return hipGetLastError();
}


} // namespace rblbm
