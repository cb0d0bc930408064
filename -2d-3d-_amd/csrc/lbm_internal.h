// lbm_internal.h -- host-side declarations shared by the C-ABI layer and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include "lbm_lattice.h"

namespace rblbm {

struct InitParams {
    int use_T0;           // T already in mac[4N..] (host-supplied profile)
    int rb_perturb;       // 0 none, 3 = 3D cos*sin (Ra <= 1e4), 2 = 2D sin*sin (Ra < 1e4)
    double L, pi, perturb_amp;
    uint64_t seed;
    uint64_t instance;    // domain index within a batch (part of the noise key)
    double noise_scale;   // amp * (thot - tcold); 0 disables the noise
};

// aa: 3D in-place layout (Fin == Fout; the plain state LAY_ODD becomes LAY_EVEN).
hipError_t launch_prologue(const KParams &kp, int bx, hipStream_t s, const double *Fin,
                           double *Fout, const double *mac, double *fy_out, bool three_d,
                           bool aa = false);
// In-place (A-A) 3D passes, each a separate launch: pass B (flow collision; F and G in state
// lay, F updated) and pass A (thermal collision; G in state lay and updated, F read in the
// other state).  first: pass A of the first step (plain g, T from mac; lay = LAY_ODD).
hipError_t launch_flow3_aa(const KParams &kp, int bx, hipStream_t s, int lay, double *F,
                           const double *G, const double *aux);
hipError_t launch_thermal3_aa(const KParams &kp, int bx, hipStream_t s, int lay, bool first,
                              const double *F, double *G, double *Bp, double *aux,
                              const double *mac);
hipError_t launch_step3(const KParams &kp, int bx, hipStream_t s, bool first, const double *Fcur,
                        double *Fnext, const double *Gin, double *Gout, double *Bp, double *aux,
                        const double *mac);
hipError_t launch_step2(const KParams &kp, int bx, hipStream_t s, bool first, const double *Fcur,
                        double *Fnext, const double *Gin, double *Gout, double *Bp,
                        const double *fy_cur, double *fy_next, const double *mac);
// Reference state from the pipelined buffers; lay: layout of Fprev / Gpost (3D: LAY_PULL, or
// an in-place state with F and G in the same state).
hipError_t launch_state(const KParams &kp, int bx, hipStream_t s, bool three_d, const double *Fprev,
                        const double *Gpost, const double *aux_or_fy, const double *Gprev,
                        double *mac, double *fo, double *go, int k0, int nk, int lay = 0);

// One-pass form: tile of (TX - 2) x (TY - 2) nodes plus a one-node ring per block; the thread
// tile TX x TY is one of five shapes (FusedGeom::tile): 32x16, 16x16, 32x8, 32x32 (the
// reduced-precision forms only) and 34x15 (32 x 13 nodes: rows of 32 nodes start on a 256-byte
// boundary, and a 512-wide domain splits into 16 tiles across).
constexpr int FUSED_TILES = 6;
inline void fused_tile_dims(int tile, int *tx, int *ty)
{
    *tx = tile == 1 ? 16 : (tile == 4 ? 34 : (tile == 5 ? 66 : 32));
    *ty = tile == 0 ? 16 : (tile == 1 ? 16 : (tile == 2 ? 8 : (tile == 3 ? 32 : 15)));
}
hipError_t launch_init(const KParams &kp, int bx, hipStream_t s, const InitParams &ip, double *F,
                       double *G, double *Bp, double *mac, int qf, int qt);
hipError_t launch_transpose(bool to_soa, hipStream_t s, const double *in, double *out, int q,
                            uint64_t N, uint64_t c0, uint64_t n);
hipError_t launch_reconstruct(const KParams &kp, hipStream_t s, const double *F, const double *G,
                              double *Bp, double *mac, int qf, int qt, int *bad);
// side = 1 selects the side-heated cell's formulas (hot wall x = 0; wall profile along y).
hipError_t launch_reduce_volume(const KParams &kp, hipStream_t s, const double *mac,
                                double L_over_kappa, double dy, int side, double *out,
                                unsigned blocks);
hipError_t launch_reduce_walls(const KParams &kp, hipStream_t s, const double *mac,
                               double L_over_kappa, double dy, int side, double *tbot,
                               double *nubot, double *out);
// Plane of a macro field across `axis` at 0-based indices (l, r), see k_plane.
hipError_t launch_plane(const KParams &kp, hipStream_t s, const double *f, int axis, int l, int r,
                        double w, int mode, double *out);
// Fine-grid |psi| maximum (2D): row splines rows / y2rows (ny x (nx + 2)) sampled at nfx fine
// x positions into the fine columns, their splines, then |psi| at nfx x nfy points; per block the
// maximum and the first linear index (column-major fine grid) where it occurs.
hipError_t launch_psi_absmax(hipStream_t s, int nx, int ny, int nfx, int nfy, double xlen, double ylen,
                             const double *xe, const double *ye, const double *rows,
                             const double *y2rows, double *cols, double *y2cols, double *wk,
                             double *bval, unsigned long long *bidx, unsigned blocks);
hipError_t launch_reduce_check(const KParams &kp, hipStream_t s, const double *mac, double *prev,
                               double *out, unsigned blocks);
// Diagnostics series: append n1 doubles of a and n2 of b at slot *count of `series` (slots of
// `stride` doubles, `cap` of them; nothing is written past the last) and advance *count.  The
// slot is read on the device, so a captured graph appends to the next slot at each replay.
hipError_t launch_series_append(hipStream_t s, const double *a, size_t n1, const double *b,
                                size_t n2, double *series, size_t stride, int cap, int *count);
hipError_t launch_series_reset(hipStream_t s, int *count, int n);

// Reduced-precision (FP32 / FP16 deviation storage) 3D kernels, lbm_kernels_lp.hip.
enum {
    LP_PROLOGUE,      // flow collision of step 1: Fc (plain) -> Fn, macros from mac
    LP_STEP,          // passes A and B: Fc, Gi -> Go, Fn; Bp, aux updated
    LP_FLOW_AA,       // in place, state lay: flow collision, Fn updated, Gi read
    LP_THERMAL_AA,    // in place, state lay: thermal collision, Go updated, Fc read (first: plain g)
    LP_STATE,         // reference state from Fc (= f_post^n), Gi (= g_post^n), aux -> mac, fo/go
    LP_INIT,          // plain state from mac: Fn = 0, Go = wT (T - T0), Bp = u T
    LP_FROM_ABS,      // plain state from an absolute AoS chunk fo/go [c0, c0 + n) -> Fn, Go
    LP_TO_ABS,        // plain state Fc, Gi [c0, c0 + n) -> absolute AoS chunk fo/go
    LP_BPREV,         // Bp (3 float planes) -> fo (FP64), n values
    LP_RECON_AOS,     // reconstruct_macro_from_fg on the AoS chunk fo/go -> mac, *flag on rho <= 0
    LP_BP_FROM_MAC,   // Bp = u T from mac (zero without use_g)
};
struct LPArgs {
    const void *Fc = nullptr;
    void *Fn = nullptr;
    const void *Gi = nullptr;
    void *Go = nullptr;
    float *Bp = nullptr, *aux = nullptr;
    double *mac = nullptr, *fo = nullptr, *go = nullptr;
    int first = 0, k0 = 0, nk = 0;
    uint64_t c0 = 0, n = 0;
    int *flag = nullptr;
    const void *Gprev = nullptr;   // LP_STATE: g_post^{n-1} when aux is not valid
    int lay = 0;                   // population layout (LAY_PULL, or an in-place state)
};
hipError_t launch_lp(int precision, int op, const KParams &kp, int bx, hipStream_t s, const LPArgs &a);

// One-pass launches read their domains from a table in device memory (lbm_fused.h).  The
// per-domain entry is form-specific; these helpers size and fill it on the host, then launch
// over `nblocks` blocks whose domain is `owner[block]` (owner null: the table's one domain).
struct FusedPtrs {
    const void *Fcur;
    void *Fnext;
    const void *Gin;
    void *Gout;
    const void *Bp;
    void *Bpout;
};
size_t fused_instance_bytes_f64();
void fused_instance_fill_f64(void *dst, const KParams &kp, const FusedGeom &fg, const FusedPtrs &p,
                             unsigned first_block);
hipError_t launch_fused_batch_f64(const FusedGeom &fg0, const void *inst, const uint16_t *owner,
                                  unsigned nblocks, hipStream_t s);
// Batched two-pass step (3D): per-domain entries of 256-node blocks, pass A then pass B.
// p.Bp is the history updated in place; aux holds T^n between the passes.
size_t twopass_entry_bytes_f64();
void twopass_entry_fill_f64(void *dst, const KParams &kp, const FusedPtrs &p, double *aux,
                            unsigned first_block);
hipError_t launch_twopass_batch_f64(const void *tab, const uint16_t *owner, unsigned nblocks,
                                    hipStream_t s);
size_t twopass_entry_bytes_lp(int precision);
void twopass_entry_fill_lp(int precision, void *dst, const KParams &kp, const FusedPtrs &p,
                           float *aux, unsigned first_block);
hipError_t launch_twopass_batch_lp(int precision, const void *tab, const uint16_t *owner,
                                   unsigned nblocks, hipStream_t s);
size_t fused_instance_bytes_lp(int precision);
void fused_instance_fill_lp(int precision, void *dst, const KParams &kp, const FusedGeom &fg,
                            const FusedPtrs &p, unsigned first_block);
hipError_t launch_fused_batch_lp(int precision, const FusedGeom &fg0, const void *inst,
                                 const uint16_t *owner, unsigned nblocks, hipStream_t s);

} // namespace rblbm
