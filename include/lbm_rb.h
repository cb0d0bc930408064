/*
 * lbm_rb.h -- C ABI of the MI355X (gfx950) thermal MRT lattice-Boltzmann Rayleigh-Benard solver.
 *
 * The reference (19134984/-2D-3D-, Fortran 90) has no API: its hot path is the body of the time
 * loop in 2DRB/3DRB.F90:225-264 (eight argument-less subroutines over module globals) plus the
 * end-of-run diagnostics at 2DRB/3DRB.F90:289-297, and likewise 2DRB/2DRB.F90:236-304 in 2D.
 * Each entry point below names the reference code it replaces.  A Fortran driver binds these
 * through iso_c_binding (-2d-3d-_amd/fortran/lbm_rb.f90); Python binds them through ctypes.
 *
 * Conventions
 *   - Every function returns LBM_OK (0) or a negative LBM_ERR_* code; nothing throws across
 *     the ABI.  lbm_last_error() gives a message for the last failure on a handle.
 *   - Host arrays use the reference's Fortran order: macros (nx,ny,nz) with x fastest;
 *     populations f(0:q-1,nx,ny,nz) with the population index fastest, in the reference's
 *     direction order (2DRB/3DRB.F90:656-658, 2DRB/2DRB.F90:188-190).
 *   - One handle owns one GPU domain and one HIP stream; lbm_step is asynchronous on it.
 *     A handle is not re-entrant; distinct handles may be used from distinct threads.
 */
#ifndef LBM_RB_H
#define LBM_RB_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_ERR_ARG (-1)          /* invalid argument or configuration */
#define LBM_ERR_HIP (-2)          /* a HIP runtime call failed */
#define LBM_ERR_NOMEM (-3)        /* device or host allocation failed */
#define LBM_ERR_STATE (-4)        /* call out of order (e.g. step before init) */
#define LBM_ERR_UNSUPPORTED (-5)  /* configuration not built (e.g. legacy thermal branch) */
#define LBM_ERR_RHO (-6)          /* non-positive density during restart reconstruction */

/* Wall conditions: the reference's cpp switches (2DRB/3DRB.F90:13-37, 2DRB/2DRB.F90:14-30). */
#define LBM_VEL_NOSLIP 0          /* *WallsNoslip: half-way bounce-back */
#define LBM_VEL_PERIODIC 1        /* *WallsPeriodicalU: same-row periodic copy */
#define LBM_T_CONST 0             /* *WallsConstT: anti-bounce-back, hot at the low index */
#define LBM_T_ADIABATIC 1         /* *WallsAdiabatic: bounce-back */
#define LBM_T_PERIODIC 2          /* *WallsPeriodicalT */

/* Population storage forms.  FP16-shift rounds the stored deviations to 11 bits; with the flow
 * relaxation time tau_f within about 0.004 of 1/2 (high Ra on a coarse grid) that rounding can
 * outgrow the non-equilibrium part of f once the convection is strong: a 256^3 Ra=1e8 cavity
 * (tau_f = 0.5037) diverged there after 67,000 steps in FP16-shift but not in FP32-shift. */
#define LBM_FP64 0                /* plain FP64 populations and arithmetic */
#define LBM_FP32_SHIFT 1          /* FP32 deviations f - w rho0, g - wT T0; FP32 arithmetic */
#define LBM_FP16_SHIFT 2          /* FP16 scaled deviations; FP32 arithmetic */

typedef struct lbm_config {
    int32_t dims;                 /* 2 (D2Q9 + D2Q5) or 3 (D3Q19 + D3Q7) */
    int32_t nx, ny, nz;           /* interior nodes (nz ignored in 2D) */
    double rayleigh, prandtl, mach;
    double thot, tcold;
    int32_t side_heated;          /* 0: RayleighBenardCell (L = ny); 1: SideHeatedCell (L = nx) */
    int32_t vel_x, vel_z;         /* LBM_VEL_* for x (vertical) and z (spanwise) walls; y is no-slip */
    int32_t temp_x, temp_y, temp_z; /* LBM_T_* */
    int32_t use_g;                /* EnableUseG: M1G heat-flux correction */
    int32_t precision;            /* LBM_FP64 | LBM_FP32_SHIFT | LBM_FP16_SHIFT */
    int32_t device;               /* HIP device ordinal */
    int32_t legacy_thermal;       /* EnableLegacyThermalScheme (FP64 form only): fixed Qk, Qnu,
                                     rest weight from paraA (2DRB/3DRB.F90:93, :116-118) */
    int32_t layout;               /* LBM_LAYOUT_*: population arrays of the step (no reference
                                     counterpart; results do not depend on it) */
} lbm_config;

/* Population layouts.  TWO_ARRAY: two arrays per lattice, every kernel form.  AA: one array per
 * lattice updated in place (half the population memory; 3D with a wall on every face, two-pass
 * form only).  AUTO: TWO_ARRAY when it fits in free device memory, else AA where allowed;
 * lbm_autotune may switch an AUTO domain that has not stepped yet to the faster of the two. */
#define LBM_LAYOUT_TWO_ARRAY 0
#define LBM_LAYOUT_AA 1
#define LBM_LAYOUT_AUTO 2

/* Derived parameters: module commondata3d, 2DRB/3DRB.F90:73-123 (2D: 2DRB/2DRB.F90:61-115). */
typedef struct lbm_params {
    double L, tref, tauf, viscosity, diffusivity, gbeta, time_unit, velocity_unit;
    double s_e, s_eps, s_nu, s_pi, s_q, s_m;
    double taug, q_k, q_nu, sg;
    double para_a;                /* legacy thermal scheme's paraA (0 otherwise) */
} lbm_params;

/* Diagnostics, FP64 with compensated sums. */
typedef struct lbm_diag {
    double nu_vol, re_vol;        /* calNuRe, 2DRB/3DRB.F90:1712-1778 */
    double nu_global;             /* RBcalc_Nu_global, 2DRB/3DRB.F90:1839-1878 */
    double nu_hot, nu_cold, nu_middle;       /* RBcalc_Nu_wall_avg, 2DRB/3DRB.F90:2049-2201 */
    double nu_hot_max, nu_hot_max_pos, nu_hot_min, nu_hot_min_pos;
    double mass, heat, kinetic;   /* sum rho, sum T, 0.5 sum rho|u|^2 (conservation checks) */
} lbm_diag;

/* Mid-plane velocity maxima, scaled by velocityScaleCompare = L / kappa as the reference
 * reports them.  Per component: the maximum, its position (3D: the two in-plane coordinates;
 * 2D: one, pos2 = 0) and the plane's coordinate, all divided by L. */
typedef struct lbm_midplane {
    double u_max, u_pos1, u_pos2, u_plane;   /* u on the x mid-plane (3D: pos = y, z) */
    double v_max, v_pos1, v_pos2, v_plane;   /* v on the y mid-plane (3D: pos = x, z) */
    double w_max, w_pos1, w_pos2, w_plane;   /* w on the z mid-plane (3D: pos = x, y; 2D: 0) */
} lbm_midplane;

/* 2D stream-function statistics (calc_psi_vort_and_output, 2DRB/2DRB.F90:2657-2770):
 * |psi| at the cavity centre from 4x4 Lagrange interpolation (output_psi_center_abs) and from
 * natural cubic splines, and the largest |psi| on the fine spline grid (10000 intervals per unit
 * length, calc_psi_absmax_fine_spline) with its position. */
typedef struct lbm_psi {
    double psi_center_coarse, psi_center_fine;
    double psi_abs_max, x_at_max, y_at_max;
} lbm_psi;

typedef struct lbm_handle lbm_handle;

/* Allocate the domain on cfg->device.  Replaces the allocate block of `initial`
 * (2DRB/3DRB.F90:467-473) and the parameter module. */
int lbm_create(const lbm_config *cfg, lbm_handle **out);
int lbm_destroy(lbm_handle *h);

/* Exact-initialisation branch of `initial` (2DRB/3DRB.F90:517-600; 2D 2DRB/2DRB.F90:540-608):
 * rho = 1, u = 0, f = w, g = wT T, B_prev = u T.  T0 (Fortran order, nx*ny*nz) replaces the
 * reference temperature profile when non-NULL. */
int lbm_init_exact(lbm_handle *h, const double *T0);
/* Exact initialisation plus seeded uniform noise amp*(thot-tcold)*[-1,1) on every node
 * (counter-based SplitMix64 of seed and linear node index; no reference counterpart). */
int lbm_init_seeded(lbm_handle *h, uint64_t seed, double amp);

/* Advance nsteps time steps: each is one pass of collision, streaming, bounceback, macro,
 * collisionT, streamingT, bouncebackT, macroT (2DRB/3DRB.F90:227-243). Asynchronous.  On
 * domains of up to 2^22 nodes, 128 or more steps replay a HIP graph of 32 steps. */
int lbm_step(lbm_handle *h, int32_t nsteps);
/* Same, bracketed by HIP events on the handle's stream; *ms is the elapsed device time. */
int lbm_step_timed(lbm_handle *h, int32_t nsteps, double *ms);
int lbm_sync(lbm_handle *h);
int lbm_itc(lbm_handle *h, int64_t *itc);
int lbm_get_params(lbm_handle *h, lbm_params *p);

/* rho, u, v, w, T after the last step (module arrays of 2DRB/3DRB.F90:155-157); any pointer
 * may be NULL.  Synchronous.  In 2D, w is filled with zeros. */
int lbm_macros(lbm_handle *h, double *rho, double *u, double *v, double *w, double *T);

/* Steady-state convergence check, `check` (2DRB/3DRB.F90:1475-1528; 2D 2DRB/2DRB.F90:1213-1260):
 * relative L2 change of u and relative L1 change of T since the previous call. */
int lbm_check(lbm_handle *h, double *error_u, double *error_t);

/* Nusselt / Reynolds / conservation diagnostics of the current state. */
int lbm_diagnostics(lbm_handle *h, lbm_diag *d);
/* Runs nsteps steps with the diagnostics every `every` steps (nsteps a multiple of every) into
 * out[nsteps / every] (no reference counterpart: the time series that calNuRe appends to,
 * 2DRB/3DRB.F90:1712-1778, at the reference's check cadence or any other).  The first period
 * runs eagerly; then one or two periods of steps and the diagnostics' device reductions are
 * captured in a HIP graph and replayed, each replay appending its partial sums to a device
 * series; the host finishes the compensated sums after the run.  Results equal lbm_step +
 * lbm_diagnostics at the same steps. */
int lbm_run_diagnostics(lbm_handle *h, int32_t nsteps, int32_t every, lbm_diag *out);
/* lbm_run_diagnostics that allocates no device memory (no counterpart in the reference; for
 * callers that cannot allocate: serving loops, harnesses that forbid hipMalloc).  The series
 * lives in a 32 MiB device scratch array of the library (one per GPU, held until the process
 * ends); every period runs eagerly (no graph is instantiated).  Returns LBM_ERR_ARG, with a
 * message naming the limit, when (nsteps / every) samples of this domain do not fit the scratch.
 * Results equal lbm_run_diagnostics.  All calls on one GPU share its scratch and the last call's
 * contents stay there: calls that overlap on one GPU (other threads, other handles' streams)
 * race. */
int lbm_run_diagnostics_noalloc(lbm_handle *h, int32_t nsteps, int32_t every, lbm_diag *out);
/* Mid-plane velocity maxima of the current state.  3D: calc_umid_max_common,
 * calc_vmid_max_common, calc_wmid_max_common (2DRB/3DRB.F90:2399-2518; the field interpolated
 * onto the plane at half the domain extent, the first grid maximum).  2D: RBcalc_umid_max /
 * RBcalc_vmid_max (2DRB/2DRB.F90:2484-2642) or, for the side-heated cell,
 * SideHeatedcalc_umid_max / SideHeatedcalc_vmid_max (:2081-2238): a 5-point least-squares
 * parabola around the grid maximum of the sampled line.  Needs nx, ny >= 5. */
int lbm_midplane_extrema(lbm_handle *h, lbm_midplane *m);
/* 3D mid-plane slice across `axis` (0: x, 1: y, 2: z) at half the domain extent:
 * u, v, w scaled by L / kappa and T, interpolated onto the plane (write_midplane_x/y/z,
 * 2DRB/3DRB.F90:2611-2713), and the slice's stream function and plane-normal vorticity
 * (write_midplane_stream_x/y/z, :2721-3011: psi integrates the first in-plane axis's velocity
 * along the second axis from the wall).  Each output holds n1 * n2 values, first in-plane axis
 * fastest: x-plane (y, z), y-plane (x, z), z-plane (x, y).  Any output may be NULL. */
int lbm_midplane_slice(lbm_handle *h, int32_t axis, double *u, double *v, double *w, double *T,
                       double *psi, double *vort);
/* 2D stream function and vorticity (nx*ny each, x fastest; psi integrates u along y from the
 * bottom wall, the vorticity takes one-sided differences through the resting walls) and, if st
 * is not NULL, the |psi| statistics; the fine-grid search runs on the GPU.  psi and vort may be
 * NULL.  2DRB/2DRB.F90:2657-3160. */
int lbm_psi_vorticity_2d(lbm_handle *h, double *psi, double *vort, lbm_psi *st);

/* f, g (reference AoS order, absolute FP64) and, if non-NULL, B_prev (3 planes of nx*ny*nz:
 * Bx_prev, By_prev, Bz_prev, 2DRB/3DRB.F90:163).  Restart data of backupData
 * (2DRB/3DRB.F90:1638-1665). */
int lbm_gather(lbm_handle *h, double *f, double *g, double *bprev);
/* Load f, g and rebuild the macros like reconstruct_macro_from_fg (2DRB/3DRB.F90:1392-1463,
 * 2DRB/2DRB.F90:1144-1200).  Returns LBM_ERR_RHO if any rho <= 0. */
int lbm_scatter(lbm_handle *h, const double *f, const double *g);

/* Bytes per cell update that the step moves at algorithmic minimum (both lattices read and
 * written once, plus the M1G history): 464 (FP64 3D), 232 (FP32), 128 (FP16), 256 (FP64 2D). */
int lbm_bytes_per_update(lbm_handle *h, double *bytes);

/* Kernel form of lbm_step (no reference counterpart: the reference runs eight sweeps).
 * TWO_PASS: a thermal and a flow launch per step.  ONE_PASS: one launch per step (3D without
 * periodic thermal walls only); it moves fewer bytes but recomputes a ring of nodes per tile.
 * AUTO (the default) runs ONE_PASS for FP64 domains of 2^22 nodes or more and FP32-shift
 * domains of 2^24 nodes or more (the faster form there on MI355X) and TWO_PASS otherwise
 * (FP16-shift: TWO_PASS at every size); lbm_autotune measures both per shape.  Both give
 * bitwise identical results in every precision form.  ONE_PASS needs the two-array layout.
 * kchunk > 0 sets the number of z-planes per block of the one-pass form, LBM_KCHUNK_AUTO the
 * default (the largest power of two up to nz / 16 within 8..32, nz below 8), 0 keeps the
 * current value. */
#define LBM_KERNEL_AUTO 0
#define LBM_KERNEL_TWO_PASS 1
#define LBM_KERNEL_ONE_PASS 2
#define LBM_KCHUNK_AUTO (-1)
int lbm_set_kernel_variant(lbm_handle *h, int32_t variant, int32_t kchunk);

/* Thread tile of the one-pass form (no reference counterpart): 32x16 threads, 16x16 or 32x8
 * (256-thread blocks, more of them per CU), 32x32 (1024 threads, the smallest ring share),
 * 34x15 (32 x 13 nodes, rows aligned to 32 nodes) or 66x15 (64 x 13 nodes; the x-edges of a
 * tile fetch half as many extra cache lines per row).  32x32 and 66x15: FP32 / FP16 forms only,
 * else LBM_ERR_UNSUPPORTED.  prefetch = LBM_PREFETCH_REGS issues the next plane's loads into
 * registers before the current flow collision; LBM_PREFETCH_STAGED loads them straight into
 * LDS (FP32-shift form with the 34x15 or 66x15 tile only, else LBM_ERR_UNSUPPORTED).  Defaults:
 * 34x15 with LBM_PREFETCH_REGS (FP64, FP16-shift), 66x15 with LBM_PREFETCH_NONE (FP32-shift).
 * Results do not depend on the choice.
 * Batch domains reject this and lbm_set_kernel_variant (LBM_ERR_STATE): a batch is tuned with
 * lbm_batch_autotune or lbm_batch_set_kernel_variant. */
#define LBM_TILE_32X16 0
#define LBM_TILE_16X16 1
#define LBM_TILE_32X8 2
#define LBM_TILE_32X32 3
#define LBM_TILE_34X15 4
#define LBM_TILE_66X15 5
#define LBM_PREFETCH_NONE 0
#define LBM_PREFETCH_REGS 1
#define LBM_PREFETCH_STAGED 2
int lbm_set_fused_tile(lbm_handle *h, int32_t tile, int32_t prefetch);
/* The population layout in effect (LBM_LAYOUT_TWO_ARRAY or LBM_LAYOUT_AA). */
int lbm_get_layout(lbm_handle *h, int32_t *layout);
/* The kernel form steps run (after the first step from an initial state, which is two-pass):
 * variant LBM_KERNEL_TWO_PASS or LBM_KERNEL_ONE_PASS (AUTO resolved), and the one-pass form's
 * z-chunk, tile and prefetch. */
int lbm_get_kernel_form(lbm_handle *h, int32_t *variant, int32_t *kchunk, int32_t *tile,
                        int32_t *prefetch);

/* Autotuner (no reference counterpart).  Picks the kernel form, the one-pass z-chunk (nodes
 * marched per thread), tile shape and prefetch for this handle's shape by timing `steps` steps
 * of each candidate on a scratch domain of the same configuration (with fewer z-planes when the
 * device cannot hold a second copy), applies the winner and appends it to the text cache
 * `cache_path` (NULL: no cache) as "key variant kchunk tile prefetch ms".  A shape already in
 * the cache is applied without timing.  The handle's own state is not advanced.  *ms_per_step
 * (may be NULL) receives the winner's time per step (0 when it came from the cache). */
int lbm_autotune(lbm_handle *h, const char *cache_path, int32_t steps, double *ms_per_step);

/* ---- Batched ensembles (no reference counterpart: the reference runs one cavity per
 * process).  A batch owns n independent domains of any shape and parameters; they share one
 * HIP stream and advance together.  Domain i is an ordinary handle for the per-domain calls that
 * read it (diagnostics, check, macros, gather, mid-plane and slice output); its seeded noise uses
 * i as the instance key.  The calls that change a domain's state on their own (lbm_step,
 * lbm_step_timed, lbm_init_exact, lbm_init_seeded, lbm_scatter, lbm_set_kernel_variant,
 * lbm_autotune, lbm_run_diagnostics, lbm_destroy) return LBM_ERR_STATE for a batch domain: the
 * batch's launch tables describe its buffers, and only the lbm_batch_* calls keep them current.
 * The one-launch modes run the one-pass form (one launch a step; the tile shape and prefetch of
 * domain 0 for all, each domain's own z-chunk) when every domain admits it, none is set to
 * LBM_KERNEL_TWO_PASS, and either every domain is set to LBM_KERNEL_ONE_PASS
 * (lbm_batch_set_kernel_variant or lbm_batch_autotune) or, under AUTO, the batch's total node
 * count reaches the size at which AUTO runs a single domain in the one-pass form (2^22 FP64,
 * 2^24 FP32-shift; never FP16-shift); otherwise the two-pass form (two launches a step over
 * 256-node blocks of all domains).  One launch per domain runs each domain's own form.  In the
 * one-pass table a domain not set to LBM_KERNEL_ONE_PASS takes z-chunks of min(nz, 32) planes
 * (the table launch fills the device, so fewer halo planes win); one set to it keeps its own. */
typedef struct lbm_batch lbm_batch;
#define LBM_BATCH_PER_DOMAIN 0    /* one launch per domain per step */
#define LBM_BATCH_ONE_LAUNCH 1    /* one launch per step for all domains (per-domain table) */
#define LBM_BATCH_GRAPH 2         /* ONE_LAUNCH, two steps captured in a HIP graph and replayed */
int lbm_batch_create(const lbm_config *cfgs, int32_t n, lbm_batch **out);
int lbm_batch_destroy(lbm_batch *b);
int lbm_batch_domain(lbm_batch *b, int32_t i, lbm_handle **h);
int lbm_batch_init_seeded(lbm_batch *b, uint64_t seed, double amp);
int lbm_batch_set_mode(lbm_batch *b, int32_t mode);
/* lbm_set_kernel_variant for every domain of the batch (LBM_KERNEL_*; z-chunks kept). */
int lbm_batch_set_kernel_variant(lbm_batch *b, int32_t variant);
int lbm_batch_step(lbm_batch *b, int32_t nsteps);
int lbm_batch_step_timed(lbm_batch *b, int32_t nsteps, double *ms);
int lbm_batch_nodes(lbm_batch *b, int64_t *nodes);
/* lbm_run_diagnostics for a batch: out[k * n + i] is sample k of domain i.  The captured
 * period holds the batch's steps in its mode (table launches, or per-domain launches in
 * LBM_BATCH_PER_DOMAIN) and every domain's diagnostics reductions. */
int lbm_batch_run_diagnostics(lbm_batch *b, int32_t nsteps, int32_t every, lbm_diag *out);
/* lbm_autotune for each distinct domain shape (the one-launch modes use the tuned z-chunks),
 * then, unless every domain already runs the one-pass form, the batch's one-launch step with
 * those choices is timed against every domain in the one-pass form on a scratch batch of the
 * same configurations (when free device memory holds it): a small cavity alone cannot fill the
 * device with one-pass tiles, the whole batch can.  The form AUTO would run for the batch is
 * kept unless the other is more than 3% faster.  The per-shape choices go to the cache; the
 * batch-level choice does not. */
int lbm_batch_autotune(lbm_batch *b, const char *cache_path, int32_t steps);

const char *lbm_last_error(lbm_handle *h);
const char *lbm_error_string(int code);

#ifdef __cplusplus
}
#endif
#endif
