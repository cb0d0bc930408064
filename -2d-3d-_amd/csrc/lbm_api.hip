// lbm_api.hip -- the C ABI of include/lbm_rb.h: handle life cycle, stepping, diagnostics,
// gather/scatter.  Host code only; the kernels live in lbm_kernels.hip.
#include "lbm_internal.h"
#include "../../include/lbm_rb.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <mutex>
#include <utility>
#include <vector>

using namespace rblbm;

struct lbm_handle {
    lbm_config cfg;
    lbm_params p;
    KParams kp;
    int bx;                 // threads per block of the step kernels
    int qf, qt;
    uint64_t N;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    int es;                 // bytes per stored population (8 / 4 / 2)
    bool lp;                // reduced-precision deviation storage (3D only)
    // F, G, Bp, aux are double in the FP64 form; F, G are float / _Float16 and Bp, aux float
    // in the reduced-precision forms.
    // Bp is the current M1G history; Bp_alt receives the one-pass form's update (ring nodes of a
    // tile read B_prev of nodes that a neighbouring tile updates in the same launch).
    double *F[2], *G[2], *Bp, *Bp_alt, *aux[2], *mac, *prev;
    double *red;            // reduction partials (device)
    double *wall;           // per-x wall profiles (device, 2 * nx)
    int *flag;              // device error flag
    unsigned red_blocks;
    int fcur, gcur, fy;     // buffer indices, see k*_ kernels
    int phase;              // -1: not initialised, 0: plain state, 1: pipelined
    int first_thermal;      // next thermal pass reads the plain g / T
    int aux_valid;          // aux holds T^n (last step ran the two-pass form)
    int variant;            // LBM_KERNEL_* requested
    bool fused_ok;          // the configuration admits the one-pass form
    bool aa_ok;             // the configuration admits the in-place layout
    bool aa;                // in-place layout: F[0], G[0] only, states fpar / gpar
    int fpar, gpar;         // LAY_ODD / LAY_EVEN of F[0] and G[0] (in-place layout, phase 1)
    bool capturing;         // the handle's stream is being captured into a HIP graph
    FusedGeom fg;
    FusedGeom bfg;          // geometry of the domain's entry in its batch's one-pass table
    int64_t itc;
    bool owns_stream;       // false for the domains of a batch (they share the batch's stream)
    bool in_batch;          // destroyed with its batch, not by lbm_destroy
    uint64_t instance;      // index within a batch (seeded-noise key)
    size_t dev_bytes;       // device memory held by the domain
    // one-pass launch table (lbm_fused.h): entry 0 for the state (ftab_fcur, ftab_gcur,
    // ftab_bp), entry 1 for the state one one-pass step later
    void *ftab;
    bool ftab_ok;
    int ftab_fcur, ftab_gcur;
    const double *ftab_bp;
    char err[256];
};

static int fail(lbm_handle *h, int code, const char *fmt, ...)
{
    if (h) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(h->err, sizeof(h->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define HIPCHK(h, call)                                                                        \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((h), LBM_ERR_HIP, "%s failed: %s", #call, hipGetErrorString(e_));     \
    } while (0)

// Returns false when the legacy thermal scheme's paraA is outside its range (the reference
// stops there: 2DRB/3DRB.F90:378-390, 2DRB/2DRB.F90:404-413).
static bool compute_params(const lbm_config *c, lbm_params *p)
{
    // module commondata3d, 2DRB/3DRB.F90:73-123 (2D: 2DRB/2DRB.F90:61-115)
    memset(p, 0, sizeof(*p));
    p->L = c->side_heated ? (double)c->nx : (double)c->ny;
    p->tref = 0.5 * (c->thot + c->tcold);
    p->tauf = 0.5 + c->mach * p->L * sqrt(3.0 * c->prandtl / c->rayleigh);
    p->viscosity = (p->tauf - 0.5) / 3.0;
    p->diffusivity = p->viscosity / c->prandtl;
    const double gbeta1 = c->rayleigh * p->viscosity * p->diffusivity / p->L;
    p->gbeta = gbeta1 / p->L / p->L;
    p->time_unit = sqrt(p->L / p->gbeta);
    p->velocity_unit = sqrt(p->gbeta * p->L);
    p->s_e = p->s_eps = p->s_nu = p->s_pi = 1.0 / p->tauf;
    p->s_q = p->s_m = 8.0 * (2.0 * p->tauf - 1.0) / (8.0 * p->tauf - 1.0);
    p->taug = c->dims == 3 ? 0.5 + p->diffusivity / 0.25 : 0.5 + (p->tauf - 0.5) / c->prandtl;
    p->q_nu = 1.0;
    p->q_k = 1.0 / p->taug;
    bool ok = true;
    if (c->legacy_thermal) {
        if (c->dims == 3) {
            p->para_a = 42.0 * sqrt(3.0) * p->diffusivity - 6.0;
            ok = !(p->para_a >= 1.0 || p->para_a <= -6.0);
        } else {
            p->para_a = 20.0 * sqrt(3.0) * p->diffusivity - 4.0;
            ok = !(p->para_a >= 1.0 || p->para_a <= -4.0);
        }
        p->taug = 0.0;
        p->q_k = 3.0 - sqrt(3.0);
        p->q_nu = 4.0 * sqrt(3.0) - 6.0;
    }
    p->sg = 1.0 - 0.5 * p->q_k;
    return ok;
}

extern "C" const char *lbm_error_string(int code)
{
    switch (code) {
    case LBM_OK: return "ok";
    case LBM_ERR_ARG: return "invalid argument";
    case LBM_ERR_HIP: return "HIP runtime error";
    case LBM_ERR_NOMEM: return "out of memory";
    case LBM_ERR_STATE: return "call out of order";
    case LBM_ERR_UNSUPPORTED: return "unsupported configuration";
    case LBM_ERR_RHO: return "non-positive density in restart data";
    default: return "unknown error";
    }
}

extern "C" const char *lbm_last_error(lbm_handle *h) { return h ? h->err : "null handle"; }

static void release(lbm_handle *h)
{
    double **bufs[] = {&h->F[0], &h->F[1], &h->G[0], &h->G[1], &h->Bp, &h->Bp_alt, &h->aux[0],
                       &h->aux[1], &h->mac, &h->prev, &h->red, &h->wall};
    for (auto b : bufs)
        if (*b) { (void)hipFree(*b); *b = nullptr; }
    if (h->flag) { (void)hipFree(h->flag); h->flag = nullptr; }
    if (h->ftab) { (void)hipFree(h->ftab); h->ftab = nullptr; }
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    if (h->stream && h->owns_stream) (void)hipStreamDestroy(h->stream);
    h->ev0 = h->ev1 = nullptr;
    h->stream = nullptr;
}

static int create_impl(const lbm_config *cfg, hipStream_t shared, lbm_handle **out);

// Plane pitch (elements) of the plane-structured arrays.  A step streams up to 40 planes at
// once; with a pitch of N nodes at 512^3 every plane is exactly 2^30 bytes from the next and all
// streams start at the same offset modulo every power of two.  The pitch adds 1 KiB plus up to
// 1 MiB (a sixteenth of a plane for small planes) of population elements of size es, so that
// the planes spread over the memory channels.
static uint64_t plane_pitch(uint64_t N, int es)
{
    const uint64_t plane = N * (uint64_t)es;
    uint64_t pad = std::min<uint64_t>(plane / 16, (uint64_t)1 << 20) / 1024 * 1024 + 1024;
    return N + (pad + (uint64_t)es - 1) / (uint64_t)es;
}

// Device bytes per node of a domain: populations (one or two arrays per lattice), B_prev (and
// the one-pass form's second copy), aux, the five FP64 macro planes.
static double node_bytes(const lbm_config &c, bool aa, bool fused)
{
    const int q = c.dims == 3 ? 26 : 14;
    const double es = c.precision == LBM_FP64 ? 8 : (c.precision == LBM_FP32_SHIFT ? 4 : 2);
    const double ea = c.precision == LBM_FP64 ? 8 : 4;
    return (aa ? 1 : 2) * q * es + (fused ? 6 : 3) * ea + (c.dims == 2 ? 2 : 1) * ea + 40.0;
}

// Default z-chunk of the one-pass form: the largest power of two up to nz / 16, within 8..32
// planes (nz itself below 8).  Measured with the default tiles, ms per step:
//   512^3 FP64: 16.8 with chunks of 8 planes, 14.3-15.0 at 16, 14.1-14.9 at 32, 15.1-15.2 at 64;
//   512^3 FP32: 7.1-7.2 at 16, 6.7-6.9 at 32, 6.9-7.8 at 64;
//   256^3 FP64: 2.14 at 8, 1.91 at 16, 1.99 at 32; FP32: 0.93, 0.89, 0.96;
//   192^3 FP64: 0.90 at 8, 0.92 at 16, 1.00 at 32; FP32: 0.467, 0.469, 0.563.
// Shorter chunks give a smaller domain enough blocks to fill the device (the 66x15 FP32 tile
// at 256^3 has 80 tiles a chunk) at the price of more halo planes.
static int default_kchunk(int nz)
{
    int k = 8;
    while (k < 32 && 2 * k <= nz / 16) k *= 2;
    return k > nz ? nz : k;
}

// Default z-chunk of a domain in its batch's one-pass table: min(nz, 32).  The table launch
// fills the device with the tiles of all its domains, so fewer halo planes win over more
// blocks.  The seeded 64-cavity ensemble (edges 32-128) in the one-pass table form, graph
// replay, two processes each: FP32-shift 2.83-2.84 ms per step with the single-domain default
// (8 planes for these depths), 2.70-2.76 with 16, 2.53-2.76 with 32; FP64 4.86-5.34, 4.98-5.15,
// 4.62-5.15.  A domain set to the one-pass form keeps its own z-chunk (lbm_batch_autotune).
static int batch_kchunk(int nz) { return nz < 32 ? nz : 32; }

// The one-pass kernel stores a tile node's values through buffer descriptors based at its plane
// with the node's in-plane byte offset, a 32-bit value: the form needs nx * ny * 8 < 2^32 (every
// plane of populations and history in every precision form).
static bool fused_plane_fits(const lbm_config &c)
{
    return (uint64_t)c.nx * (uint64_t)c.ny * 8u < ((uint64_t)1 << 32);
}

// The one-pass form's default tile and prefetching per precision form (see create_impl).
static void default_tile(int precision, int *tile, int *prefetch)
{
    *tile = precision == LBM_FP32_SHIFT ? LBM_TILE_66X15 : LBM_TILE_34X15;
    *prefetch = precision == LBM_FP32_SHIFT ? LBM_PREFETCH_NONE : LBM_PREFETCH_REGS;
}

// One-pass launch geometry of configuration c for z-chunks of `kchunk` planes
// (LBM_KCHUNK_AUTO: default_kchunk) and thread tile shape `tile`.
static FusedGeom make_fused_geom(const lbm_config &c, int kchunk, int tile, int prefetch)
{
    int tx, ty;
    fused_tile_dims(tile, &tx, &ty);
    FusedGeom g;
    g.tile = tile;
    g.pf = prefetch;
    g.ntx = (unsigned)((c.nx + tx - 3) / (tx - 2));
    g.nty = (unsigned)((c.ny + ty - 3) / (ty - 2));
    if (kchunk <= 0) kchunk = default_kchunk(c.nz);
    g.kchunk = kchunk > c.nz ? c.nz : kchunk;
    g.nzc = (unsigned)((c.nz + g.kchunk - 1) / g.kchunk);
    g.nblocks = g.ntx * g.nty * g.nzc;
    return g;
}
static void set_fused_geom(lbm_handle *h, int kchunk, int tile, int prefetch)
{
    h->fg = make_fused_geom(h->cfg, kchunk, tile, prefetch);
    h->ftab_ok = false;
}

// Pointers of one one-pass step from the buffer state (fcur, gcur, Bp) advanced by `par` steps.
static FusedPtrs fused_ptrs(const lbm_handle *h, int fcur, int gcur, double *bp, double *bp_alt,
                            int par)
{
    const bool sw = par == 1 && h->cfg.use_g;
    FusedPtrs p;
    p.Fcur = h->F[fcur ^ par];
    p.Fnext = h->F[fcur ^ par ^ 1];
    p.Gin = h->G[gcur ^ par];
    p.Gout = h->G[gcur ^ par ^ 1];
    p.Bp = sw ? bp_alt : bp;
    p.Bpout = sw ? bp : bp_alt;
    return p;
}

static size_t fused_entry_bytes(const lbm_handle *h)
{
    return h->lp ? fused_instance_bytes_lp(h->cfg.precision) : fused_instance_bytes_f64();
}

static void fused_fill(const lbm_handle *h, void *dst, const FusedPtrs &p, unsigned first_block,
                       const FusedGeom &g)
{
    if (h->lp) fused_instance_fill_lp(h->cfg.precision, dst, h->kp, g, p, first_block);
    else fused_instance_fill_f64(dst, h->kp, g, p, first_block);
}

// One-pass step of a single domain: launch from the handle's two-entry table, rebuilding it
// when the buffers are not in either state it describes.
// Entry of the one-pass launch table that matches the domain's buffers (-1: out of date).
static int fused_table_parity(const lbm_handle *h)
{
    if (!h->ftab_ok) return -1;
    // a one-pass step flips F and G, and swaps the two history buffers when use_g
    const bool bp_same = h->Bp == h->ftab_bp;
    if (h->fcur == h->ftab_fcur && h->gcur == h->ftab_gcur && bp_same) return 0;
    if (h->fcur != h->ftab_fcur && h->gcur != h->ftab_gcur && bp_same != (bool)h->cfg.use_g) return 1;
    return -1;
}

static int launch_fused_single(lbm_handle *h)
{
    const size_t isz = fused_entry_bytes(h);
    int par = fused_table_parity(h);
    if (par < 0) {
        if (h->capturing)   // the rebuild below synchronises: not inside a capture
            return fail(h, LBM_ERR_STATE, "one-pass launch table out of date during graph capture");
        if (!h->ftab) return fail(h, LBM_ERR_STATE, "one-pass launch table missing");
        // synchronous: the table is rebuilt only after an initialisation, a two-pass step or a
        // change of the kernel form, never inside a run of one-pass steps
        HIPCHK(h, hipStreamSynchronize(h->stream));
        std::vector<unsigned char> host(2 * isz);
        for (int q = 0; q < 2; ++q)
            fused_fill(h, host.data() + q * isz, fused_ptrs(h, h->fcur, h->gcur, h->Bp, h->Bp_alt, q), 0, h->fg);
        HIPCHK(h, hipMemcpy(h->ftab, host.data(), 2 * isz, hipMemcpyHostToDevice));
        h->ftab_ok = true;
        h->ftab_fcur = h->fcur;
        h->ftab_gcur = h->gcur;
        h->ftab_bp = h->Bp;
        par = 0;
    }
    const void *entry = (const unsigned char *)h->ftab + par * isz;
    if (h->lp) HIPCHK(h, launch_fused_batch_lp(h->cfg.precision, h->fg, entry, nullptr, h->fg.nblocks, h->stream));
    else HIPCHK(h, launch_fused_batch_f64(h->fg, entry, nullptr, h->fg.nblocks, h->stream));
    return LBM_OK;
}

extern "C" int lbm_create(const lbm_config *cfg, lbm_handle **out)
{
    return create_impl(cfg, nullptr, out);
}

static int create_impl(const lbm_config *cfg, hipStream_t shared, lbm_handle **out)
{
    if (!cfg || !out) return LBM_ERR_ARG;
    *out = nullptr;
    lbm_config c = *cfg;
    if (c.dims != 2 && c.dims != 3) return LBM_ERR_ARG;
    if (c.dims == 2) c.nz = 1;
    if (c.nx < 4 || c.ny < 4 || (c.dims == 3 && c.nz < 4)) return LBM_ERR_ARG;
    if (c.nx > 65536 || c.ny > 65536 || c.nz > 65536) return LBM_ERR_ARG;
    if (!(c.rayleigh > 0) || !(c.prandtl > 0) || !(c.mach > 0)) return LBM_ERR_ARG;
    if (c.vel_x < 0 || c.vel_x > 1 || c.vel_z < 0 || c.vel_z > 1) return LBM_ERR_ARG;
    if (c.temp_x < 0 || c.temp_x > 2 || c.temp_z < 0 || c.temp_z > 2) return LBM_ERR_ARG;
    if (c.temp_y < 0 || c.temp_y > 1) return LBM_ERR_ARG;  // no periodic option for y walls
    if (c.dims == 3 && c.temp_z == LBM_T_CONST) return LBM_ERR_UNSUPPORTED;
    if (c.precision < LBM_FP64 || c.precision > LBM_FP16_SHIFT) return LBM_ERR_ARG;
    if (c.precision != LBM_FP64 && c.dims != 3) return LBM_ERR_UNSUPPORTED;
    if (c.legacy_thermal < 0 || c.legacy_thermal > 1) return LBM_ERR_ARG;
    if (c.legacy_thermal && c.precision != LBM_FP64) return LBM_ERR_UNSUPPORTED;

    lbm_handle *h = (lbm_handle *)calloc(1, sizeof(lbm_handle));
    if (!h) return LBM_ERR_NOMEM;
    h->cfg = c;
    if (!compute_params(&c, &h->p)) {   // legacy thermal scheme outside its paraA range
        free(h);
        return LBM_ERR_ARG;
    }
    h->qf = c.dims == 3 ? 19 : 9;
    h->qt = c.dims == 3 ? 7 : 5;
    h->N = (uint64_t)c.nx * c.ny * c.nz;
    h->phase = -1;
    if ((uint64_t)c.nx * c.ny * c.nz >= (1ULL << 32)) { free(h); return LBM_ERR_ARG; }

    KParams &k = h->kp;
    k.nx = c.nx; k.ny = c.ny; k.nz = c.nz;
    k.vel_x = c.vel_x; k.vel_z = c.dims == 3 ? c.vel_z : 0;
    k.temp_x = c.temp_x; k.temp_y = c.temp_y; k.temp_z = c.dims == 3 ? c.temp_z : 1;
    k.use_g = c.use_g ? 1 : 0;
    k.N = h->N;
    k.P = plane_pitch(h->N, c.precision == LBM_FP64 ? 8 : (c.precision == LBM_FP32_SHIFT ? 4 : 2));
    k.gbeta = h->p.gbeta; k.tref = h->p.tref; k.thot = c.thot; k.tcold = c.tcold;
    k.sv = 1.0 / h->p.tauf; k.sq = h->p.s_q;
    k.qk = h->p.q_k; k.qnu = h->p.q_nu; k.sg = h->p.sg;
    // thermal scheme constants (2DRB/3DRB.F90:116-122, :671-676, :1153-1157, :1267-1316;
    // 2DRB/2DRB.F90:108-114, :514-522, :949-953, :1048-1098)
    if (c.dims == 3) {
        k.neqc = c.legacy_thermal ? h->p.para_a : -0.75;
        k.abb = c.legacy_thermal ? (6.0 + h->p.para_a) / 21.0 : 2.0 * (1.0 / 8.0);
        k.wt0 = c.legacy_thermal ? (1.0 - h->p.para_a) / 7.0 : 1.0 / 4.0;
        k.wt1 = c.legacy_thermal ? (6.0 + h->p.para_a) / 42.0 : 1.0 / 8.0;
    } else {
        k.neqc = c.legacy_thermal ? h->p.para_a : -2.0 / 3.0;
        k.abb = c.legacy_thermal ? (4.0 + h->p.para_a) / 10.0 : 2.0 * (1.0 / 6.0);
        k.wt0 = c.legacy_thermal ? (1.0 - h->p.para_a) / 5.0 : 1.0 / 3.0;
        k.wt1 = c.legacy_thermal ? (h->p.para_a + 4.0) / 20.0 : 1.0 / 6.0;
    }
    h->bx = c.nx >= 256 ? 256 : ((c.nx + 63) / 64) * 64;
    k.nxb = (c.nx + h->bx - 1) / h->bx;
    const uint64_t nb = (uint64_t)k.nxb * c.ny * c.nz;
    if (nb >= (1ULL << 31)) { free(h); return LBM_ERR_ARG; }
    k.nblocks = (unsigned)nb;
    // one-pass form: 3D without periodic thermal walls (see lbm_fused.h); in-place layout: 3D
    // with a wall on every face (a same-row periodic copy would be read by two nodes)
    h->aa_ok = c.dims == 3 && c.vel_x != LBM_VEL_PERIODIC && c.vel_z != LBM_VEL_PERIODIC &&
               c.temp_x != LBM_T_PERIODIC && c.temp_z != LBM_T_PERIODIC;
    if (c.layout == LBM_LAYOUT_AA) {
        if (!h->aa_ok) { free(h); return LBM_ERR_UNSUPPORTED; }
        h->aa = true;
    } else if (c.layout == LBM_LAYOUT_AUTO && h->aa_ok) {
        // two arrays when they fit in free memory (with 2% to spare), else in place
        size_t fr = 0, tot = 0;
        const double two = (double)h->N * node_bytes(c, false, true);
        if (hipSetDevice(c.device) == hipSuccess && hipMemGetInfo(&fr, &tot) == hipSuccess)
            h->aa = two > 0.98 * (double)fr;
    } else if (c.layout != LBM_LAYOUT_TWO_ARRAY && c.layout != LBM_LAYOUT_AUTO) {
        free(h);
        return LBM_ERR_ARG;
    }
    h->fused_ok = c.dims == 3 && c.temp_x != LBM_T_PERIODIC && c.temp_z != LBM_T_PERIODIC && !h->aa &&
                  fused_plane_fits(c);
    h->variant = LBM_KERNEL_AUTO;
    // the one-pass form's default shape, the fastest one-pass form at 512^3: FP64 the 34x15 tile
    // with prefetching into registers; FP32-shift the 66x15 tile without prefetching (95
    // registers, no spills: 6.72-6.94 ms per step against 6.93-7.58 ms staged, alternating
    // processes on one box); the FP16-shift form runs the two-pass form under AUTO
    int tile0, pf0;
    default_tile(c.precision, &tile0, &pf0);
    set_fused_geom(h, LBM_KCHUNK_AUTO, tile0, pf0);

    int rc = LBM_OK;
    hipError_t e = hipSetDevice(c.device);
    if (e != hipSuccess) { free(h); return LBM_ERR_HIP; }
    h->owns_stream = shared == nullptr;
    if (shared) h->stream = shared;
    if ((!shared && hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) ||
        hipEventCreate(&h->ev0) != hipSuccess || hipEventCreate(&h->ev1) != hipSuccess) {
        release(h);
        free(h);
        return LBM_ERR_HIP;
    }
    const size_t N = h->kp.P;   // plane pitch
    h->lp = c.precision != LBM_FP64;
    h->es = c.precision == LBM_FP64 ? 8 : (c.precision == LBM_FP32_SHIFT ? 4 : 2);
    const size_t ea = h->lp ? sizeof(float) : sizeof(double);   // B_prev / aux element
    const size_t two = h->aa ? 0 : 1;   // second population arrays
    struct { double **p; size_t bytes; } plan[] = {
        {&h->F[0], N * h->qf * h->es}, {&h->F[1], two * N * h->qf * h->es},
        {&h->G[0], N * h->qt * h->es}, {&h->G[1], two * N * h->qt * h->es},
        {&h->Bp, N * 3 * ea},          {&h->aux[0], N * ea},
        {&h->Bp_alt, (h->fused_ok ? N * 3 : 1) * ea},
        {&h->aux[1], (c.dims == 2 ? N : 1) * ea},
        {&h->mac, N * 5 * sizeof(double)},
        {&h->wall, (size_t)2 * (c.nx > c.ny ? c.nx : c.ny) * sizeof(double)},
    };
    // A large domain is allocated behind a block of its own size, which is freed afterwards.
    // The memory a process allocates first maps slower: on a freshly started GPU, a 512^3 FP64
    // domain allocated first ran the two-pass step in 18.0-18.1 ms for as long as it lived,
    // the same domain allocated after an 84 GiB block (or after another domain) in 16.5-16.8 ms,
    // after an 8 GiB block in 18.1 ms, and a domain later allocated into the freed first block
    // in 17.6-17.9 ms.  Best effort: skipped when free memory cannot hold both.
    size_t total = 0;
    for (auto &b : plan) total += b.bytes;
    void *ahead = nullptr;
    size_t free_b = 0, all_b = 0;
    if (total >= ((size_t)4 << 30) && hipMemGetInfo(&free_b, &all_b) == hipSuccess &&
        free_b >= 2 * total + ((size_t)2 << 30) && hipMalloc(&ahead, total) != hipSuccess) {
        ahead = nullptr;
        (void)hipGetLastError();   // not an error of the domain
    }
    for (auto &b : plan) {
        if (b.bytes == 0) continue;
        if (hipMalloc((void **)b.p, b.bytes) != hipSuccess) { rc = LBM_ERR_NOMEM; break; }
        h->dev_bytes += b.bytes;
    }
    if (ahead) (void)hipFree(ahead);
    h->red_blocks = 1024;
    // the one-pass launch table is allocated with the domain, so that no step allocates
    if (rc == LBM_OK && h->fused_ok && hipMalloc(&h->ftab, 2 * fused_entry_bytes(h)) != hipSuccess)
        rc = LBM_ERR_NOMEM;
    if (rc == LBM_OK) {
        // volume partials (6 sums) then wall partials (4 sums per wall-profile block)
        const unsigned nw = (unsigned)(c.nx > c.ny ? c.nx : c.ny);   // wall-profile blocks
        size_t nred = (size_t)2 * 6 * h->red_blocks + (size_t)2 * 4 * nw;
        if (hipMalloc((void **)&h->red, nred * sizeof(double)) != hipSuccess ||
            hipMalloc((void **)&h->flag, sizeof(int)) != hipSuccess)
            rc = LBM_ERR_NOMEM;
    }
    if (rc != LBM_OK) {
        release(h);
        free(h);
        return rc;
    }
    *out = h;
    return LBM_OK;
}

extern "C" int lbm_destroy(lbm_handle *h)
{
    if (!h) return LBM_ERR_ARG;
    if (h->in_batch) return fail(h, LBM_ERR_STATE, "domain belongs to a batch: use lbm_batch_destroy");
    (void)hipSetDevice(h->cfg.device);
    (void)hipStreamSynchronize(h->stream);
    release(h);
    free(h);
    return LBM_OK;
}

static int do_init(lbm_handle *h, const double *T0, uint64_t seed, double amp, bool seeded)
{
    HIPCHK(h, hipSetDevice(h->cfg.device));
    const lbm_config &c = h->cfg;
    InitParams ip{};
    ip.use_T0 = T0 != nullptr;
    ip.L = h->p.L;
    ip.pi = acos(-1.0);
    ip.perturb_amp = 1.0e-3 * (c.thot - c.tcold);
    ip.rb_perturb = 0;
    if (!c.side_heated && c.temp_y == LBM_T_CONST) {
        if (c.dims == 3 && c.rayleigh <= 1.0e4) ip.rb_perturb = 3;   // 2DRB/3DRB.F90:542
        if (c.dims == 2 && c.rayleigh < 1.0e4) ip.rb_perturb = 2;    // 2DRB/2DRB.F90:568
    }
    ip.seed = seed;
    ip.instance = h->instance;
    ip.noise_scale = seeded ? amp * (c.thot - c.tcold) : 0.0;
    if (T0) HIPCHK(h, hipMemcpyAsync(h->mac + 4 * h->kp.P, T0, h->N * sizeof(double),
                                     hipMemcpyHostToDevice, h->stream));
    if (!h->lp) {
        HIPCHK(h, launch_init(h->kp, h->bx, h->stream, ip, h->F[0], h->G[0], h->Bp, h->mac, h->qf, h->qt));
    } else {
        // macros (T profile, noise) in FP64 exactly as the FP64 form, then the deviation state
        HIPCHK(h, launch_init(h->kp, h->bx, h->stream, ip, nullptr, nullptr, nullptr, h->mac, h->qf, h->qt));
        LPArgs a;
        a.Fn = h->F[0];
        a.Go = h->G[0];
        a.Bp = (float *)h->Bp;
        a.mac = h->mac;
        HIPCHK(h, launch_lp(h->cfg.precision, LP_INIT, h->kp, h->bx, h->stream, a));
    }
    if (h->prev) HIPCHK(h, hipMemsetAsync(h->prev, 0, 4 * h->kp.P * sizeof(double), h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->phase = 0;
    h->itc = 0;
    return LBM_OK;
}

// A domain of a batch changes state only through the lbm_batch_* calls: its buffers are also
// described by the batch's launch tables, which a step, initialisation or scatter of the domain
// alone would leave out of date.
static int refuse_batch_domain(lbm_handle *h, const char *what)
{
    return fail(h, LBM_ERR_STATE, "%s on a domain of a batch: use the lbm_batch_* calls", what);
}

extern "C" int lbm_init_exact(lbm_handle *h, const double *T0)
{
    if (!h) return LBM_ERR_ARG;
    if (h->in_batch) return refuse_batch_domain(h, "lbm_init_exact");
    return do_init(h, T0, 0, 0.0, false);
}

extern "C" int lbm_init_seeded(lbm_handle *h, uint64_t seed, double amp)
{
    if (!h || !(amp >= 0.0)) return LBM_ERR_ARG;
    if (h->in_batch) return refuse_batch_domain(h, "lbm_init_seeded");
    return do_init(h, nullptr, seed, amp, true);
}

// The kernel form a step of domain h runs (after its first step, which is always two-pass):
// ONE_PASS as requested; under AUTO the one-pass form for FP64 domains of 2^22 nodes or more
// and FP32-shift domains of 2^24 nodes or more, else the two-pass form.  Measured, ms per step
// (two-pass against one-pass in its default shape): FP64 512^3 16.4 / 14.2, 256^3 2.28 / 1.91,
// 192^3 0.98 / 0.90, 128^3 0.301 / 0.302; FP32 512^3 8.1 / 6.8, 256^3 1.11 / 0.89, 192^3
// 0.455 / 0.467, 128^3 0.143 / 0.174.  The FP16-shift one-pass kernel is 1.2x slower at 512^3.
// The one-pass form needs the configuration to admit it.
static uint64_t one_pass_min_nodes(int precision)
{
    return precision == LBM_FP64 ? (uint64_t)1 << 22 : (uint64_t)1 << 24;
}
static bool auto_one_pass(const lbm_handle *h)
{
    return h->fused_ok && h->cfg.precision != LBM_FP16_SHIFT &&
           h->N >= one_pass_min_nodes(h->cfg.precision);
}
static bool runs_one_pass(const lbm_handle *h)
{
    if (!h->fused_ok) return false;
    if (h->variant == LBM_KERNEL_ONE_PASS) return true;
    return h->variant == LBM_KERNEL_AUTO && auto_one_pass(h);
}

// In-place layout.  A step is the flow collision (pass B) of the step before, then pass A, so
// that between steps F[0] and G[0] hold f_post^n and g_post^n (both in state fpar == gpar) and
// the reference state f^{n+1}, g^{n+1} can still be read from them.  The first step has no
// pass B: the prologue did that collision.
static int enqueue_steps_aa(lbm_handle *h, int32_t nsteps)
{
    if (h->phase == 0) {
        if (h->lp) {
            LPArgs a;
            a.Fn = h->F[0];
            a.mac = h->mac;
            a.lay = LAY_ODD;
            HIPCHK(h, launch_lp(h->cfg.precision, LP_PROLOGUE, h->kp, h->bx, h->stream, a));
        } else {
            HIPCHK(h, launch_prologue(h->kp, h->bx, h->stream, nullptr, h->F[0], h->mac, nullptr,
                                      true, true));
        }
        h->fpar = LAY_EVEN;   // the plain state is LAY_ODD; the prologue turned F over
        h->gpar = LAY_ODD;
        h->first_thermal = 1;
        h->phase = 1;
    }
    for (int32_t n = 0; n < nsteps; ++n) {
        if (!h->first_thermal) {
            if (h->lp) {
                LPArgs a;
                a.Fn = h->F[0];
                a.Gi = h->G[0];
                a.aux = (float *)h->aux[0];
                a.lay = h->fpar;
                HIPCHK(h, launch_lp(h->cfg.precision, LP_FLOW_AA, h->kp, h->bx, h->stream, a));
            } else {
                HIPCHK(h, launch_flow3_aa(h->kp, h->bx, h->stream, h->fpar, h->F[0], h->G[0], h->aux[0]));
            }
            h->fpar = 3 - h->fpar;
        }
        if (h->lp) {
            LPArgs a;
            a.Fc = h->F[0];
            a.Go = h->G[0];
            a.Bp = (float *)h->Bp;
            a.aux = (float *)h->aux[0];
            a.mac = h->mac;
            a.first = h->first_thermal;
            a.lay = h->gpar;
            HIPCHK(h, launch_lp(h->cfg.precision, LP_THERMAL_AA, h->kp, h->bx, h->stream, a));
        } else {
            HIPCHK(h, launch_thermal3_aa(h->kp, h->bx, h->stream, h->gpar, h->first_thermal != 0,
                                         h->F[0], h->G[0], h->Bp, h->aux[0], h->mac));
        }
        h->gpar = 3 - h->gpar;
        h->aux_valid = 1;
        h->first_thermal = 0;
        h->itc += 1;
    }
    return LBM_OK;
}

static int enqueue_steps(lbm_handle *h, int32_t nsteps)
{
    if (h->aa) return enqueue_steps_aa(h, nsteps);
    const bool d3 = h->cfg.dims == 3;
    if (h->phase == 0) {
        // flow collision of step 1 from the plain state (f^1 in F[0], macros in mac)
        if (h->lp) {
            LPArgs a;
            a.Fc = h->F[0];
            a.Fn = h->F[1];
            a.mac = h->mac;
            HIPCHK(h, launch_lp(h->cfg.precision, LP_PROLOGUE, h->kp, h->bx, h->stream, a));
        } else {
            HIPCHK(h, launch_prologue(h->kp, h->bx, h->stream, h->F[0], h->F[1], h->mac, h->aux[1], d3));
        }
        h->fcur = 1;
        h->gcur = 0;
        h->fy = 1;
        h->first_thermal = 1;
        h->phase = 1;
    }
    for (int32_t n = 0; n < nsteps; ++n) {
        const double *Fc = h->F[h->fcur];
        double *Fn = h->F[h->fcur ^ 1];
        const double *Gi = h->G[h->gcur];
        double *Go = h->G[h->gcur ^ 1];
        // the first step after (re)initialisation reads the plain state: two-pass form.  AUTO
        // runs the two-pass form, the faster one at 512^3 on MI355X in every precision form;
        // lbm_autotune measures both per shape.
        const bool fused = !h->first_thermal && runs_one_pass(h);
        if (fused) {
            int rc = launch_fused_single(h);
            if (rc) return rc;
            h->aux_valid = 0;
            if (h->cfg.use_g) std::swap(h->Bp, h->Bp_alt);
        } else if (h->lp) {
            LPArgs a;
            a.Fc = Fc;
            a.Fn = Fn;
            a.Gi = Gi;
            a.Go = Go;
            a.Bp = (float *)h->Bp;
            a.aux = (float *)h->aux[0];
            a.mac = h->mac;
            a.first = h->first_thermal;
            HIPCHK(h, launch_lp(h->cfg.precision, LP_STEP, h->kp, h->bx, h->stream, a));
            h->aux_valid = 1;
        } else if (d3) {
            HIPCHK(h, launch_step3(h->kp, h->bx, h->stream, h->first_thermal, Fc, Fn, Gi, Go, h->Bp,
                                   h->aux[0], h->mac));
            h->aux_valid = 1;
        } else {
            HIPCHK(h, launch_step2(h->kp, h->bx, h->stream, h->first_thermal, Fc, Fn, Gi, Go, h->Bp,
                                   h->aux[h->fy], h->aux[h->fy ^ 1], h->mac));
            h->fy ^= 1;
        }
        h->first_thermal = 0;
        h->fcur ^= 1;
        h->gcur ^= 1;
        h->itc += 1;
    }
    return LBM_OK;
}

// Runs of many steps on a small domain replay a HIP graph of kReplaySteps steps: there a step is
// launch-bound (a 257 x 129 2D step is two launches of a few microseconds).  kReplaySteps is
// even, so every alternating buffer, in-place state and history buffer is back in its place
// after each replay; the first step after an initialisation reads the plain state and runs
// eagerly.  The graph holds the same launches on the same buffers as the eager loop.
constexpr int kReplaySteps = 32;
constexpr uint64_t kReplayNodes = (uint64_t)1 << 22;

// A step can be captured once the domain has left its plain state (the first step after an
// initialisation runs the two-pass form from it) and, for the one-pass form, once the launch
// table matches the buffers (rebuilding it synchronises, which a capture cannot).
static bool capture_ready(const lbm_handle *h)
{
    if (h->phase == 0 || h->first_thermal) return false;
    const bool one_pass = !h->aa && runs_one_pass(h);
    return !one_pass || fused_table_parity(h) >= 0;
}

static int enqueue_steps_replayed(lbm_handle *h, int32_t nsteps)
{
    if (h->capturing || h->in_batch || h->N > kReplayNodes || nsteps < 4 * kReplaySteps)
        return enqueue_steps(h, nsteps);
    int rc = LBM_OK;
    for (int k = 0; k < 2 && !capture_ready(h); ++k) {   // at most two eager steps
        rc = enqueue_steps(h, 1);
        if (rc) return rc;
        --nsteps;
    }
    if (!capture_ready(h)) return enqueue_steps(h, nsteps);
    const int replays = nsteps / kReplaySteps;
    hipGraph_t g = nullptr;
    hipGraphExec_t ge = nullptr;
    const int64_t itc0 = h->itc;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    h->capturing = true;
    rc = enqueue_steps(h, kReplaySteps);
    h->capturing = false;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (!rc && e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
    int launched = 0;
    for (; launched < replays && !rc && e == hipSuccess; ++launched) {
        e = hipGraphLaunch(ge, h->stream);
        if (e != hipSuccess) break;
    }
    if (ge) (void)hipGraphExecDestroy(ge);
    if (g) (void)hipGraphDestroy(g);
    // the capture advanced the host state by an even number of steps (buffers back in place);
    // the step counter follows the replays that were launched
    h->itc = itc0 + (int64_t)launched * kReplaySteps;
    if (rc) return rc;
    if (e != hipSuccess) return fail(h, LBM_ERR_HIP, "step graph: %s", hipGetErrorString(e));
    return enqueue_steps(h, nsteps - replays * kReplaySteps);
}

extern "C" int lbm_step(lbm_handle *h, int32_t nsteps)
{
    if (!h || nsteps < 0) return LBM_ERR_ARG;
    if (h->in_batch) return refuse_batch_domain(h, "lbm_step");
    if (h->phase < 0) return fail(h, LBM_ERR_STATE, "lbm_step before initialisation");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    return enqueue_steps_replayed(h, nsteps);
}

extern "C" int lbm_step_timed(lbm_handle *h, int32_t nsteps, double *ms)
{
    if (!h || nsteps < 0 || !ms) return LBM_ERR_ARG;
    if (h->in_batch) return refuse_batch_domain(h, "lbm_step_timed");
    if (h->phase < 0) return fail(h, LBM_ERR_STATE, "lbm_step before initialisation");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    if (h->phase == 0) {            // keep the one-off prologue out of the timed region
        int rc = enqueue_steps(h, 0);
        if (rc) return rc;
    }
    HIPCHK(h, hipEventRecord(h->ev0, h->stream));
    int rc = enqueue_steps_replayed(h, nsteps);
    if (rc) return rc;
    HIPCHK(h, hipEventRecord(h->ev1, h->stream));
    HIPCHK(h, hipEventSynchronize(h->ev1));
    float t = 0.f;
    HIPCHK(h, hipEventElapsedTime(&t, h->ev0, h->ev1));
    *ms = t;
    return LBM_OK;
}

extern "C" int lbm_sync(lbm_handle *h)
{
    if (!h) return LBM_ERR_ARG;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

extern "C" int lbm_itc(lbm_handle *h, int64_t *itc)
{
    if (!h || !itc) return LBM_ERR_ARG;
    *itc = h->itc;
    return LBM_OK;
}

extern "C" int lbm_get_params(lbm_handle *h, lbm_params *p)
{
    if (!h || !p) return LBM_ERR_ARG;
    *p = h->p;
    return LBM_OK;
}

extern "C" int lbm_bytes_per_update(lbm_handle *h, double *bytes)
{
    if (!h || !bytes) return LBM_ERR_ARG;
    const double s_pop = h->es, s_aux = h->lp ? 4.0 : 8.0;
    const int q = h->qf + h->qt, naux = h->cfg.dims == 3 ? 3 : 2;
    *bytes = 2.0 * q * s_pop + 2.0 * naux * s_aux;
    return LBM_OK;
}

// Kernel form of one domain (no batch check: the batch code retunes its own domains).
static int set_form(lbm_handle *h, int variant, int kchunk, int tile, int prefetch)
{
    if (variant < LBM_KERNEL_AUTO || variant > LBM_KERNEL_ONE_PASS || kchunk < LBM_KCHUNK_AUTO ||
        tile < LBM_TILE_32X16 || tile > LBM_TILE_66X15 || prefetch < LBM_PREFETCH_NONE ||
        prefetch > LBM_PREFETCH_STAGED)
        return LBM_ERR_ARG;
    if ((tile == LBM_TILE_32X32 || tile == LBM_TILE_66X15) && !h->lp)
        return fail(h, LBM_ERR_UNSUPPORTED, "the 32x32 and 66x15 tiles need a reduced-precision form");
    if (prefetch == LBM_PREFETCH_STAGED &&
        (h->cfg.precision != LBM_FP32_SHIFT || (tile != LBM_TILE_34X15 && tile != LBM_TILE_66X15)))
        return fail(h, LBM_ERR_UNSUPPORTED,
                    "staged prefetching needs the FP32-shift form and the 34x15 or 66x15 tile");
    if (variant == LBM_KERNEL_ONE_PASS && !h->fused_ok)
        return fail(h, LBM_ERR_UNSUPPORTED,
                    "one-pass form needs 3D, non-periodic thermal walls and two population arrays");
    h->variant = variant;
    set_fused_geom(h, kchunk != 0 ? kchunk : h->fg.kchunk, tile, prefetch);
    return LBM_OK;
}

extern "C" int lbm_set_kernel_variant(lbm_handle *h, int32_t variant, int32_t kchunk)
{
    if (!h) return LBM_ERR_ARG;
    if (h->in_batch)
        return fail(h, LBM_ERR_STATE, "kernel form of a batch domain: use lbm_batch_autotune");
    return set_form(h, variant, kchunk, h->fg.tile, h->fg.pf);
}

extern "C" int lbm_set_fused_tile(lbm_handle *h, int32_t tile, int32_t prefetch)
{
    if (!h) return LBM_ERR_ARG;
    if (h->in_batch)
        return fail(h, LBM_ERR_STATE, "kernel form of a batch domain: use lbm_batch_autotune");
    return set_form(h, h->variant, 0, tile, prefetch);
}

extern "C" int lbm_get_layout(lbm_handle *h, int32_t *layout)
{
    if (!h || !layout) return LBM_ERR_ARG;
    *layout = h->aa ? LBM_LAYOUT_AA : LBM_LAYOUT_TWO_ARRAY;
    return LBM_OK;
}

extern "C" int lbm_get_kernel_form(lbm_handle *h, int32_t *variant, int32_t *kchunk, int32_t *tile,
                                   int32_t *prefetch)
{
    if (!h) return LBM_ERR_ARG;
    if (variant) *variant = runs_one_pass(h) ? LBM_KERNEL_ONE_PASS : LBM_KERNEL_TWO_PASS;
    if (kchunk) *kchunk = h->fg.kchunk;
    if (tile) *tile = h->fg.tile;
    if (prefetch) *prefetch = h->fg.pf;
    return LBM_OK;
}

// ---- autotuner ---------------------------------------------------------------------------
// An AUTO-layout domain that has not stepped may still change its layout.
static bool layout_open(const lbm_handle *h)
{
    return h->cfg.layout == LBM_LAYOUT_AUTO && h->aa_ok && h->phase <= 0;
}

static void tune_key(const lbm_handle *h, char *key, size_t n)
{
    const lbm_config &c = h->cfg;
    snprintf(key, n, "d%d_%dx%dx%d_p%d_w%d%d%d%d%d_l%d", c.dims, c.nx, c.ny, c.nz, c.precision,
             c.vel_x, c.vel_z, c.temp_x, c.temp_y, c.temp_z,
             layout_open(h) ? LBM_LAYOUT_AUTO : (h->aa ? LBM_LAYOUT_AA : LBM_LAYOUT_TWO_ARRAY));
}

// Cache lines: "<key> <variant> <kchunk> <tile> <prefetch> <ms per step> <layout>"; the last
// line for a key wins.
struct TuneChoice { int variant, kchunk, tile, pf, aa; };

// Switch the layout of a domain in its plain state (F[0], G[0] hold the same data in both).
static int convert_layout(lbm_handle *h, bool aa)
{
    if (aa == h->aa) return LBM_OK;
    if (h->phase > 0 || !h->aa_ok) return fail(h, LBM_ERR_STATE, "layout change after stepping");
    const size_t N = h->kp.P, ea = h->lp ? sizeof(float) : sizeof(double);   // plane pitch
    const size_t fb = N * h->qf * h->es, gb = N * h->qt * h->es;
    if (aa) {
        (void)hipStreamSynchronize(h->stream);
        (void)hipFree(h->F[1]);
        (void)hipFree(h->G[1]);
        (void)hipFree(h->Bp_alt);
        h->F[1] = h->G[1] = h->Bp_alt = nullptr;
        h->dev_bytes -= fb + gb + N * 3 * ea;
        h->aa = true;
        h->fused_ok = false;
        h->ftab_ok = false;
        if (h->variant == LBM_KERNEL_ONE_PASS) h->variant = LBM_KERNEL_AUTO;
        return LBM_OK;
    }
    double *f1 = nullptr, *g1 = nullptr, *b1 = nullptr;
    if (hipMalloc((void **)&f1, fb) != hipSuccess || hipMalloc((void **)&g1, gb) != hipSuccess ||
        hipMalloc((void **)&b1, N * 3 * ea) != hipSuccess) {
        (void)hipFree(f1);
        (void)hipFree(g1);
        (void)hipFree(b1);
        return fail(h, LBM_ERR_NOMEM, "second population arrays");
    }
    h->F[1] = f1;
    h->G[1] = g1;
    h->Bp_alt = b1;
    h->dev_bytes += fb + gb + N * 3 * ea;
    h->aa = false;
    h->fused_ok = fused_plane_fits(h->cfg);   // aa_ok implies the wall conditions
    return LBM_OK;
}

static bool tune_lookup(const char *path, const char *key, TuneChoice *out)
{
    if (!path) return false;
    FILE *fp = fopen(path, "r");
    if (!fp) return false;
    char line[256], k[128];
    bool found = false;
    while (fgets(line, sizeof(line), fp)) {
        TuneChoice c;
        double ms;
        int lay = LBM_LAYOUT_TWO_ARRAY;
        const int got = sscanf(line, "%127s %d %d %d %d %lf %d", k, &c.variant, &c.kchunk, &c.tile,
                               &c.pf, &ms, &lay);
        if (got >= 6 && strcmp(k, key) == 0) {
            c.aa = lay == LBM_LAYOUT_AA;
            *out = c;
            found = true;
        }
    }
    fclose(fp);
    return found;
}

// Time every candidate form on a scratch domain of the same configuration (z shortened when
// the device cannot hold a second copy) and keep the fastest.
static int autotune_impl(lbm_handle *h, const char *cache_path, int32_t steps, double *ms_per_step)
{
    char key[128];
    tune_key(h, key, sizeof(key));
    if (ms_per_step) *ms_per_step = 0.0;
    TuneChoice win{};
    const bool open = layout_open(h);
    if (tune_lookup(cache_path, key, &win)) {
        int rc = open ? convert_layout(h, win.aa != 0) : LBM_OK;
        return rc ? rc : set_form(h, win.variant, win.kchunk, win.tile, win.pf);
    }

    size_t free_b = 0, total_b = 0;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    HIPCHK(h, hipMemGetInfo(&free_b, &total_b));
    // time per full-size step of the fastest candidate over the layouts open to the domain.  The
    // first candidate (the two-pass form in the first layout timed) is the incumbent; another
    // replaces the best only when it is more than 3% faster: the same kernels run up to 11%
    // apart depending on where an allocation lands, and the scratch domain's placement is not
    // the domain's own.
    constexpr double margin = 0.97;
    double best = 1e300;
    int rc = LBM_OK;
    bool timed = false;
    const bool lays[2] = {false, true};
    for (bool aa : lays) {
        if (rc) break;
        if (open ? false : aa != h->aa) continue;
        lbm_config sc = h->cfg;
        sc.layout = aa ? LBM_LAYOUT_AA : LBM_LAYOUT_TWO_ARRAY;
        const double per_plane = node_bytes(sc, aa, !aa && h->aa_ok) * sc.nx * (sc.dims == 3 ? sc.ny : 1);
        const int fit = (int)(0.8 * (double)free_b / per_plane);
        if (sc.dims == 3 && fit < sc.nz) {
            if (fit < 16) continue;   // no room for a scratch domain of this layout
            sc.nz = fit;
        }
        lbm_handle *t = nullptr;
        rc = create_impl(&sc, nullptr, &t);
        if (rc) return fail(h, rc, "autotune: scratch domain: %s", lbm_error_string(rc));
        rc = lbm_init_seeded(t, 1, 1e-3);
        if (!rc) rc = lbm_step(t, 1);   // leave the plain state: the candidates time steady steps
        std::vector<TuneChoice> cands;
        // the incumbent: the form AUTO runs in this layout (its default one-pass shape for large
        // FP64 and FP32-shift domains; the scratch domain is judged by the main domain's size)
        if (!aa && t->fused_ok && h->cfg.precision != LBM_FP16_SHIFT &&
            h->N >= one_pass_min_nodes(h->cfg.precision))
            cands.push_back({LBM_KERNEL_ONE_PASS, t->fg.kchunk, t->fg.tile, t->fg.pf, 0});
        cands.push_back({LBM_KERNEL_TWO_PASS, h->fg.kchunk, h->fg.tile, h->fg.pf, aa ? 1 : 0});
        if (t->fused_ok)
            for (int tile = 0; tile < FUSED_TILES; ++tile)
                for (int pf = LBM_PREFETCH_NONE; pf <= LBM_PREFETCH_STAGED; ++pf) {
                    const bool lp_tile = tile == LBM_TILE_32X32 || tile == LBM_TILE_66X15;
                    const bool staged_ok = h->cfg.precision == LBM_FP32_SHIFT &&
                                           (tile == LBM_TILE_34X15 || tile == LBM_TILE_66X15);
                    if ((lp_tile && !h->lp) || (pf == LBM_PREFETCH_STAGED && !staged_ok)) continue;
                    for (int kc : {8, 16, 32, 64, 128})
                        if (kc <= sc.nz) cands.push_back({LBM_KERNEL_ONE_PASS, kc, tile, pf, 0});
                }
        const double scale = (double)h->cfg.nz / (double)sc.nz;
        for (const TuneChoice &cd : cands) {
            if (rc) break;
            rc = set_form(t, cd.variant, cd.kchunk, cd.tile, cd.pf);
            double ms = 0.0;
            if (!rc) rc = lbm_step(t, 2);                   // warm-up with this form
            if (!rc) rc = lbm_step_timed(t, steps, &ms);
            if (!rc && (!timed || ms / steps * scale < margin * best)) {
                best = ms / steps * scale;
                win = cd;
                timed = true;
            }
        }
        lbm_destroy(t);
    }
    if (rc) return fail(h, rc, "autotune: timing failed");
    if (!timed) return fail(h, LBM_ERR_NOMEM, "autotune: no room for a scratch domain");
    if (open) rc = convert_layout(h, win.aa != 0);
    if (!rc) rc = set_form(h, win.variant, win.kchunk, win.tile, win.pf);
    if (rc) return rc;
    if (ms_per_step) *ms_per_step = best;
    if (cache_path) {
        FILE *fp = fopen(cache_path, "a");
        if (fp) {
            fprintf(fp, "%s %d %d %d %d %.6f %d\n", key, win.variant, win.kchunk, win.tile, win.pf,
                    best, win.aa ? LBM_LAYOUT_AA : LBM_LAYOUT_TWO_ARRAY);
            fclose(fp);
        }
    }
    return LBM_OK;
}

extern "C" int lbm_autotune(lbm_handle *h, const char *cache_path, int32_t steps, double *ms_per_step)
{
    if (!h || steps <= 0) return LBM_ERR_ARG;
    if (h->in_batch)
        return fail(h, LBM_ERR_STATE, "kernel form of a batch domain: use lbm_batch_autotune");
    return autotune_impl(h, cache_path, steps, ms_per_step);
}

// Bring the macro planes up to date with the current state (phase 1) or leave the plain ones.
static int refresh_macros(lbm_handle *h)
{
    if (h->phase < 0) return fail(h, LBM_ERR_STATE, "no state: call lbm_init_* first");
    if (h->phase == 1 && h->first_thermal) {
        // after the prologue alone: the macro planes still hold the plain state; in place, the
        // prologue has overwritten the plain f
        if (h->aa) return fail(h, LBM_ERR_STATE, "in-place layout between the prologue and the first step");
        return LBM_OK;
    }
    if (h->phase == 1 && h->aa && h->lp) {
        LPArgs a;
        a.Fc = h->F[0];
        a.Gi = h->G[0];
        a.aux = (float *)h->aux[0];
        a.mac = h->mac;
        a.lay = h->fpar;
        HIPCHK(h, launch_lp(h->cfg.precision, LP_STATE, h->kp, h->bx, h->stream, a));
    } else if (h->phase == 1 && h->aa) {
        HIPCHK(h, launch_state(h->kp, h->bx, h->stream, true, h->F[0], h->G[0], h->aux[0], nullptr,
                               h->mac, nullptr, nullptr, 0, 0, h->fpar));
    } else if (h->phase == 1 && h->lp) {
        LPArgs a;
        a.Fc = h->F[h->fcur ^ 1];
        a.Gi = h->G[h->gcur];
        a.aux = h->aux_valid ? (float *)h->aux[0] : nullptr;
        a.Gprev = h->G[h->gcur ^ 1];
        a.mac = h->mac;
        HIPCHK(h, launch_lp(h->cfg.precision, LP_STATE, h->kp, h->bx, h->stream, a));
    } else if (h->phase == 1) {
        const bool d3 = h->cfg.dims == 3;
        const double *aux = d3 ? (h->aux_valid ? h->aux[0] : nullptr) : h->aux[h->fy ^ 1];
        HIPCHK(h, launch_state(h->kp, h->bx, h->stream, d3, h->F[h->fcur ^ 1], h->G[h->gcur], aux,
                               h->G[h->gcur ^ 1], h->mac, nullptr, nullptr, 0, 0));
    }
    return LBM_OK;
}

extern "C" int lbm_macros(lbm_handle *h, double *rho, double *u, double *v, double *w, double *T)
{
    if (!h) return LBM_ERR_ARG;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    int rc = refresh_macros(h);
    if (rc) return rc;
    double *dst[5] = {rho, u, v, w, T};
    for (int q = 0; q < 5; ++q)
        if (dst[q])
            HIPCHK(h, hipMemcpyAsync(dst[q], h->mac + q * h->kp.P, h->N * sizeof(double),
                                     hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return LBM_OK;
}

// Compensated host combination of per-block (sum, carry) pairs.
struct HostKS {
    double s = 0, c = 0;
    void add(double x)
    {
        const double t = s + x;
        if (fabs(s) >= fabs(x)) c += (s - t) + x;
        else c += (x - t) + s;
        s = t;
    }
    double val() const { return s + c; }
};

static void combine(const std::vector<double> &part, unsigned blocks, int nq, double *out)
{
    for (int q = 0; q < nq; ++q) {
        HostKS k;
        for (unsigned b = 0; b < blocks; ++b) {
            k.add(part[(b * nq + q) * 2]);
            k.add(part[(b * nq + q) * 2 + 1]);
        }
        out[q] = k.val();
    }
}

extern "C" int lbm_check(lbm_handle *h, double *error_u, double *error_t)
{
    if (!h || !error_u || !error_t) return LBM_ERR_ARG;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    if (!h->prev) {
        if (hipMalloc((void **)&h->prev, 4 * h->kp.P * sizeof(double)) != hipSuccess)
            return fail(h, LBM_ERR_NOMEM, "prev arrays");
        HIPCHK(h, hipMemsetAsync(h->prev, 0, 4 * h->kp.P * sizeof(double), h->stream));
    }
    int rc = refresh_macros(h);
    if (rc) return rc;
    const unsigned nb = h->red_blocks;
    HIPCHK(h, launch_reduce_check(h->kp, h->stream, h->mac, h->prev, h->red, nb));
    std::vector<double> part((size_t)nb * 4 * 2);
    HIPCHK(h, hipMemcpyAsync(part.data(), h->red, part.size() * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    double s[4];
    combine(part, nb, 4, s);
    if (h->cfg.dims == 3) {  // 2DRB/3DRB.F90:1512-1521
        *error_u = s[1] > 1.0e-30 ? sqrt(s[0]) / sqrt(s[1]) : sqrt(s[0]);
        *error_t = s[3] > 1.0e-30 ? s[2] / s[3] : s[2];
    } else {                 // 2DRB/2DRB.F90:1244-1245 (no guard)
        *error_u = sqrt(s[0]) / sqrt(s[1]);
        *error_t = s[2] / s[3];
    }
    return LBM_OK;
}

// fit_adiabatic_wall_T4, 2DRB/3DRB.F90:2212-2241
static double fit_wall_T4(double y0, const double *y, const double *t)
{
    double sq[4], S1 = 0, S2 = 0, B0 = 0, B1 = 0;
    const double S0 = 4.0;
    for (int k = 0; k < 4; ++k) sq[k] = (y[k] - y0) * (y[k] - y0);
    for (int k = 0; k < 4; ++k) {
        S1 += sq[k];
        S2 += sq[k] * sq[k];
        B0 += t[k];
        B1 += t[k] * sq[k];
    }
    return (B0 * S2 - B1 * S1) / (S0 * S2 - S1 * S1);
}

// fit_parabola_ls5, 2DRB/3DRB.F90:2252-2334
static void fit_parabola5(const double *y, const double *f, int mode, double *fs, double *ys)
{
    int kb = 0;
    for (int k = 1; k < 5; ++k)
        if ((mode == 1 && f[k] > f[kb]) || (mode != 1 && f[k] < f[kb])) kb = k;
    double S0 = 0, S1 = 0, S2 = 0, S3 = 0, S4 = 0, F0 = 0, F1 = 0, F2 = 0;
    for (int k = 0; k < 5; ++k) {
        S0 += 1.0;
        S1 += y[k];
        S2 += y[k] * y[k];
        S3 += y[k] * y[k] * y[k];
        S4 += y[k] * y[k] * y[k] * y[k];
        F0 += f[k];
        F1 += f[k] * y[k];
        F2 += f[k] * y[k] * y[k];
    }
    const double D = S0 * (S2 * S4 - S3 * S3) - S1 * (S1 * S4 - S2 * S3) + S2 * (S1 * S3 - S2 * S2);
    const double DA = F0 * (S2 * S4 - S3 * S3) - S1 * (F1 * S4 - S3 * F2) + S2 * (F1 * S3 - S2 * F2);
    const double DB = S0 * (F1 * S4 - S3 * F2) - F0 * (S1 * S4 - S2 * S3) + S2 * (S1 * F2 - F1 * S2);
    const double DC = S0 * (S2 * F2 - F1 * S3) - S1 * (S1 * F2 - F1 * S2) + F0 * (S1 * S3 - S2 * S2);
    *ys = y[kb];
    *fs = f[kb];
    if (fabs(D) <= 1.0e-20) return;
    const double A = DA / D, B = DB / D, C = DC / D;
    double ymin = y[0], ymax = y[0];
    for (int k = 1; k < 5; ++k) {
        ymin = y[k] < ymin ? y[k] : ymin;
        ymax = y[k] > ymax ? y[k] : ymax;
    }
    if (fabs(C) <= 1.0e-14 || (mode == 1 && C >= 0.0) || (mode == -1 && C <= 0.0)) return;
    const double yst = -B / (2.0 * C);
    if (yst < ymin || yst > ymax) return;
    *ys = yst;
    *fs = A + B * yst + C * yst * yst;
}

// Device part of the diagnostics: the macro planes brought up to date, then the volume sums and
// the wall sums and profiles into the partials buffer and h->wall (no host synchronisation; a
// HIP graph may capture it).
static int diag_device(lbm_handle *h)
{
    int rc = refresh_macros(h);
    if (rc) return rc;
    const lbm_config &c = h->cfg;
    const double L = h->p.L, dy = 1.0 / L, coef = L / h->p.diffusivity;
    // Rayleigh-Benard cell: hot wall y = 0, wall profile along x.  Side-heated cell: hot wall
    // x = 0, wall profile along y (SideHeatedcalc_Nu_wall_avg, 2DRB/3DRB.F90:1889-2038).
    const int side = c.side_heated ? 1 : 0;
    const int nx = side ? c.ny : c.nx;   // length of the hot-wall profile
    const unsigned nb = h->red_blocks;
    HIPCHK(h, launch_reduce_volume(h->kp, h->stream, h->mac, coef, dy, side, h->red, nb));
    HIPCHK(h, launch_reduce_walls(h->kp, h->stream, h->mac, coef, dy, side, h->wall, h->wall + nx,
                                  h->red + (size_t)nb * 6 * 2));
    return LBM_OK;
}

// Doubles of one diagnostics sample as diag_device leaves them: volume partials, wall
// partials, the two wall profiles.
static size_t diag_sample_doubles(const lbm_handle *h)
{
    const int nx = h->cfg.side_heated ? h->cfg.ny : h->cfg.nx;
    return (size_t)h->red_blocks * 6 * 2 + (size_t)nx * 4 * 2 + (size_t)2 * nx;
}

// Host part: compensated sums of the partials and the wall fits of one sample laid out as
// diag_sample_doubles describes.
static void diag_finish(const lbm_handle *h, const double *sample, lbm_diag *d)
{
    const lbm_config &c = h->cfg;
    const double L = h->p.L, dy = 1.0 / L;
    const double dT = c.thot - c.tcold, N = (double)h->N;
    const int side = c.side_heated ? 1 : 0;
    const int nx = side ? c.ny : c.nx;
    const unsigned nb = h->red_blocks;
    memset(d, 0, sizeof(*d));
    std::vector<double> part(sample, sample + (size_t)nb * 6 * 2);
    double s[6];
    combine(part, nb, 6, s);
    d->nu_vol = s[0] / N * L / h->p.diffusivity + 1.0;
    d->re_vol = sqrt(s[1] / N) * L / h->p.viscosity;
    d->nu_global = (s[2] / N) / dT;
    d->mass = s[3];
    d->heat = s[4];
    d->kinetic = s[5];

    const double *wpart = sample + (size_t)nb * 6 * 2;
    const double *prof = wpart + (size_t)nx * 4 * 2;
    HostKS hot, cold, mid;
    for (int i = 0; i < nx; ++i) {
        hot.add(prof[nx + i]);
        cold.add(wpart[(i * 4 + 2) * 2]);
        cold.add(wpart[(i * 4 + 2) * 2 + 1]);
        mid.add(wpart[(i * 4 + 3) * 2]);
        mid.add(wpart[(i * 4 + 3) * 2 + 1]);
    }
    d->nu_hot = hot.val() / (double)nx;
    d->nu_cold = cold.val() / ((double)nx * c.nz);
    d->nu_middle = mid.val() / ((double)nx * c.nz);

    // Corner extension and 5-point fits of the hot-wall profile (2DRB/3DRB.F90:2089-2150;
    // side-heated cell :1929-1989, along y).
    std::vector<double> xp(nx + 2), ext(nx + 2);
    xp[0] = 0.0;
    xp[nx + 1] = (double)nx / L;
    for (int i = 1; i <= nx; ++i) xp[i] = ((double)i - 0.5) / L;
    for (int i = 0; i < nx; ++i) ext[i + 1] = prof[nx + i];
    const double *tb = prof;
    ext[0] = (2.0 * (c.thot - fit_wall_T4(0.0, &xp[1], tb)) / dy) / dT;
    ext[nx + 1] = (2.0 * (c.thot - fit_wall_T4(xp[nx + 1], &xp[nx - 3], tb + nx - 4)) / dy) / dT;
    int imax = 0, imin = 0;
    for (int i = 1; i <= nx + 1; ++i) {
        if (ext[i] > ext[imax]) imax = i;
        if (ext[i] < ext[imin]) imin = i;
    }
    for (int pass = 0; pass < 2; ++pass) {
        const int ic = pass == 0 ? imax : imin;
        int i0 = ic <= 2 ? 0 : (ic >= nx - 1 ? nx - 3 : ic - 2);
        if (side && pass == 1) i0 = ic >= 4 ? ic - 4 : 0;   // the side-heated minimum's window
        double fs, ys;
        fit_parabola5(&xp[i0], &ext[i0], pass == 0 ? 1 : -1, &fs, &ys);
        if (pass == 0) { d->nu_hot_max = fs; d->nu_hot_max_pos = ys; }
        else { d->nu_hot_min = fs; d->nu_hot_min_pos = ys; }
    }
}

extern "C" int lbm_diagnostics(lbm_handle *h, lbm_diag *d)
{
    if (!h || !d) return LBM_ERR_ARG;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    int rc = diag_device(h);
    if (rc) return rc;
    const int nx = h->cfg.side_heated ? h->cfg.ny : h->cfg.nx;
    const size_t nred = (size_t)h->red_blocks * 6 * 2 + (size_t)nx * 4 * 2;
    std::vector<double> sample(diag_sample_doubles(h));
    HIPCHK(h, hipMemcpyAsync(sample.data(), h->red, nred * sizeof(double), hipMemcpyDeviceToHost,
                             h->stream));
    HIPCHK(h, hipMemcpyAsync(sample.data() + nred, h->wall, (size_t)2 * nx * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    diag_finish(h, sample.data(), d);
    return LBM_OK;
}

// ---- runs with periodic diagnostics in a HIP graph ------------------------------------------
// A device series of diagnostics samples (the partials and wall profiles of diag_device).
struct DiagSeries {
    double *buf = nullptr;
    int *count = nullptr;
    ~DiagSeries()
    {
        if (buf) (void)hipFree(buf);
        if (count) (void)hipFree(count);
    }
};

// Appends the diagnostics of domain h's current state to slot count[0] of series.
static int diag_append(lbm_handle *h, double *series, int cap, int *count)
{
    int rc = diag_device(h);
    if (rc) return rc;
    const int nx = h->cfg.side_heated ? h->cfg.ny : h->cfg.nx;
    const size_t nred = (size_t)h->red_blocks * 6 * 2 + (size_t)nx * 4 * 2;
    HIPCHK(h, launch_series_append(h->stream, h->red, nred, h->wall, (size_t)2 * nx, series,
                                   diag_sample_doubles(h), cap, count));
    return LBM_OK;
}

// The steps of one period return every buffer to its place when the period is even (two-array
// buffers, in-place states, the one-pass history buffers and the 2D force planes all alternate
// from step to step).
static int graph_periods(int every) { return every % 2 == 0 ? 1 : 2; }

extern "C" int lbm_run_diagnostics(lbm_handle *h, int32_t nsteps, int32_t every, lbm_diag *out)
{
    if (!h || !out || every <= 0 || nsteps < every || nsteps % every) return LBM_ERR_ARG;
    if (h->phase < 0) return fail(h, LBM_ERR_STATE, "lbm_run_diagnostics before initialisation");
    if (h->in_batch) return fail(h, LBM_ERR_STATE, "batch domain: use lbm_batch_run_diagnostics");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    const int ns = nsteps / every;
    const size_t stride = diag_sample_doubles(h);
    DiagSeries ser;
    if (hipMalloc((void **)&ser.buf, (size_t)ns * stride * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&ser.count, sizeof(int)) != hipSuccess)
        return fail(h, LBM_ERR_NOMEM, "diagnostics series");
    HIPCHK(h, launch_series_reset(h->stream, ser.count, 1));
    auto period = [&]() -> int {
        int rc = enqueue_steps(h, every);
        return rc ? rc : diag_append(h, ser.buf, ns, ser.count);
    };
    // the first period eagerly: it leaves the plain state and settles the buffer parities (and
    // a second one when that first period was a single two-pass step before one-pass steps)
    int rc = period();
    if (rc) return rc;
    int done = 1;
    if (!capture_ready(h) && done < ns) {
        rc = period();
        if (rc) return rc;
        ++done;
    }
    const int P = graph_periods(every);
    const int replays = (ns - done) / P;
    if (replays > 0) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        const int64_t itc0 = h->itc;
        HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
        h->capturing = true;
        for (int p = 0; p < P && !rc; ++p) rc = period();
        h->capturing = false;
        hipError_t e = hipStreamEndCapture(h->stream, &g);
        if (!rc && e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        // host bookkeeping now stands one replay ahead (buffers back in place); the step counter
        // follows the replays that were launched
        int launched = 0;
        for (; launched < replays && !rc && e == hipSuccess; ++launched) {
            e = hipGraphLaunch(ge, h->stream);
            if (e != hipSuccess) break;
        }
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        h->itc = itc0 + (int64_t)launched * P * every;
        if (rc) return rc;
        if (e != hipSuccess) return fail(h, LBM_ERR_HIP, "diagnostics graph: %s", hipGetErrorString(e));
        done += replays * P;
    }
    for (; done < ns; ++done) {
        rc = period();
        if (rc) return rc;
    }
    std::vector<double> host((size_t)ns * stride);
    HIPCHK(h, hipMemcpyAsync(host.data(), ser.buf, host.size() * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < ns; ++k) diag_finish(h, host.data() + (size_t)k * stride, &out[k]);
    return LBM_OK;
}

// ---- periodic diagnostics without allocation -----------------------------------------------
// Device scratch of lbm_run_diagnostics_noalloc: the series counter in the first 256 bytes, the
// series after it.  Every GPU the library runs on holds its own copy of the array until the
// process ends; host code reaches a device's copy only through the address hipGetSymbolAddress
// returns for it with that device current (scratch_dev), never through the array's name.
constexpr size_t kScratchBytes = (size_t)32 << 20;
constexpr size_t kScratchHead = 256;
constexpr int kScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];
static std::mutex scratch_mu;
static unsigned char *scratch_dev[kScratchDevices];

// Address of device h->cfg.device's scratch (looked up once per device, that device current).
static int scratch_setup(lbm_handle *h)
{
    const int d = h->cfg.device;
    if (d < 0 || d >= kScratchDevices) return fail(h, LBM_ERR_ARG, "device %d beyond the scratch table", d);
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_dev[d]) return LBM_OK;
    HIPCHK(h, hipSetDevice(d));
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return fail(h, LBM_ERR_HIP, "device scratch address: %s", hipGetErrorString(e));
    scratch_dev[d] = (unsigned char *)p;
    return LBM_OK;
}

// lbm_run_diagnostics with its series in the device scratch: every period runs eagerly (steps,
// then the device part of the diagnostics appending to the scratch series), the host finishes
// the samples after one copy of the series.  No device memory is allocated or freed.
extern "C" int lbm_run_diagnostics_noalloc(lbm_handle *h, int32_t nsteps, int32_t every, lbm_diag *out)
{
    if (!h || !out || every <= 0 || nsteps < every || nsteps % every) return LBM_ERR_ARG;
    if (h->phase < 0) return fail(h, LBM_ERR_STATE, "lbm_run_diagnostics_noalloc before initialisation");
    if (h->in_batch) return refuse_batch_domain(h, "lbm_run_diagnostics_noalloc");
    int rc = scratch_setup(h);
    if (rc) return rc;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    const int ns = nsteps / every;
    const size_t stride = diag_sample_doubles(h);
    const size_t need = kScratchHead + (size_t)ns * stride * sizeof(double);
    if (need > kScratchBytes)
        return fail(h, LBM_ERR_ARG,
                    "lbm_run_diagnostics_noalloc: %d samples of %zu bytes need %zu bytes of scratch, "
                    "which holds %zu (at most %zu samples for this domain)",
                    ns, stride * sizeof(double), need, kScratchBytes,
                    (kScratchBytes - kScratchHead) / (stride * sizeof(double)));
    unsigned char *base = scratch_dev[h->cfg.device];
    int *count = (int *)base;
    double *series = (double *)(base + kScratchHead);
    HIPCHK(h, launch_series_reset(h->stream, count, 1));   // the last call left its count
    for (int k = 0; k < ns; ++k) {
        rc = enqueue_steps(h, every);
        if (!rc) rc = diag_append(h, series, ns, count);
        if (rc) return rc;
    }
    std::vector<double> host((size_t)ns * stride);
    HIPCHK(h, hipMemcpyAsync(host.data(), series, host.size() * sizeof(double),
                             hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int k = 0; k < ns; ++k) diag_finish(h, host.data() + (size_t)k * stride, &out[k]);
    return LBM_OK;
}

// ---- mid-plane velocity maxima --------------------------------------------------------------
// find_bracketing_index (3DRB.F90:2525-2547) on 1-based coordinates coord[0..n+1].
static void bracket_index(const std::vector<double> &coord, int n, double target, int *iL, int *iR,
                          double *weight)
{
    int l = 1;
    while (l < n && coord[l + 1] < target) ++l;
    const int r = l + 1 < n ? l + 1 : n;
    double w;
    if (fabs(coord[r] - coord[l]) <= 1.0e-14) w = 0.0;
    else w = (target - coord[l]) / (coord[r] - coord[l]);
    *iL = l;
    *iR = r;
    *weight = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
}

// The 2D routines' 5-point least-squares parabola around the grid maximum of line[1..n].
static void line_max_fit(const std::vector<double> &coord, const std::vector<double> &line, int n,
                         double *fmax, double *pos)
{
    int i0 = 1;
    for (int i = 2; i <= n; ++i)
        if (line[i] > line[i0]) i0 = i;
    const int a = i0 <= 2 ? 1 : (i0 >= n - 1 ? n - 4 : i0 - 2);
    fit_parabola5(&coord[a], &line[a], 1, fmax, pos);
}

extern "C" int lbm_midplane_extrema(lbm_handle *h, lbm_midplane *m)
{
    if (!h || !m) return LBM_ERR_ARG;
    HIPCHK(h, hipSetDevice(h->cfg.device));
    int rc = refresh_macros(h);
    if (rc) return rc;
    const lbm_config &c = h->cfg;
    if (c.nx < 5 || c.ny < 5) return fail(h, LBM_ERR_ARG, "mid-plane fits need nx, ny >= 5");
    const double L = h->p.L, coef = L / h->p.diffusivity;
    const int n[3] = {c.nx, c.ny, c.dims == 3 ? c.nz : 1};
    std::vector<double> coord[3];
    for (int a = 0; a < 3; ++a) {   // node coordinates / L, walls at 0 and n / L
        coord[a].assign(n[a] + 2, 0.0);
        coord[a][n[a] + 1] = (double)n[a] / L;
        for (int i = 1; i <= n[a]; ++i) coord[a][i] = ((double)i - 0.5) / L;
    }
    const size_t np = (size_t)std::max(n[1] * n[2], std::max(n[0] * n[2], n[0] * n[1]));
    double *dplane = nullptr;
    if (hipMalloc((void **)&dplane, np * sizeof(double)) != hipSuccess)
        return fail(h, LBM_ERR_NOMEM, "mid-plane buffer");
    std::vector<double> plane(np);
    // one plane of component `comp` (1 u, 2 v, 3 w) across `axis` -> host vector
    auto extract = [&](int comp, int axis, int l, int r, double w, int mode, size_t cnt) -> int {
        hipError_t e = launch_plane(h->kp, h->stream, h->mac + (size_t)comp * h->kp.P, axis, l - 1,
                                    r - 1, w, mode, dplane);
        if (e == hipSuccess)
            e = hipMemcpyAsync(plane.data(), dplane, cnt * sizeof(double), hipMemcpyDeviceToHost,
                               h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        return e == hipSuccess ? LBM_OK : fail(h, LBM_ERR_HIP, "mid-plane: %s", hipGetErrorString(e));
    };
    memset(m, 0, sizeof(*m));
    if (c.dims == 3) {
        // calc_{u,v,w}mid_max_common (3DRB.F90:2399-2518): the first grid maximum in loop order
        double *out[3] = {&m->u_max, &m->v_max, &m->w_max};
        for (int a = 0; a < 3 && !rc; ++a) {
            const int pa = a == 0 ? 1 : 0, qa = a == 2 ? 1 : 2;
            const double target = 0.5 * coord[a][n[a] + 1];
            int iL, iR;
            double w;
            bracket_index(coord[a], n[a], target, &iL, &iR, &w);
            const size_t cnt = (size_t)n[pa] * n[qa];
            rc = extract(a + 1, a, iL, iR, w, 0, cnt);
            if (rc) break;
            size_t best = 0;
            for (size_t t = 1; t < cnt; ++t)
                if (plane[t] > plane[best]) best = t;
            out[a][0] = plane[best] * coef;
            out[a][1] = coord[pa][1 + best % n[pa]];
            out[a][2] = coord[qa][1 + best / n[pa]];
            out[a][3] = target;
        }
    } else {
        const int nx = n[0], ny = n[1];
        std::vector<double> line;
        double fmax = 0.0, pos = 0.0;
        // u along y: the middle column (side-heated, 2DRB.F90:2081-2153) or x = 0.5
        // (Rayleigh-Benard, 2DRB.F90:2484-2564, position folded into the lower half)
        if (c.side_heated) {
            if (nx % 2 == 1) {
                rc = extract(1, 0, (nx + 1) / 2, (nx + 1) / 2, 0.0, 0, ny);
                m->u_plane = coord[0][(nx + 1) / 2];
            } else {
                rc = extract(1, 0, nx / 2, nx / 2 + 1, 0.0, 1, ny);
                m->u_plane = 0.5 * (coord[0][nx / 2] + coord[0][nx / 2 + 1]);
            }
        } else {
            const double target = 0.5;
            int il = 1;
            while (il < nx && coord[0][il + 1] < target) ++il;
            const int ir = il + 1 < nx ? il + 1 : nx;
            if (fabs(coord[0][il] - target) <= 1.0e-14 || ir == il)
                rc = extract(1, 0, il, il, 0.0, 0, ny);
            else
                rc = extract(1, 0, il, ir, (target - coord[0][il]) / (coord[0][ir] - coord[0][il]), 0, ny);
            m->u_plane = target;
        }
        if (!rc) {
            line.assign(plane.begin(), plane.begin() + ny);
            line.insert(line.begin(), 0.0);
            line_max_fit(coord[1], line, ny, &fmax, &pos);
            if (!c.side_heated && pos > 0.5 * coord[1][ny + 1]) pos = coord[1][ny + 1] - pos;
            m->u_max = fmax * coef;
            m->u_pos1 = pos;
            // v along x on the middle row (2DRB.F90:2575-2642, :2166-2238)
            if (ny % 2 == 1) {
                rc = extract(2, 1, (ny + 1) / 2, (ny + 1) / 2, 0.0, 0, nx);
                m->v_plane = coord[1][(ny + 1) / 2];
            } else {
                rc = extract(2, 1, ny / 2, ny / 2 + 1, 0.0, 1, nx);
                m->v_plane = 0.5 * (coord[1][ny / 2] + coord[1][ny / 2 + 1]);
            }
        }
        if (!rc) {
            line.assign(plane.begin(), plane.begin() + nx);
            line.insert(line.begin(), 0.0);
            line_max_fit(coord[0], line, nx, &fmax, &pos);
            m->v_max = fmax * coef;
            m->v_pos1 = pos;
        }
    }
    (void)hipFree(dplane);
    return rc;
}

// ---- mid-plane slices: fields, stream function, vorticity (3D) ------------------------------
// d/ds of a line q[0], q[st], ..., q[(n-1) st] with spacing h: central inside, second-order
// one-sided at the ends, first order for n = 2, zero for n = 1 (3DRB.F90:2776-2805).
static double line_derivative(const double *q, int n, ptrdiff_t st, int t, double h)
{
    if (n == 1) return 0.0;
    if (t == 0) {
        if (n == 2) return (q[st] - q[0]) / h;
        return (-3.0 * q[0] + 4.0 * q[st] - q[2 * st]) / (2.0 * h);
    }
    if (t == n - 1) {
        if (n == 2) return (q[t * st] - q[(t - 1) * st]) / h;
        return (3.0 * q[t * st] - 4.0 * q[(t - 1) * st] + q[(t - 2) * st]) / (2.0 * h);
    }
    return (q[(t + 1) * st] - q[(t - 1) * st]) / (2.0 * h);
}

extern "C" int lbm_midplane_slice(lbm_handle *h, int32_t axis, double *us, double *vs, double *ws,
                                  double *ts, double *psi, double *vort)
{
    if (!h || axis < 0 || axis > 2) return LBM_ERR_ARG;
    if (h->cfg.dims != 3) return fail(h, LBM_ERR_UNSUPPORTED, "mid-plane slices are 3D");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    int rc = refresh_macros(h);
    if (rc) return rc;
    const lbm_config &c = h->cfg;
    const double L = h->p.L, coef = L / h->p.diffusivity;
    const int n[3] = {c.nx, c.ny, c.nz};
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    const int n1 = n[a1], n2 = n[a2];
    const size_t cnt = (size_t)n1 * n2;
    std::vector<double> coord(n[axis] + 2, 0.0);
    coord[n[axis] + 1] = (double)n[axis] / L;
    for (int i = 1; i <= n[axis]; ++i) coord[i] = ((double)i - 0.5) / L;
    int iL, iR;
    double wgt;
    bracket_index(coord, n[axis], 0.5 * coord[n[axis] + 1], &iL, &iR, &wgt);

    // write_midplane_x/y/z (3DRB.F90:2611-2713): components scaled by L / kappa, T as is
    std::vector<double> slice[5];
    double *dplane = nullptr;
    if (hipMalloc((void **)&dplane, cnt * sizeof(double)) != hipSuccess)
        return fail(h, LBM_ERR_NOMEM, "mid-plane buffer");
    for (int comp = 1; comp <= 4 && !rc; ++comp) {
        slice[comp].resize(cnt);
        hipError_t e = launch_plane(h->kp, h->stream, h->mac + (size_t)comp * h->kp.P, axis, iL - 1,
                                    iR - 1, wgt, 0, dplane);
        if (e == hipSuccess)
            e = hipMemcpyAsync(slice[comp].data(), dplane, cnt * sizeof(double),
                               hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = fail(h, LBM_ERR_HIP, "mid-plane slice: %s", hipGetErrorString(e));
    }
    (void)hipFree(dplane);
    if (rc) return rc;
    double *outs[4] = {us, vs, ws, ts};
    for (int comp = 1; comp <= 4; ++comp)
        if (outs[comp - 1])
            for (size_t t = 0; t < cnt; ++t)
                outs[comp - 1][t] = comp == 4 ? slice[4][t] : coef * slice[comp][t];

    // write_midplane_stream_x/y/z (3DRB.F90:2721-3011): psi integrates the component along the
    // first in-plane axis over the second one from the wall; the vorticity is the plane-normal
    // component (sign of the right-handed axis triple)
    const std::vector<double> &c1 = slice[1 + a1], &c2 = slice[1 + a2];
    const double d = 1.0 / L;
    if (psi) {
        for (int p1 = 0; p1 < n1; ++p1) {
            const double *q = c1.data() + p1;   // stride n1 along the second axis
            double *o = psi + p1;
            if (n2 == 1) { o[0] = 0.0; continue; }
            const double q1 = q[0] * coef, q2 = q[n1] * coef;
            o[0] = d * (21.0 * q1 - q2) / 72.0;
            for (int k = 2; k <= n2; ++k) {
                const int m = k - 1;   // 1-based midpoint between m and m + 1
                double mid;
                if (n2 == 2) mid = 0.5 * (q[0] + q[n1]) * coef;
                else if (m == 1)
                    mid = (3.0 / 8.0 * q[0] + 3.0 / 4.0 * q[n1] - 1.0 / 8.0 * q[2 * n1]) * coef;
                else
                    mid = (-1.0 / 8.0 * q[(m - 2) * n1] + 3.0 / 4.0 * q[(m - 1) * n1] +
                           3.0 / 8.0 * q[m * n1]) * coef;
                const double inc = d / 6.0 * (q[(k - 2) * n1] * coef + 4.0 * mid + q[(k - 1) * n1] * coef);
                o[(k - 1) * n1] = o[(k - 2) * n1] + inc;
            }
        }
    }
    if (vort) {
        for (int p2 = 0; p2 < n2; ++p2)
            for (int p1 = 0; p1 < n1; ++p1) {
                const double dc2 = line_derivative(c2.data() + (size_t)p2 * n1, n1, 1, p1, d);
                const double dc1 = line_derivative(c1.data() + p1, n2, n1, p2, d);
                vort[p1 + (size_t)n1 * p2] = axis == 1 ? coef * (dc1 - dc2) : coef * (dc2 - dc1);
            }
    }
    return LBM_OK;
}

// ---- 2D stream function, vorticity and |psi| statistics --------------------------------------
// natural cubic spline (spline_natural, 2DRB/2DRB.F90:3000-3031) and its evaluation (splint)
static void spline_natural(int n, const double *x, const double *y, double *y2, double *u)
{
    y2[0] = 0.0;
    u[0] = 0.0;
    for (int i = 1; i < n - 1; ++i) {
        const double sig = (x[i] - x[i - 1]) / (x[i + 1] - x[i - 1]);
        const double p = sig * y2[i - 1] + 2.0;
        y2[i] = (sig - 1.0) / p;
        u[i] = (6.0 * ((y[i + 1] - y[i]) / (x[i + 1] - x[i]) - (y[i] - y[i - 1]) / (x[i] - x[i - 1])) /
                    (x[i + 1] - x[i - 1]) - sig * u[i - 1]) / p;
    }
    y2[n - 1] = 0.0;
    for (int k = n - 2; k >= 0; --k) y2[k] = y2[k] * y2[k + 1] + u[k];
}

static double splint_host(int n, const double *xa, const double *ya, const double *y2a, double x)
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

static double lagrange4(double xq, const double *xk, const double *fk)
{
    double fq = 0.0;
    for (int a = 0; a < 4; ++a) {
        double basis = 1.0;
        for (int b = 0; b < 4; ++b)
            if (b != a) basis = basis * (xq - xk[b]) / (xk[a] - xk[b]);
        fq = fq + fk[a] * basis;
    }
    return fq;
}

extern "C" int lbm_psi_vorticity_2d(lbm_handle *h, double *psi_out, double *vort_out, lbm_psi *st)
{
    if (!h) return LBM_ERR_ARG;
    const lbm_config &c = h->cfg;
    if (c.dims != 2) return fail(h, LBM_ERR_UNSUPPORTED, "stream function: 2D only");
    if (c.nx < 3 || c.ny < 3) return fail(h, LBM_ERR_ARG, "stream function needs nx, ny >= 3");
    HIPCHK(h, hipSetDevice(c.device));
    int rc = refresh_macros(h);
    if (rc) return rc;
    const int nx = c.nx, ny = c.ny;
    const size_t N = (size_t)nx * ny;
    std::vector<double> uu(N), vv(N), psi(N);
    HIPCHK(h, hipMemcpyAsync(uu.data(), h->mac + h->kp.P, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipMemcpyAsync(vv.data(), h->mac + 2 * h->kp.P, N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const double L = h->p.L, dx = 1.0 / L, dy = 1.0 / L, coef = L / h->p.diffusivity;
    auto U = [&](int i, int j) { return uu[(i - 1) + (size_t)nx * (j - 1)]; };
    auto V = [&](int i, int j) { return vv[(i - 1) + (size_t)nx * (j - 1)]; };
    auto P = [&](int i, int j) -> double & { return psi[(i - 1) + (size_t)nx * (j - 1)]; };
    // calc_psi_vort_and_output (2DRB/2DRB.F90:2657-2770): psi integrates u along y from the wall
    for (int i = 1; i <= nx; ++i) {
        const double u1 = U(i, 1) * coef, u2 = U(i, 2) * coef;
        P(i, 1) = dy * (21.0 * u1 - u2) / 72.0;
        for (int j = 2; j <= ny; ++j) {
            const int m = j - 1;
            double um;
            if (m == 1) um = (3.0 / 8.0 * U(i, 1) + 3.0 / 4.0 * U(i, 2) - 1.0 / 8.0 * U(i, 3)) * coef;
            else um = (-1.0 / 8.0 * U(i, m - 1) + 3.0 / 4.0 * U(i, m) + 3.0 / 8.0 * U(i, m + 1)) * coef;
            const double inc = dy / 6.0 * ((U(i, j - 1) * coef) + 4.0 * um + (U(i, j) * coef));
            P(i, j) = P(i, j - 1) + inc;
        }
    }
    if (psi_out) memcpy(psi_out, psi.data(), N * sizeof(double));
    if (vort_out)   // walls at rest: one-sided differences through the wall value 0
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double dvdx, dudy;
                if (i == 1) dvdx = (3.0 * V(1, j) + V(2, j) - 4.0 * 0.0) / (3.0 * dx);
                else if (i == nx) dvdx = (-3.0 * V(nx, j) - V(nx - 1, j) + 4.0 * 0.0) / (3.0 * dx);
                else dvdx = (V(i + 1, j) - V(i - 1, j)) / (2.0 * dx);
                if (j == 1) dudy = (3.0 * U(i, 1) + U(i, 2) - 4.0 * 0.0) / (3.0 * dy);
                else if (j == ny) dudy = (-3.0 * U(i, ny) - U(i, ny - 1) + 4.0 * 0.0) / (3.0 * dy);
                else dudy = (U(i, j + 1) - U(i, j - 1)) / (2.0 * dy);
                vort_out[(i - 1) + (size_t)nx * (j - 1)] = coef * (dvdx - dudy);
            }
    if (!st) return LBM_OK;
    memset(st, 0, sizeof(*st));
    std::vector<double> xp(nx + 2), yp(ny + 2);
    xp[nx + 1] = (double)nx / L;
    yp[ny + 1] = (double)ny / L;
    for (int i = 1; i <= nx; ++i) xp[i] = ((double)i - 0.5) / L;
    for (int j = 1; j <= ny; ++j) yp[j] = ((double)j - 0.5) / L;
    // output_psi_center_abs (:3085-3160): 4-point Lagrange interpolation, x then y
    const double x0 = 0.5 * xp[nx + 1], y0 = 0.5 * yp[ny + 1];
    double pc;
    if (nx < 4 || ny < 4) {
        pc = P((nx + 1) / 2, (ny + 1) / 2);
    } else {
        int i0 = 1;
        while (i0 < nx && xp[i0 + 1] <= x0) ++i0;
        i0 = std::max(1, std::min(i0 - 1, nx - 3));
        int j0 = 1;
        while (j0 < ny && yp[j0 + 1] <= y0) ++j0;
        j0 = std::max(1, std::min(j0 - 1, ny - 3));
        double x4[4], y4[4], f4[4], gx[4];
        for (int q = 0; q < 4; ++q) { x4[q] = xp[i0 + q]; y4[q] = yp[j0 + q]; }
        for (int q = 0; q < 4; ++q) {
            for (int pp = 0; pp < 4; ++pp) f4[pp] = P(i0 + pp, j0 + q);
            gx[q] = lagrange4(x0, x4, f4);
        }
        pc = lagrange4(y0, y4, gx);
    }
    st->psi_center_coarse = fabs(pc);
    // calc_psi_absmax_fine_spline (:2873-2989): row splines on the host, the fine grid on the GPU
    const double xlen = xp[nx + 1], ylen = yp[ny + 1];
    const int nfx = (int)std::max(2.0, (double)lround(xlen * 10000.0) + 1.0);
    const int nfy = (int)std::max(2.0, (double)lround(ylen * 10000.0) + 1.0);
    const int nxe = nx + 2, nye = ny + 2;
    std::vector<double> xe(nxe), ye(nye), rows((size_t)ny * nxe), y2r((size_t)ny * nxe), wk(std::max(nxe, nye));
    xe[0] = 0.0; xe[nxe - 1] = xp[nx + 1];
    for (int i = 1; i <= nx; ++i) xe[i] = xp[i];
    ye[0] = 0.0; ye[nye - 1] = yp[ny + 1];
    for (int j = 1; j <= ny; ++j) ye[j] = yp[j];
    std::vector<double> pcx(ny), colc(nye), y2c(nye);
    for (int j = 1; j <= ny; ++j) {
        double *row = &rows[(size_t)(j - 1) * nxe], *d2 = &y2r[(size_t)(j - 1) * nxe];
        row[0] = 0.0;
        row[nxe - 1] = 0.0;
        for (int i = 1; i <= nx; ++i) row[i] = P(i, j);
        spline_natural(nxe, xe.data(), row, d2, wk.data());
        pcx[j - 1] = splint_host(nxe, xe.data(), row, d2, 0.5 * xlen);
    }
    colc[0] = 0.0;
    colc[nye - 1] = 0.0;
    for (int j = 1; j <= ny; ++j) colc[j] = pcx[j - 1];
    spline_natural(nye, ye.data(), colc.data(), y2c.data(), wk.data());
    st->psi_center_fine = fabs(splint_host(nye, ye.data(), colc.data(), y2c.data(), 0.5 * ylen));

    const unsigned blocks = 2048;
    double *d = nullptr;
    const size_t ncol = (size_t)nfx * nye;
    const size_t nd = nxe + nye + 2 * (size_t)ny * nxe + 3 * ncol + blocks;
    if (hipMalloc((void **)&d, nd * sizeof(double) + blocks * sizeof(unsigned long long)) != hipSuccess)
        return fail(h, LBM_ERR_NOMEM, "stream-function buffers");
    double *dxe = d, *dye = dxe + nxe, *drows = dye + nye, *dy2r = drows + (size_t)ny * nxe;
    double *dcols = dy2r + (size_t)ny * nxe, *dy2c = dcols + ncol, *dwk = dy2c + ncol;
    double *dbv = dwk + ncol;
    unsigned long long *dbi = (unsigned long long *)(dbv + blocks);
    std::vector<double> bv(blocks);
    std::vector<unsigned long long> bi(blocks);
    hipError_t e = hipMemcpyAsync(dxe, xe.data(), nxe * sizeof(double), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(dye, ye.data(), nye * sizeof(double), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(drows, rows.data(), rows.size() * sizeof(double), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(dy2r, y2r.data(), y2r.size() * sizeof(double), hipMemcpyHostToDevice, h->stream);
    if (e == hipSuccess)
        e = launch_psi_absmax(h->stream, nx, ny, nfx, nfy, xlen, ylen, dxe, dye, drows, dy2r, dcols,
                              dy2c, dwk, dbv, dbi, blocks);
    if (e == hipSuccess)
        e = hipMemcpyAsync(bv.data(), dbv, blocks * sizeof(double), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(bi.data(), dbi, blocks * sizeof(unsigned long long), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(d);
    if (e != hipSuccess) return fail(h, LBM_ERR_HIP, "stream function: %s", hipGetErrorString(e));
    double best = -1.0;
    unsigned long long idx = ~0ULL;
    for (unsigned b = 0; b < blocks; ++b)
        if (bv[b] > best || (bv[b] == best && bi[b] < idx)) {
            best = bv[b];
            idx = bi[b];
        }
    st->psi_abs_max = best;
    st->x_at_max = xlen * (double)(idx / (unsigned long long)nfy) / (double)(nfx - 1);
    st->y_at_max = ylen * (double)(idx % (unsigned long long)nfy) / (double)(nfy - 1);
    return LBM_OK;
}

// Chunk of z-planes whose AoS f and g fit a staging buffer of about 256 MiB.
static int chunk_planes(const lbm_handle *h)
{
    const size_t plane = (size_t)h->cfg.nx * h->cfg.ny * (h->qf + h->qt) * sizeof(double);
    size_t n = ((size_t)256 << 20) / (plane ? plane : 1);
    if (n < 1) n = 1;
    if (n > (size_t)h->cfg.nz) n = h->cfg.nz;
    return (int)n;
}

extern "C" int lbm_gather(lbm_handle *h, double *f, double *g, double *bprev)
{
    if (!h) return LBM_ERR_ARG;
    if (h->phase < 0) return fail(h, LBM_ERR_STATE, "no state: call lbm_init_* first");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    const bool d3 = h->cfg.dims == 3;
    const int nk = chunk_planes(h);
    const size_t plane = (size_t)h->cfg.nx * h->cfg.ny;
    // after the prologue alone the plain arrays still hold the state (two arrays) or the
    // prologue has overwritten the plain f (in place)
    const bool plain = h->phase == 0 || (h->phase == 1 && h->first_thermal);
    if (plain && h->phase == 1 && h->aa && (f || g))
        return fail(h, LBM_ERR_STATE, "in-place layout between the prologue and the first step");
    double *stage = nullptr;
    if (f || g) {
        if (hipMalloc((void **)&stage, (size_t)nk * plane * (h->qf + h->qt) * sizeof(double)) != hipSuccess)
            return fail(h, LBM_ERR_NOMEM, "gather staging");
    }
    double *fo = stage, *go = stage ? stage + (size_t)nk * plane * h->qf : nullptr;
    int rc = LBM_OK;
    for (int k0 = 0; stage && k0 < h->cfg.nz && rc == LBM_OK; k0 += nk) {
        const int n = (k0 + nk <= h->cfg.nz) ? nk : h->cfg.nz - k0;
        const uint64_t c0 = (uint64_t)k0 * plane, cnt = (uint64_t)n * plane;
        hipError_t e;
        if (h->lp) {
            LPArgs a;
            a.fo = fo;
            a.go = go;
            if (plain) {
                a.Fc = h->F[0];
                a.Gi = h->G[0];
                a.c0 = c0;
                a.n = cnt;
                e = launch_lp(h->cfg.precision, LP_TO_ABS, h->kp, h->bx, h->stream, a);
            } else if (h->aa) {
                a.Fc = h->F[0];
                a.Gi = h->G[0];
                a.aux = (float *)h->aux[0];
                a.lay = h->fpar;
                a.k0 = k0;
                a.nk = n;
                e = launch_lp(h->cfg.precision, LP_STATE, h->kp, h->bx, h->stream, a);
            } else {
                a.Fc = h->F[h->fcur ^ 1];
                a.Gi = h->G[h->gcur];
                a.aux = h->aux_valid ? (float *)h->aux[0] : nullptr;
                a.Gprev = h->G[h->gcur ^ 1];
                a.k0 = k0;
                a.nk = n;
                e = launch_lp(h->cfg.precision, LP_STATE, h->kp, h->bx, h->stream, a);
            }
        } else if (plain) {
            e = launch_transpose(false, h->stream, h->F[0], fo, h->qf, h->kp.P, c0, cnt);
            if (e == hipSuccess) e = launch_transpose(false, h->stream, h->G[0], go, h->qt, h->kp.P, c0, cnt);
        } else if (h->aa) {
            e = launch_state(h->kp, h->bx, h->stream, true, h->F[0], h->G[0], h->aux[0], nullptr,
                             nullptr, fo, go, k0, n, h->fpar);
        } else {
            const double *aux = d3 ? (h->aux_valid ? h->aux[0] : nullptr) : h->aux[h->fy ^ 1];
            e = launch_state(h->kp, h->bx, h->stream, d3, h->F[h->fcur ^ 1], h->G[h->gcur], aux,
                             h->G[h->gcur ^ 1], nullptr, fo, go, k0, n);
        }
        if (e == hipSuccess && f)
            e = hipMemcpyAsync(f + c0 * h->qf, fo, cnt * h->qf * sizeof(double), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess && g)
            e = hipMemcpyAsync(g + c0 * h->qt, go, cnt * h->qt * sizeof(double), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = fail(h, LBM_ERR_HIP, "gather: %s", hipGetErrorString(e));
    }
    if (stage) (void)hipFree(stage);
    if (rc) return rc;
    if (bprev && !h->lp) {
        for (int q = 0; q < 3; ++q)   // planes at the pitch, packed on the host
            HIPCHK(h, hipMemcpyAsync(bprev + q * h->N, h->Bp + q * h->kp.P, h->N * sizeof(double),
                                     hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
    } else if (bprev) {
        double *tmp = nullptr;
        if (hipMalloc((void **)&tmp, 3 * h->N * sizeof(double)) != hipSuccess)
            return fail(h, LBM_ERR_NOMEM, "gather B_prev staging");
        hipError_t e = hipSuccess;
        for (int q = 0; q < 3 && e == hipSuccess; ++q) {
            LPArgs a;
            a.Bp = (float *)h->Bp + q * h->kp.P;
            a.fo = tmp + q * h->N;
            a.n = h->N;
            e = launch_lp(h->cfg.precision, LP_BPREV, h->kp, h->bx, h->stream, a);
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(bprev, tmp, 3 * h->N * sizeof(double), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        (void)hipFree(tmp);
        if (e != hipSuccess) return fail(h, LBM_ERR_HIP, "gather B_prev: %s", hipGetErrorString(e));
    }
    return LBM_OK;
}

extern "C" int lbm_scatter(lbm_handle *h, const double *f, const double *g)
{
    if (!h || !f || !g) return LBM_ERR_ARG;
    if (h->in_batch) return refuse_batch_domain(h, "lbm_scatter");
    HIPCHK(h, hipSetDevice(h->cfg.device));
    const int nk = chunk_planes(h);
    const size_t plane = (size_t)h->cfg.nx * h->cfg.ny;
    double *stage = nullptr;
    if (hipMalloc((void **)&stage, (size_t)nk * plane * (h->qf + h->qt) * sizeof(double)) != hipSuccess)
        return fail(h, LBM_ERR_NOMEM, "scatter staging");
    double *fo = stage, *go = stage + (size_t)nk * plane * h->qf;
    int rc = LBM_OK;
    for (int k0 = 0; k0 < h->cfg.nz && rc == LBM_OK; k0 += nk) {
        const int n = (k0 + nk <= h->cfg.nz) ? nk : h->cfg.nz - k0;
        const uint64_t c0 = (uint64_t)k0 * plane, cnt = (uint64_t)n * plane;
        hipError_t e = hipMemcpyAsync(fo, f + c0 * h->qf, cnt * h->qf * sizeof(double), hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) e = hipMemcpyAsync(go, g + c0 * h->qt, cnt * h->qt * sizeof(double), hipMemcpyHostToDevice, h->stream);
        if (h->lp) {
            if (e == hipSuccess && k0 == 0) e = hipMemsetAsync(h->flag, 0, sizeof(int), h->stream);
            LPArgs a;
            a.fo = fo;
            a.go = go;
            a.Fn = h->F[0];
            a.Go = h->G[0];
            a.mac = h->mac;
            a.c0 = c0;
            a.n = cnt;
            a.flag = h->flag;
            if (e == hipSuccess) e = launch_lp(h->cfg.precision, LP_FROM_ABS, h->kp, h->bx, h->stream, a);
            if (e == hipSuccess) e = launch_lp(h->cfg.precision, LP_RECON_AOS, h->kp, h->bx, h->stream, a);
        } else {
            if (e == hipSuccess) e = launch_transpose(true, h->stream, fo, h->F[0], h->qf, h->kp.P, c0, cnt);
            if (e == hipSuccess) e = launch_transpose(true, h->stream, go, h->G[0], h->qt, h->kp.P, c0, cnt);
        }
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) rc = fail(h, LBM_ERR_HIP, "scatter: %s", hipGetErrorString(e));
    }
    (void)hipFree(stage);
    if (rc) return rc;
    if (h->lp) {
        LPArgs a;
        a.mac = h->mac;
        a.Bp = (float *)h->Bp;
        HIPCHK(h, launch_lp(h->cfg.precision, LP_BP_FROM_MAC, h->kp, h->bx, h->stream, a));
    } else {
        HIPCHK(h, hipMemsetAsync(h->flag, 0, sizeof(int), h->stream));
        HIPCHK(h, launch_reconstruct(h->kp, h->stream, h->F[0], h->G[0], h->Bp, h->mac, h->qf, h->qt, h->flag));
    }
    int bad = 0;
    HIPCHK(h, hipMemcpyAsync(&bad, h->flag, sizeof(int), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->phase = 0;
    if (bad) return fail(h, LBM_ERR_RHO, "non-positive rho found during restart reconstruction");
    return LBM_OK;
}

// =========================================================================================
// Batched ensembles
// =========================================================================================
struct lbm_batch {
    int n;
    int precision;
    lbm_handle **dom;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    int mode;
    // one-launch tables, one per buffer parity, built at the first batched step; two-pass form
    // (two launches a step, 256-node blocks) unless every domain runs the one-pass form
    bool twopass;
    void *table[2];
    uint16_t *owner;
    unsigned nblocks;
    int table_parity;      // parity whose table matches the domains' current buffers
    bool tables_ready;
    hipGraph_t graph;
    hipGraphExec_t gexec;
    bool graph_ready;
    int table_parity_graph;   // parity the graph's first step starts from
};

static void batch_release(lbm_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (int p = 0; p < 2; ++p)
        if (b->table[p]) (void)hipFree(b->table[p]);
    if (b->owner) (void)hipFree(b->owner);
    if (b->dom) {
        for (int i = 0; i < b->n; ++i)
            if (b->dom[i]) {
                b->dom[i]->in_batch = false;
                lbm_destroy(b->dom[i]);
            }
        free(b->dom);
    }
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    free(b);
}

extern "C" int lbm_batch_create(const lbm_config *cfgs, int32_t n, lbm_batch **out)
{
    if (!cfgs || !out || n <= 0 || n > 65535) return LBM_ERR_ARG;
    *out = nullptr;
    for (int i = 0; i < n; ++i)
        if (cfgs[i].dims != 3 || cfgs[i].precision != cfgs[0].precision ||
            cfgs[i].device != cfgs[0].device || cfgs[i].temp_x == LBM_T_PERIODIC ||
            cfgs[i].temp_z == LBM_T_PERIODIC || cfgs[i].layout == LBM_LAYOUT_AA)
            return LBM_ERR_UNSUPPORTED;   // table forms: 3D walls, one precision, two arrays
    lbm_batch *b = (lbm_batch *)calloc(1, sizeof(lbm_batch));
    if (!b) return LBM_ERR_NOMEM;
    b->n = n;
    b->precision = cfgs[0].precision;
    b->mode = LBM_BATCH_GRAPH;
    if (hipSetDevice(cfgs[0].device) != hipSuccess ||
        hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev0) != hipSuccess || hipEventCreate(&b->ev1) != hipSuccess) {
        batch_release(b);
        return LBM_ERR_HIP;
    }
    b->dom = (lbm_handle **)calloc(n, sizeof(lbm_handle *));
    if (!b->dom) { batch_release(b); return LBM_ERR_NOMEM; }
    for (int i = 0; i < n; ++i) {
        lbm_config ci = cfgs[i];
        ci.layout = LBM_LAYOUT_TWO_ARRAY;
        int rc = create_impl(&ci, b->stream, &b->dom[i]);
        if (rc) { batch_release(b); return rc; }
        b->dom[i]->in_batch = true;
        b->dom[i]->instance = (uint64_t)i;
    }
    *out = b;
    return LBM_OK;
}

extern "C" int lbm_batch_destroy(lbm_batch *b)
{
    if (!b) return LBM_ERR_ARG;
    (void)hipStreamSynchronize(b->stream);
    batch_release(b);
    return LBM_OK;
}

extern "C" int lbm_batch_domain(lbm_batch *b, int32_t i, lbm_handle **h)
{
    if (!b || !h || i < 0 || i >= b->n) return LBM_ERR_ARG;
    *h = b->dom[i];
    return LBM_OK;
}

extern "C" int lbm_batch_nodes(lbm_batch *b, int64_t *nodes)
{
    if (!b || !nodes) return LBM_ERR_ARG;
    int64_t s = 0;
    for (int i = 0; i < b->n; ++i) s += (int64_t)b->dom[i]->N;
    *nodes = s;
    return LBM_OK;
}

static void batch_invalidate(lbm_batch *b)
{
    b->tables_ready = false;
    for (int p = 0; p < 2; ++p)
        if (b->table[p]) { (void)hipFree(b->table[p]); b->table[p] = nullptr; }
    if (b->owner) { (void)hipFree(b->owner); b->owner = nullptr; }
    if (b->gexec) { (void)hipGraphExecDestroy(b->gexec); b->gexec = nullptr; }
    if (b->graph) { (void)hipGraphDestroy(b->graph); b->graph = nullptr; }
    b->graph_ready = false;
}

extern "C" int lbm_batch_init_seeded(lbm_batch *b, uint64_t seed, double amp)
{
    if (!b) return LBM_ERR_ARG;
    if (!(amp >= 0.0)) return LBM_ERR_ARG;
    for (int i = 0; i < b->n; ++i) {
        int rc = do_init(b->dom[i], nullptr, seed, amp, true);
        if (rc) return rc;
    }
    batch_invalidate(b);
    return LBM_OK;
}

// The per-shape choices of lbm_batch_autotune time each shape alone, where a small cavity
// cannot fill the device with one-pass tiles; a one-launch step of the whole batch can.  This
// times the batch's one-launch step with those choices against every domain in the one-pass
// form (domain 0's one-pass tile and prefetch when it chose the one-pass form, else the
// precision form's default; each domain's one-pass z-chunk, else the batch default) on
// a scratch batch of the same configurations.  The incumbent is the form the batch runs under
// AUTO (the one-pass table once the batch's total node count reaches the one-pass size, see
// batch_table_one_pass); the other is taken only when it is more than 3% faster, since
// placement alone moves a form's time by several per cent.  Skipped when a domain cannot run
// the one-pass form, when every domain already does, or when free device memory cannot hold
// the scratch batch.
static int batch_tune_table_form(lbm_batch *b, int steps)
{
    lbm_handle *h0 = b->dom[0];
    bool all_one = true;
    size_t need = 0;
    uint64_t total = 0;
    for (int i = 0; i < b->n; ++i) {
        if (!b->dom[i]->fused_ok) return LBM_OK;
        all_one = all_one && b->dom[i]->variant == LBM_KERNEL_ONE_PASS;
        need += b->dom[i]->dev_bytes;
        total += b->dom[i]->N;
    }
    const bool auto_one = b->precision != LBM_FP16_SHIFT && total >= one_pass_min_nodes(b->precision);
    if (all_one) return LBM_OK;
    size_t free_b = 0, all_b = 0;
    if (hipMemGetInfo(&free_b, &all_b) != hipSuccess) {
        (void)hipGetLastError();
        return LBM_OK;
    }
    if ((double)need > 0.8 * (double)free_b) return LBM_OK;
    int tile0, pf0;
    if (h0->variant == LBM_KERNEL_ONE_PASS) {
        tile0 = h0->fg.tile;
        pf0 = h0->fg.pf;
    } else {
        default_tile(b->precision, &tile0, &pf0);
    }
    auto one_pass_kchunk = [](const lbm_handle *h) {
        return h->variant == LBM_KERNEL_ONE_PASS ? h->fg.kchunk : batch_kchunk(h->cfg.nz);
    };
    std::vector<lbm_config> cf(b->n);
    for (int i = 0; i < b->n; ++i) cf[i] = b->dom[i]->cfg;
    lbm_batch *t = nullptr;
    int rc = lbm_batch_create(cf.data(), b->n, &t);
    if (rc == LBM_ERR_NOMEM) return LBM_OK;   // no room for the scratch batch after all
    if (rc) return fail(h0, rc, "batch autotune: scratch batch: %s", lbm_error_string(rc));
    t->mode = LBM_BATCH_ONE_LAUNCH;
    rc = lbm_batch_init_seeded(t, 1, 1e-3);
    double ms[2] = {0.0, 0.0};
    for (int c = 0; c < 2 && !rc; ++c) {
        for (int i = 0; i < b->n && !rc; ++i) {
            const lbm_handle *h = b->dom[i];
            rc = c == 0 ? set_form(t->dom[i], h->variant, h->fg.kchunk, h->fg.tile, h->fg.pf)
                        : set_form(t->dom[i], LBM_KERNEL_ONE_PASS, one_pass_kchunk(h), tile0, pf0);
        }
        batch_invalidate(t);
        if (!rc) rc = lbm_batch_step(t, 2);   // the first leaves the plain state
        if (!rc) rc = lbm_batch_step_timed(t, steps, &ms[c]);
    }
    lbm_batch_destroy(t);
    if (rc) return fail(h0, rc, "batch autotune: timing failed");
    if (auto_one ? !(ms[0] < 0.97 * ms[1]) : ms[1] < 0.97 * ms[0])
        for (int i = 0; i < b->n && !rc; ++i) {
            lbm_handle *h = b->dom[i];
            rc = set_form(h, LBM_KERNEL_ONE_PASS, one_pass_kchunk(h), tile0, pf0);
        }
    return rc;
}

extern "C" int lbm_batch_autotune(lbm_batch *b, const char *cache_path, int32_t steps)
{
    if (!b || steps <= 0) return LBM_ERR_ARG;
    HIPCHK(b->dom[0], hipStreamSynchronize(b->stream));
    // tune each distinct shape once, then copy the choice to every domain of that shape
    for (int i = 0; i < b->n; ++i) {
        int j = 0;
        for (; j < i; ++j)
            if (b->dom[j]->cfg.nx == b->dom[i]->cfg.nx && b->dom[j]->cfg.ny == b->dom[i]->cfg.ny &&
                b->dom[j]->cfg.nz == b->dom[i]->cfg.nz)
                break;
        int rc;
        const lbm_handle *src = b->dom[j];
        if (j < i) rc = set_form(b->dom[i], src->variant, src->fg.kchunk, src->fg.tile, src->fg.pf);
        else rc = autotune_impl(b->dom[i], cache_path, steps, nullptr);
        if (rc) return rc;
    }
    // One launch has one block shape: every domain takes the tile shape and prefetch choice of
    // domain 0 and keeps the z-chunk tuned for its own shape.
    for (int i = 1; i < b->n; ++i) {
        lbm_handle *h = b->dom[i];
        int rc = set_form(h, h->variant, h->fg.kchunk, b->dom[0]->fg.tile, b->dom[0]->fg.pf);
        if (rc) return rc;
    }
    int rc = batch_tune_table_form(b, steps);
    batch_invalidate(b);
    return rc;
}

extern "C" int lbm_batch_set_kernel_variant(lbm_batch *b, int32_t variant)
{
    if (!b || variant < LBM_KERNEL_AUTO || variant > LBM_KERNEL_ONE_PASS) return LBM_ERR_ARG;
    HIPCHK(b->dom[0], hipStreamSynchronize(b->stream));
    for (int i = 0; i < b->n; ++i) {
        lbm_handle *h = b->dom[i];
        int rc = set_form(h, variant, 0, h->fg.tile, h->fg.pf);
        if (rc) return rc;
    }
    batch_invalidate(b);
    return LBM_OK;
}

extern "C" int lbm_batch_set_mode(lbm_batch *b, int32_t mode)
{
    if (!b || mode < LBM_BATCH_PER_DOMAIN || mode > LBM_BATCH_GRAPH) return LBM_ERR_ARG;
    b->mode = mode;
    return LBM_OK;
}

// The form of a one-launch step: one-pass when every domain admits it and none is set to
// TWO_PASS, and either every domain is set to ONE_PASS or, under AUTO, the batch as a whole has
// the node count at which a single domain runs the one-pass form: the combined launch fills the
// device although each cavity alone would not.  The seeded 64-cavity ensemble (edges 32-128,
// 44.6 M nodes), graph replay: FP32-shift 2.77 ms per step one-pass against 2.99 two-pass, FP64
// 5.29 against 6.25.
static bool batch_table_one_pass(const lbm_batch *b)
{
    uint64_t total = 0;
    bool all_one = true;
    for (int i = 0; i < b->n; ++i) {
        const lbm_handle *h = b->dom[i];
        if (!h->fused_ok || h->variant == LBM_KERNEL_TWO_PASS) return false;
        all_one = all_one && h->variant == LBM_KERNEL_ONE_PASS;
        total += h->N;
    }
    return all_one || (b->precision != LBM_FP16_SHIFT && total >= one_pass_min_nodes(b->precision));
}

// Blocks of one domain in the batch's form.
static unsigned batch_blocks(const lbm_batch *b, const lbm_handle *h)
{
    return b->twopass ? (unsigned)((h->N + 255) / 256) : h->bfg.nblocks;
}


// Tables for both buffer parities from the domains' current pointers: parity p is the current
// state, parity p^1 the state after one step (F and G swapped; in the one-pass form also
// B_prev when use_g, which the two-pass form updates in place).
static int batch_tables(lbm_batch *b)
{
    lbm_handle *h0 = b->dom[0];
    b->twopass = !batch_table_one_pass(b);
    for (int i = 0; i < b->n; ++i) {
        lbm_handle *h = b->dom[i];
        h->bfg = h->variant == LBM_KERNEL_ONE_PASS
                     ? h->fg
                     : make_fused_geom(h->cfg, batch_kchunk(h->cfg.nz), h->fg.tile, h->fg.pf);
    }
    const size_t isz = !b->twopass ? fused_entry_bytes(h0)
                       : h0->lp    ? twopass_entry_bytes_lp(b->precision)
                                   : twopass_entry_bytes_f64();
    unsigned total = 0;
    for (int i = 0; i < b->n; ++i) total += batch_blocks(b, b->dom[i]);
    if (!b->owner) {
        if (hipMalloc((void **)&b->owner, total * sizeof(uint16_t)) != hipSuccess ||
            hipMalloc(&b->table[0], isz * b->n) != hipSuccess ||
            hipMalloc(&b->table[1], isz * b->n) != hipSuccess)
            return fail(h0, LBM_ERR_NOMEM, "batch tables");
        std::vector<uint16_t> own(total);
        unsigned at = 0;
        for (int i = 0; i < b->n; ++i)
            for (unsigned k = 0; k < batch_blocks(b, b->dom[i]); ++k) own[at++] = (uint16_t)i;
        HIPCHK(h0, hipMemcpy(b->owner, own.data(), total * sizeof(uint16_t), hipMemcpyHostToDevice));
        b->nblocks = total;
    }
    std::vector<unsigned char> host(isz * b->n);
    for (int par = 0; par < 2; ++par) {
        unsigned first = 0;
        for (int i = 0; i < b->n; ++i) {
            lbm_handle *h = b->dom[i];
            void *dst = host.data() + isz * i;
            if (!b->twopass)
                fused_fill(h, dst, fused_ptrs(h, h->fcur, h->gcur, h->Bp, h->Bp_alt, par), first, h->bfg);
            else if (h->lp)
                twopass_entry_fill_lp(b->precision, dst, h->kp,
                                      fused_ptrs(h, h->fcur, h->gcur, h->Bp, h->Bp, par),
                                      (float *)h->aux[0], first);
            else
                twopass_entry_fill_f64(dst, h->kp, fused_ptrs(h, h->fcur, h->gcur, h->Bp, h->Bp, par),
                                       h->aux[0], first);
            first += batch_blocks(b, h);
        }
        HIPCHK(h0, hipMemcpy(b->table[par], host.data(), host.size(), hipMemcpyHostToDevice));
    }
    b->table_parity = 0;
    b->tables_ready = true;
    return LBM_OK;
}

static hipError_t batch_launch(lbm_batch *b, int par)
{
    if (b->twopass)
        return b->dom[0]->lp
                   ? launch_twopass_batch_lp(b->precision, b->table[par], b->owner, b->nblocks, b->stream)
                   : launch_twopass_batch_f64(b->table[par], b->owner, b->nblocks, b->stream);
    const FusedGeom &fg0 = b->dom[0]->fg;
    if (b->dom[0]->lp)
        return launch_fused_batch_lp(b->precision, fg0, b->table[par], b->owner, b->nblocks, b->stream);
    return launch_fused_batch_f64(fg0, b->table[par], b->owner, b->nblocks, b->stream);
}

// Host bookkeeping of one batched step on every domain (buffers swap like enqueue_steps).
static void batch_advance(lbm_batch *b)
{
    for (int i = 0; i < b->n; ++i) {
        lbm_handle *h = b->dom[i];
        h->fcur ^= 1;
        h->gcur ^= 1;
        h->aux_valid = b->twopass ? 1 : 0;
        h->itc += 1;
        if (!b->twopass && h->cfg.use_g) std::swap(h->Bp, h->Bp_alt);
    }
    b->table_parity ^= 1;
}

static int batch_enqueue(lbm_batch *b, int32_t nsteps)
{
    // the first step after initialisation reads the plain state: per domain, two-pass
    for (int i = 0; i < b->n; ++i) {
        lbm_handle *h = b->dom[i];
        if (h->phase < 0) return fail(h, LBM_ERR_STATE, "batch domain %d not initialised", i);
    }
    int done = 0;
    if (nsteps > 0 && (b->dom[0]->phase == 0 || b->dom[0]->first_thermal)) {
        for (int i = 0; i < b->n; ++i) {
            int rc = enqueue_steps(b->dom[i], 1);
            if (rc) return rc;
        }
        batch_invalidate(b);
        done = 1;
    }
    if (b->mode == LBM_BATCH_PER_DOMAIN) {
        for (int s = done; s < nsteps; ++s)
            for (int i = 0; i < b->n; ++i) {
                int rc = enqueue_steps(b->dom[i], 1);
                if (rc) return rc;
            }
        batch_invalidate(b);
        return LBM_OK;
    }
    if (done >= nsteps) return LBM_OK;
    if (!b->tables_ready) {
        int rc = batch_tables(b);
        if (rc) return rc;
    }
    lbm_handle *h0 = b->dom[0];
    if (b->mode == LBM_BATCH_GRAPH) {
        if (!b->graph_ready) {
            // two steps return every buffer to its place: capture parity p then p^1
            HIPCHK(h0, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
            hipError_t e = batch_launch(b, b->table_parity);
            if (e == hipSuccess) e = batch_launch(b, b->table_parity ^ 1);
            hipError_t e2 = hipStreamEndCapture(b->stream, &b->graph);
            if (e != hipSuccess || e2 != hipSuccess)
                return fail(h0, LBM_ERR_HIP, "graph capture: %s", hipGetErrorString(e != hipSuccess ? e : e2));
            HIPCHK(h0, hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
            b->graph_ready = true;
            b->table_parity_graph = b->table_parity;
        }
        if (b->table_parity != b->table_parity_graph && nsteps - done >= 3) {
            HIPCHK(h0, batch_launch(b, b->table_parity));   // realign with the graph's parity
            batch_advance(b);
            ++done;
        }
        while (nsteps - done >= 2 && b->table_parity == b->table_parity_graph) {
            HIPCHK(h0, hipGraphLaunch(b->gexec, b->stream));
            batch_advance(b);
            batch_advance(b);
            done += 2;
        }
    }
    for (; done < nsteps; ++done) {
        HIPCHK(h0, batch_launch(b, b->table_parity));
        batch_advance(b);
    }
    return LBM_OK;
}

extern "C" int lbm_batch_run_diagnostics(lbm_batch *b, int32_t nsteps, int32_t every, lbm_diag *out)
{
    if (!b || !out || every <= 0 || nsteps < every || nsteps % every) return LBM_ERR_ARG;
    lbm_handle *h0 = b->dom[0];
    HIPCHK(h0, hipSetDevice(h0->cfg.device));
    for (int i = 0; i < b->n; ++i)
        if (b->dom[i]->phase < 0) return fail(b->dom[i], LBM_ERR_STATE, "batch domain %d not initialised", i);
    const int ns = nsteps / every, n = b->n;
    std::vector<size_t> off(n + 1, 0);
    for (int i = 0; i < n; ++i) off[i + 1] = off[i] + (size_t)ns * diag_sample_doubles(b->dom[i]);
    DiagSeries ser;
    if (hipMalloc((void **)&ser.buf, off[n] * sizeof(double)) != hipSuccess ||
        hipMalloc((void **)&ser.count, (size_t)n * sizeof(int)) != hipSuccess)
        return fail(h0, LBM_ERR_NOMEM, "diagnostics series");
    HIPCHK(h0, launch_series_reset(b->stream, ser.count, n));
    auto diag_all = [&]() -> int {
        for (int i = 0; i < n; ++i) {
            int rc = diag_append(b->dom[i], ser.buf + off[i], ns, ser.count + i);
            if (rc) return rc;
        }
        return LBM_OK;
    };
    // the first period eagerly, in the batch's own mode (it leaves the plain states)
    int rc = batch_enqueue(b, every);
    if (!rc) rc = diag_all();
    if (rc) return rc;
    int done = 1;
    const bool table = b->mode != LBM_BATCH_PER_DOMAIN;
    // one launch per domain: a second eager period when that first one was a single two-pass
    // step before one-pass steps (the domains' one-pass launch tables are built eagerly)
    auto all_ready = [&]() {
        for (int i = 0; i < n; ++i)
            if (!capture_ready(b->dom[i])) return false;
        return true;
    };
    if (!table && done < ns && !all_ready()) {
        rc = batch_enqueue(b, every);
        if (!rc) rc = diag_all();
        if (rc) return rc;
        ++done;
    }
    const int P = graph_periods(every);
    const int replays = (ns - done) / P;
    if (replays > 0 && table && !b->tables_ready) rc = batch_tables(b);   // outside the capture
    if (rc) return rc;
    if (replays > 0) {
        std::vector<int64_t> itc0(n);
        for (int i = 0; i < n; ++i) itc0[i] = b->dom[i]->itc;
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        HIPCHK(h0, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        for (int i = 0; i < n; ++i) b->dom[i]->capturing = true;
        for (int p = 0; p < P && !rc; ++p) {
            for (int s = 0; s < every && !rc; ++s) {
                if (table) {
                    if (batch_launch(b, b->table_parity) != hipSuccess)
                        rc = fail(h0, LBM_ERR_HIP, "batch launch during capture");
                    batch_advance(b);
                } else {
                    for (int i = 0; i < n && !rc; ++i) rc = enqueue_steps(b->dom[i], 1);
                }
            }
            if (!rc) rc = diag_all();
        }
        for (int i = 0; i < n; ++i) b->dom[i]->capturing = false;
        hipError_t e = hipStreamEndCapture(b->stream, &g);
        if (!rc && e == hipSuccess) e = hipGraphInstantiate(&ge, g, nullptr, nullptr, 0);
        // host bookkeeping now stands one replay ahead (buffers back in place); the step
        // counters follow the replays that were launched
        int launched = 0;
        for (; launched < replays && !rc && e == hipSuccess; ++launched) {
            e = hipGraphLaunch(ge, b->stream);
            if (e != hipSuccess) break;
        }
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
        for (int i = 0; i < n; ++i) b->dom[i]->itc = itc0[i] + (int64_t)launched * P * every;
        if (rc) return rc;
        if (e != hipSuccess) return fail(h0, LBM_ERR_HIP, "diagnostics graph: %s", hipGetErrorString(e));
        done += replays * P;
    }
    for (; done < ns; ++done) {
        rc = batch_enqueue(b, every);
        if (!rc) rc = diag_all();
        if (rc) return rc;
    }
    std::vector<double> host(off[n]);
    HIPCHK(h0, hipMemcpyAsync(host.data(), ser.buf, host.size() * sizeof(double),
                              hipMemcpyDeviceToHost, b->stream));
    HIPCHK(h0, hipStreamSynchronize(b->stream));
    for (int i = 0; i < n; ++i) {
        const size_t stride = diag_sample_doubles(b->dom[i]);
        for (int k = 0; k < ns; ++k)
            diag_finish(b->dom[i], host.data() + off[i] + (size_t)k * stride, &out[(size_t)k * n + i]);
    }
    return LBM_OK;
}

extern "C" int lbm_batch_step(lbm_batch *b, int32_t nsteps)
{
    if (!b || nsteps < 0) return LBM_ERR_ARG;
    HIPCHK(b->dom[0], hipSetDevice(b->dom[0]->cfg.device));
    return batch_enqueue(b, nsteps);
}

extern "C" int lbm_batch_step_timed(lbm_batch *b, int32_t nsteps, double *ms)
{
    if (!b || nsteps < 0 || !ms) return LBM_ERR_ARG;
    lbm_handle *h0 = b->dom[0];
    HIPCHK(h0, hipSetDevice(h0->cfg.device));
    HIPCHK(h0, hipEventRecord(b->ev0, b->stream));
    int rc = batch_enqueue(b, nsteps);
    if (rc) return rc;
    HIPCHK(h0, hipEventRecord(b->ev1, b->stream));
    HIPCHK(h0, hipEventSynchronize(b->ev1));
    float t = 0.f;
    HIPCHK(h0, hipEventElapsedTime(&t, b->ev0, b->ev1));
    *ms = t;
    return LBM_OK;
}
