/*
 * rb_oracle.c -- CPU FP64 restatement of the reference's Rayleigh-Benard MRT-LBM hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity checker and the CPU baseline); the product never calls it.
 *
 * Every routine names the reference routine it restates (file:line, paths relative to the
 * reference repository root).  The reference keeps its state in module globals; here it lives
 * in `orc_state`, with the same arrays and the same layout:
 *   f(0:q-1, nx, ny, nz)                 population index fastest, interior only
 *   f_post(0:q-1, 0:nx+1, 0:ny+1, 0:nz+1) one-cell halo whose ghosts stay zero
 * and the same eight-sweep step: collision, streaming, bounceback, macro, collisionT,
 * streamingT, bouncebackT, macroT (2DRB/3DRB.F90:225-243).
 *
 * Build with -ffp-contract=off: the reference's x86-64 build has no fused multiply-add, and
 * expressions below keep the reference's evaluation order.
 */
#include "rb_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ----------------------------------------------------------------------------------------- */
/* Lattice tables.  D3Q19 order, opp and weights: 2DRB/3DRB.F90:656-678.                     */
/* D2Q9 order: 2DRB/2DRB.F90:188-190; weights 2DRB/2DRB.F90:506-524.                        */
/* ----------------------------------------------------------------------------------------- */
static const int EX3[19] = {0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 1, -1, 1, -1, 0, 0, 0, 0};
static const int EY3[19] = {0, 0, 0, 1, -1, 0, 0, 1, 1, -1, -1, 0, 0, 0, 0, 1, -1, 1, -1};
static const int EZ3[19] = {0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, 1, -1, -1};
static const int OPP3[19] = {0, 2, 1, 4, 3, 6, 5, 10, 9, 8, 7, 14, 13, 12, 11, 18, 17, 16, 15};
static const int OPP7[7] = {0, 2, 1, 4, 3, 6, 5};

static const int EX2[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
static const int EY2[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};

struct orc_state {
    orc_cfg c;
    orc_params p;
    int qf, qt;
    long n;                      /* interior nodes */
    long itc;
    double omega[19], omegaT[7];
    double *xp, *yp, *zp;        /* 0..n+1 node coordinates / L (3DRB.F90:451-470) */
    double *f, *fpost, *g, *gpost;
    double *rho, *u, *v, *w, *T;
    double *up, *vp, *wp, *Tp;
    double *bx, *by, *bz;
    double *Fx, *Fy;             /* 2D only: forces stored by collision (2DRB.F90:718-719) */
};

/* ----------------------------------------------------------------------------------------- */
/* Parameters: 2DRB/3DRB.F90:73-123 (3D) and 2DRB/2DRB.F90:61-115 (2D).                      */
/* ----------------------------------------------------------------------------------------- */
int orc_params_compute(const orc_cfg *c, orc_params *p)
{
    memset(p, 0, sizeof(*p));
    p->L = c->side_heated ? (double)c->nx : (double)c->ny;
    p->Tref = 0.5 * (c->thot + c->tcold);
    p->tauf = 0.5 + c->mach * p->L * sqrt(3.0 * c->prandtl / c->rayleigh);
    p->viscosity = (p->tauf - 0.5) / 3.0;
    p->diffusivity = p->viscosity / c->prandtl;
    p->heat_flux_scale = p->L / p->diffusivity;
    double gbeta1 = c->rayleigh * p->viscosity * p->diffusivity / p->L;
    p->gbeta = gbeta1 / p->L / p->L;
    p->time_unit = sqrt(p->L / p->gbeta);
    p->velocity_unit = sqrt(p->gbeta * p->L);
    p->Se = 1.0 / p->tauf;
    p->Seps = 1.0 / p->tauf;
    p->Snu = 1.0 / p->tauf;
    p->Spi = 1.0 / p->tauf;
    p->Sq = 8.0 * (2.0 * p->tauf - 1.0) / (8.0 * p->tauf - 1.0);
    p->Sm = 8.0 * (2.0 * p->tauf - 1.0) / (8.0 * p->tauf - 1.0);
    if (c->dims == 3) {
        const double cs2T = 0.25;
        p->taug = 0.5 + p->diffusivity / cs2T;           /* 3DRB.F90:121 */
    } else {
        p->taug = 0.5 + (p->tauf - 0.5) / c->prandtl;    /* 2DRB.F90:111 */
    }
    p->Qnu = 1.0;
    p->Qk = 1.0 / p->taug;
    if (c->legacy) {
        /* EnableLegacyThermalScheme: 3DRB.F90:93, :116-118, :378-390; 2DRB.F90:95, :108-110,
         * :404-413.  The diffusivity enters through paraA (the rest weight) instead of Qk. */
        if (c->dims == 3) {
            p->paraA = 42.0 * sqrt(3.0) * p->diffusivity - 6.0;
            if (p->paraA >= 1.0 || p->paraA <= -6.0) return -1;
        } else {
            p->paraA = 20.0 * sqrt(3.0) * p->diffusivity - 4.0;
            if (p->paraA >= 1.0 || p->paraA <= -4.0) return -1;
        }
        p->taug = 0.0;
        p->Qk = 3.0 - sqrt(3.0);
        p->Qnu = 4.0 * sqrt(3.0) - 6.0;
    }
    p->SG = 1.0 - 0.5 * p->Qk;                           /* 3DRB.F90:1113, 2DRB.F90:915 */
    return 0;
}

/* ----------------------------------------------------------------------------------------- */
/* Compensated (Neumaier) summation used by every global reduction.                          */
/* ----------------------------------------------------------------------------------------- */
typedef struct { double s, c; } ksum;
static inline void ks_add(ksum *a, double x)
{
    double t = a->s + x;
    if (fabs(a->s) >= fabs(x)) a->c += (a->s - t) + x;
    else a->c += (x - t) + a->s;
    a->s = t;
}
static inline void ks_merge(ksum *a, ksum b)
{
    ks_add(a, b.s);
    ks_add(a, b.c);
}
static inline double ks_val(ksum a) { return a.s + a.c; }

/* ----------------------------------------------------------------------------------------- */
/* Indexing helpers (0-based interior i in [0,nx); haloed index shifts by one).               */
/* ----------------------------------------------------------------------------------------- */
#define NX (s->c.nx)
#define NY (s->c.ny)
#define NZ (s->c.nz)
static inline long IDX(const orc_state *s, long i, long j, long k)
{
    return i + (long)s->c.nx * (j + (long)s->c.ny * k);
}
static inline long HDX(const orc_state *s, long i, long j, long k)
{
    return (i + 1) + (long)(s->c.nx + 2) * ((j + 1) + (long)(s->c.ny + 2) * (k + 1));
}

static void *xcalloc(size_t n, size_t sz)
{
    void *p = calloc(n ? n : 1, sz);
    return p;
}

/* Anti-bounce-back wall term: 2 wT Tw, legacy (6 + paraA) / 21 Tw (3D, 3DRB.F90:1267-1272,
 * :1311-1316) and (4 + paraA) / 10 Tw (2D, 2DRB.F90:1048-1055, :1091-1098). */
static double abb3(const orc_state *s, double w, double Tw)
{
    return s->c.legacy ? (6.0 + s->p.paraA) / 21.0 * Tw : 2.0 * w * Tw;
}
static double abb2(const orc_state *s, double w, double Tw)
{
    return s->c.legacy ? (4.0 + s->p.paraA) / 10.0 * Tw : 2.0 * w * Tw;
}

orc_state *orc_create(const orc_cfg *cfg)
{
    if (cfg->dims != 2 && cfg->dims != 3) return NULL;
    if (cfg->nx < 4 || cfg->ny < 4) return NULL;
    if (cfg->dims == 3 && cfg->nz < 4) return NULL;
    orc_state *s = (orc_state *)xcalloc(1, sizeof(orc_state));
    s->c = *cfg;
    if (s->c.dims == 2) s->c.nz = 1;
    if (orc_params_compute(&s->c, &s->p) != 0) {   /* legacy scheme outside its paraA range */
        free(s);
        return NULL;
    }
    s->qf = (s->c.dims == 3) ? 19 : 9;
    s->qt = (s->c.dims == 3) ? 7 : 5;
    s->n = (long)s->c.nx * s->c.ny * s->c.nz;
    long nh = (long)(s->c.nx + 2) * (s->c.ny + 2) * (s->c.dims == 3 ? (s->c.nz + 2) : 1);
    s->xp = (double *)xcalloc(s->c.nx + 2, sizeof(double));
    s->yp = (double *)xcalloc(s->c.ny + 2, sizeof(double));
    s->zp = (double *)xcalloc(s->c.nz + 2, sizeof(double));
    s->f = (double *)xcalloc((size_t)s->n * s->qf, sizeof(double));
    s->fpost = (double *)xcalloc((size_t)nh * s->qf, sizeof(double));
    s->g = (double *)xcalloc((size_t)s->n * s->qt, sizeof(double));
    s->gpost = (double *)xcalloc((size_t)nh * s->qt, sizeof(double));
    double **macro[] = {&s->rho, &s->u, &s->v, &s->w, &s->T, &s->up, &s->vp, &s->wp,
                        &s->Tp, &s->bx, &s->by, &s->bz, &s->Fx, &s->Fy};
    for (unsigned m = 0; m < sizeof(macro) / sizeof(macro[0]); ++m)
        *macro[m] = (double *)xcalloc((size_t)s->n, sizeof(double));
    if (!s->f || !s->fpost || !s->g || !s->gpost || !s->Fy) {
        orc_destroy(s);
        return NULL;
    }
    if (s->c.dims == 3) {
        /* 3DRB.F90:662-678 (non-legacy thermal weights) */
        s->omega[0] = 1.0 / 3.0;
        for (int a = 1; a <= 6; ++a) s->omega[a] = 1.0 / 18.0;
        for (int a = 7; a <= 18; ++a) s->omega[a] = 1.0 / 36.0;
        s->omegaT[0] = 1.0 / 4.0;
        for (int a = 1; a <= 6; ++a) s->omegaT[a] = 1.0 / 8.0;
        if (s->c.legacy) {   /* 3DRB.F90:671-673 */
            s->omegaT[0] = (1.0 - s->p.paraA) / 7.0;
            for (int a = 1; a <= 6; ++a) s->omegaT[a] = (6.0 + s->p.paraA) / 42.0;
        }
    } else {
        /* 2DRB.F90:506-524 */
        s->omega[0] = 4.0 / 9.0;
        for (int a = 1; a <= 4; ++a) s->omega[a] = 1.0 / 9.0;
        for (int a = 5; a <= 8; ++a) s->omega[a] = 1.0 / 36.0;
        s->omegaT[0] = 1.0 / 3.0;
        for (int a = 1; a <= 4; ++a) s->omegaT[a] = 1.0 / 6.0;
        if (s->c.legacy) {   /* 2DRB.F90:514-518 */
            s->omegaT[0] = (1.0 - s->p.paraA) / 5.0;
            for (int a = 1; a <= 4; ++a) s->omegaT[a] = (s->p.paraA + 4.0) / 20.0;
        }
    }
    return s;
}

void orc_destroy(orc_state *s)
{
    if (!s) return;
    double *arr[] = {s->xp, s->yp, s->zp, s->f, s->fpost, s->g, s->gpost, s->rho, s->u,
                     s->v, s->w, s->T, s->up, s->vp, s->wp, s->Tp, s->bx, s->by, s->bz,
                     s->Fx, s->Fy};
    for (unsigned m = 0; m < sizeof(arr) / sizeof(arr[0]); ++m) free(arr[m]);
    free(s);
}

long orc_itc(const orc_state *s) { return s->itc; }

/* Thread count of the OpenMP sweeps (the reference hard-codes 24, 2DRB/3DRB.F90:215). */
int orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* ----------------------------------------------------------------------------------------- */
/* Seeded noise: SplitMix64 finaliser of (seed, instance, global linear index).              */
/* ----------------------------------------------------------------------------------------- */
static inline uint64_t mix64(uint64_t z)
{
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
double orc_noise(uint64_t seed, uint64_t instance, uint64_t idx)
{
    uint64_t key = seed + 0x9E3779B97F4A7C15ULL * (idx + 1ULL + (instance << 40));
    double r = (double)(mix64(key) >> 11) * 0x1.0p-53;
    return 2.0 * r - 1.0;
}

/* ----------------------------------------------------------------------------------------- */
/* initial, exact branch: 3DRB.F90:437-470 (coordinates), :517-600 (fields);                 */
/* 2DRB.F90:461-490, :540-608.                                                               */
/* ----------------------------------------------------------------------------------------- */
static void coords(double *xp, int n, double L)
{
    xp[0] = 0.0;
    xp[n + 1] = (double)n;
    for (int i = 1; i <= n; ++i) xp[i] = (double)i - 0.5;
    for (int i = 0; i <= n + 1; ++i) xp[i] = xp[i] / L;
}

void orc_init_exact(orc_state *s, const double *T0)
{
    const orc_cfg *c = &s->c;
    const double L = s->p.L;
    const double pi = acos(-1.0);
    s->itc = 0;
    coords(s->xp, c->nx, L);
    coords(s->yp, c->ny, L);
    if (c->dims == 3) coords(s->zp, c->nz, L);
    memset(s->fpost, 0, sizeof(double) * (size_t)s->qf * (c->nx + 2) * (c->ny + 2) *
                             (c->dims == 3 ? (c->nz + 2) : 1));
    memset(s->gpost, 0, sizeof(double) * (size_t)s->qt * (c->nx + 2) * (c->ny + 2) *
                            (c->dims == 3 ? (c->nz + 2) : 1));
    for (long n = 0; n < s->n; ++n) {
        s->rho[n] = 1.0;
        s->u[n] = s->v[n] = s->w[n] = 0.0;
        s->T[n] = 0.0;
        s->bx[n] = s->by[n] = s->bz[n] = 0.0;
        s->up[n] = s->vp[n] = s->wp[n] = s->Tp[n] = 0.0;
        s->Fx[n] = s->Fy[n] = 0.0;
    }
    for (int k = 0; k < c->nz; ++k)
        for (int j = 0; j < c->ny; ++j)
            for (int i = 0; i < c->nx; ++i) {
                long n = IDX(s, i, j, k);
                double T = 0.0;
                if (T0) {
                    T = T0[n];
                } else {
                    if (c->temp_x == ORC_T_CONST)   /* VerticalWallsConstT */
                        T = c->thot + (s->xp[i + 1] - s->xp[0]) / (s->xp[c->nx + 1] - s->xp[0]) *
                                          (c->tcold - c->thot);
                    if (c->temp_y == ORC_T_CONST) { /* HorizontalWallsConstT */
                        T = c->thot + (s->yp[j + 1] - s->yp[0]) / (s->yp[c->ny + 1] - s->yp[0]) *
                                          (c->tcold - c->thot);
                        if (!c->side_heated) {
                            /* RB perturbation: 3D if Ra <= 1e4 (3DRB.F90:542-558),
                             * 2D if Ra < 1e4 (2DRB.F90:568-581). */
                            double amp = 1.0e-3 * (c->thot - c->tcold);
                            double xLen = s->xp[c->nx + 1], yLen = s->yp[c->ny + 1];
                            if (c->dims == 3 && c->rayleigh <= 1.0e4)
                                T = T + amp * cos(2.0 * pi * s->xp[i + 1] / xLen) *
                                            sin(pi * s->yp[j + 1] / yLen);
                            if (c->dims == 2 && c->rayleigh < 1.0e4)
                                T = T + amp * sin(2.0 * pi * s->xp[i + 1] / xLen) *
                                            sin(pi * s->yp[j + 1] / yLen);
                        }
                    }
                }
                s->T[n] = T;
            }
    /* f = w rho (1 + 3 e.u + 4.5 (e.u)^2 - 1.5 u^2); g = wT T (1 + c e.u); B_prev = u T. */
    for (long n = 0; n < s->n; ++n) {
        double u = s->u[n], v = s->v[n], w = s->w[n], rho = s->rho[n], T = s->T[n];
        if (c->dims == 3) {
            double u2 = u * u + v * v + w * w;
            for (int a = 0; a < 19; ++a) {
                double eu = (double)EX3[a] * u + (double)EY3[a] * v + (double)EZ3[a] * w;
                s->f[n * 19 + a] =
                    s->omega[a] * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * u2);
            }
            for (int a = 0; a < 7; ++a) {
                double eu = (double)EX3[a] * u + (double)EY3[a] * v + (double)EZ3[a] * w;
                /* thermalGeqCoeff: 1/cs2T, legacy 7/((6 + paraA) cs2T) (3DRB.F90:118, :122) */
                const double gc = c->legacy ? 7.0 / ((6.0 + s->p.paraA) * 0.25) : 4.0;
                s->g[n * 7 + a] = s->omegaT[a] * T * (1.0 + gc * eu);
            }
        } else {
            double us2 = u * u + v * v;
            for (int a = 0; a < 9; ++a) {
                double un = u * EX2[a] + v * EY2[a];
                s->f[n * 9 + a] =
                    rho * s->omega[a] * (1.0 + 3.0 * un + 4.5 * un * un - 1.5 * us2);
            }
            for (int a = 0; a < 5; ++a) {
                double un = u * EX2[a] + v * EY2[a];
                /* thermalGeqCoeff: 3, legacy 10/(4 + paraA) (2DRB.F90:110, :114) */
                const double gc = c->legacy ? 10.0 / (4.0 + s->p.paraA) : 3.0;
                s->g[n * 5 + a] = s->omegaT[a] * T * (1.0 + gc * un);
            }
        }
        if (c->use_g) {
            s->bx[n] = u * T;
            s->by[n] = v * T;
            s->bz[n] = w * T;
        }
    }
}

void orc_init_seeded(orc_state *s, uint64_t seed, uint64_t instance, double amp)
{
    orc_init_exact(s, NULL);
    const double scale = amp * (s->c.thot - s->c.tcold);
    for (long n = 0; n < s->n; ++n) {
        double r = orc_noise(seed, instance, (uint64_t)n);
        double noise = scale * r;
        s->T[n] = s->T[n] + noise;
        for (int a = 0; a < s->qt; ++a) s->g[n * s->qt + a] = s->omegaT[a] * s->T[n];
    }
}

/* ========================================================================================= */
/* 3D: D3Q19 MRT flow + D3Q7 MRT temperature                                                 */
/* ========================================================================================= */

/* collision, 3DRB.F90:689-949 */
static void collision3d(orc_state *s)
{
    const orc_params *p = &s->p;
    const double gBeta = p->gbeta, Tref = p->Tref;
    double sr[19];
    sr[0] = 0.0; sr[1] = p->Se; sr[2] = p->Seps; sr[3] = 0.0; sr[4] = p->Sq; sr[5] = 0.0;
    sr[6] = p->Sq; sr[7] = 0.0; sr[8] = p->Sq; sr[9] = p->Snu; sr[10] = p->Spi;
    sr[11] = p->Snu; sr[12] = p->Spi; sr[13] = p->Snu; sr[14] = p->Snu; sr[15] = p->Snu;
    sr[16] = p->Sm; sr[17] = p->Sm; sr[18] = p->Sm;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                const long n = IDX(s, i, j, k);
                const double *F = s->f + n * 19;
                double m[19], meq[19], mp[19], src[19];
                const double rhoLoc = s->rho[n], uLoc = s->u[n], vLoc = s->v[n], wLoc = s->w[n];
                const double u2 = uLoc * uLoc + vLoc * vLoc + wLoc * wLoc;

                m[0] = F[0] + F[1] + F[2] + F[3] + F[4] + F[5] + F[6] + F[7] + F[8] + F[9] +
                       F[10] + F[11] + F[12] + F[13] + F[14] + F[15] + F[16] + F[17] + F[18];
                m[1] = -30.0 * F[0] - 11.0 * (F[1] + F[2] + F[3] + F[4] + F[5] + F[6]) +
                       8.0 * (F[7] + F[8] + F[9] + F[10] + F[11] + F[12] + F[13] + F[14] +
                              F[15] + F[16] + F[17] + F[18]);
                m[2] = 12.0 * F[0] - 4.0 * (F[1] + F[2] + F[3] + F[4] + F[5] + F[6]) +
                       (F[7] + F[8] + F[9] + F[10] + F[11] + F[12] + F[13] + F[14] + F[15] +
                        F[16] + F[17] + F[18]);
                m[3] = F[1] - F[2] + F[7] - F[8] + F[9] - F[10] + F[11] - F[12] + F[13] - F[14];
                m[4] = -4.0 * F[1] + 4.0 * F[2] + F[7] - F[8] + F[9] - F[10] + F[11] - F[12] +
                       F[13] - F[14];
                m[5] = F[3] - F[4] + F[7] + F[8] - F[9] - F[10] + F[15] - F[16] + F[17] - F[18];
                m[6] = -4.0 * F[3] + 4.0 * F[4] + F[7] + F[8] - F[9] - F[10] + F[15] - F[16] +
                       F[17] - F[18];
                m[7] = F[5] - F[6] + F[11] + F[12] - F[13] - F[14] + F[15] + F[16] - F[17] - F[18];
                m[8] = -4.0 * F[5] + 4.0 * F[6] + F[11] + F[12] - F[13] - F[14] + F[15] + F[16] -
                       F[17] - F[18];
                m[9] = 2.0 * (F[1] + F[2]) - (F[3] + F[4] + F[5] + F[6]) +
                       (F[7] + F[8] + F[9] + F[10] + F[11] + F[12] + F[13] + F[14]) -
                       2.0 * (F[15] + F[16] + F[17] + F[18]);
                m[10] = -4.0 * (F[1] + F[2]) + 2.0 * (F[3] + F[4] + F[5] + F[6]) +
                        (F[7] + F[8] + F[9] + F[10] + F[11] + F[12] + F[13] + F[14]) -
                        2.0 * (F[15] + F[16] + F[17] + F[18]);
                m[11] = (F[3] + F[4]) - (F[5] + F[6]) + (F[7] + F[8] + F[9] + F[10]) -
                        (F[11] + F[12] + F[13] + F[14]);
                m[12] = -2.0 * (F[3] + F[4]) + 2.0 * (F[5] + F[6]) +
                        (F[7] + F[8] + F[9] + F[10]) - (F[11] + F[12] + F[13] + F[14]);
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

                const double FxLoc = 0.0;
                const double FyLoc = rhoLoc * gBeta * (s->T[n] - Tref);
                const double FzLoc = 0.0;
                const double uDotF = uLoc * FxLoc + vLoc * FyLoc + wLoc * FzLoc;
                src[0] = 0.0;
                src[1] = (1.0 - 0.5 * sr[1]) * 38.0 * uDotF;
                src[2] = (1.0 - 0.5 * sr[2]) * (-11.0) * uDotF;
                src[3] = (1.0 - 0.5 * sr[3]) * FxLoc;
                src[4] = (1.0 - 0.5 * sr[4]) * (-2.0 / 3.0) * FxLoc;
                src[5] = (1.0 - 0.5 * sr[5]) * FyLoc;
                src[6] = (1.0 - 0.5 * sr[6]) * (-2.0 / 3.0) * FyLoc;
                src[7] = (1.0 - 0.5 * sr[7]) * FzLoc;
                src[8] = (1.0 - 0.5 * sr[8]) * (-2.0 / 3.0) * FzLoc;
                src[9] = (1.0 - 0.5 * sr[9]) *
                         (4.0 * uLoc * FxLoc - 2.0 * vLoc * FyLoc - 2.0 * wLoc * FzLoc);
                src[10] = (1.0 - 0.5 * sr[10]) * (-2.0 * uLoc * FxLoc + vLoc * FyLoc + wLoc * FzLoc);
                src[11] = (1.0 - 0.5 * sr[11]) * (2.0 * vLoc * FyLoc - 2.0 * wLoc * FzLoc);
                src[12] = (1.0 - 0.5 * sr[12]) * (-vLoc * FyLoc + wLoc * FzLoc);
                src[13] = (1.0 - 0.5 * sr[13]) * (uLoc * FyLoc + vLoc * FxLoc);
                src[14] = (1.0 - 0.5 * sr[14]) * (vLoc * FzLoc + wLoc * FyLoc);
                src[15] = (1.0 - 0.5 * sr[15]) * (uLoc * FzLoc + wLoc * FxLoc);
                src[16] = 0.0;
                src[17] = 0.0;
                src[18] = 0.0;

                for (int a = 0; a < 19; ++a) mp[a] = m[a] - sr[a] * (m[a] - meq[a]) + src[a];

                double *P = s->fpost + HDX(s, i, j, k) * 19;
                P[0] = mp[0] / 19.0 - 5.0 * mp[1] / 399.0 + mp[2] / 21.0;
                P[1] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 + mp[3] / 10.0 -
                       mp[4] / 10.0 + mp[9] / 18.0 - mp[10] / 18.0;
                P[2] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 - mp[3] / 10.0 +
                       mp[4] / 10.0 + mp[9] / 18.0 - mp[10] / 18.0;
                P[3] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 + mp[5] / 10.0 -
                       mp[6] / 10.0 - mp[9] / 36.0 + mp[10] / 36.0 + mp[11] / 12.0 -
                       mp[12] / 12.0;
                P[4] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 - mp[5] / 10.0 +
                       mp[6] / 10.0 - mp[9] / 36.0 + mp[10] / 36.0 + mp[11] / 12.0 -
                       mp[12] / 12.0;
                P[5] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 + mp[7] / 10.0 -
                       mp[8] / 10.0 - mp[9] / 36.0 + mp[10] / 36.0 - mp[11] / 12.0 +
                       mp[12] / 12.0;
                P[6] = mp[0] / 19.0 - 11.0 * mp[1] / 2394.0 - mp[2] / 63.0 - mp[7] / 10.0 +
                       mp[8] / 10.0 - mp[9] / 36.0 + mp[10] / 36.0 - mp[11] / 12.0 +
                       mp[12] / 12.0;
                P[7] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[3] / 10.0 +
                       mp[4] / 40.0 + mp[5] / 10.0 + mp[6] / 40.0 + mp[9] / 36.0 +
                       mp[10] / 72.0 + mp[11] / 12.0 + mp[12] / 24.0 + mp[13] / 4.0 +
                       mp[16] / 8.0 - mp[17] / 8.0;
                P[8] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[3] / 10.0 -
                       mp[4] / 40.0 + mp[5] / 10.0 + mp[6] / 40.0 + mp[9] / 36.0 +
                       mp[10] / 72.0 + mp[11] / 12.0 + mp[12] / 24.0 - mp[13] / 4.0 -
                       mp[16] / 8.0 - mp[17] / 8.0;
                P[9] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[3] / 10.0 +
                       mp[4] / 40.0 - mp[5] / 10.0 - mp[6] / 40.0 + mp[9] / 36.0 +
                       mp[10] / 72.0 + mp[11] / 12.0 + mp[12] / 24.0 - mp[13] / 4.0 +
                       mp[16] / 8.0 + mp[17] / 8.0;
                P[10] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[3] / 10.0 -
                        mp[4] / 40.0 - mp[5] / 10.0 - mp[6] / 40.0 + mp[9] / 36.0 +
                        mp[10] / 72.0 + mp[11] / 12.0 + mp[12] / 24.0 + mp[13] / 4.0 -
                        mp[16] / 8.0 + mp[17] / 8.0;
                P[11] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[3] / 10.0 +
                        mp[4] / 40.0 + mp[7] / 10.0 + mp[8] / 40.0 + mp[9] / 36.0 +
                        mp[10] / 72.0 - mp[11] / 12.0 - mp[12] / 24.0 + mp[15] / 4.0 -
                        mp[16] / 8.0 + mp[18] / 8.0;
                P[12] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[3] / 10.0 -
                        mp[4] / 40.0 + mp[7] / 10.0 + mp[8] / 40.0 + mp[9] / 36.0 +
                        mp[10] / 72.0 - mp[11] / 12.0 - mp[12] / 24.0 - mp[15] / 4.0 +
                        mp[16] / 8.0 + mp[18] / 8.0;
                P[13] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[3] / 10.0 +
                        mp[4] / 40.0 - mp[7] / 10.0 - mp[8] / 40.0 + mp[9] / 36.0 +
                        mp[10] / 72.0 - mp[11] / 12.0 - mp[12] / 24.0 - mp[15] / 4.0 -
                        mp[16] / 8.0 - mp[18] / 8.0;
                P[14] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[3] / 10.0 -
                        mp[4] / 40.0 - mp[7] / 10.0 - mp[8] / 40.0 + mp[9] / 36.0 +
                        mp[10] / 72.0 - mp[11] / 12.0 - mp[12] / 24.0 + mp[15] / 4.0 +
                        mp[16] / 8.0 - mp[18] / 8.0;
                P[15] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[5] / 10.0 +
                        mp[6] / 40.0 + mp[7] / 10.0 + mp[8] / 40.0 - mp[9] / 18.0 -
                        mp[10] / 36.0 + mp[14] / 4.0 + mp[17] / 8.0 - mp[18] / 8.0;
                P[16] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[5] / 10.0 -
                        mp[6] / 40.0 + mp[7] / 10.0 + mp[8] / 40.0 - mp[9] / 18.0 -
                        mp[10] / 36.0 - mp[14] / 4.0 - mp[17] / 8.0 - mp[18] / 8.0;
                P[17] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 + mp[5] / 10.0 +
                        mp[6] / 40.0 - mp[7] / 10.0 - mp[8] / 40.0 - mp[9] / 18.0 -
                        mp[10] / 36.0 - mp[14] / 4.0 + mp[17] / 8.0 + mp[18] / 8.0;
                P[18] = mp[0] / 19.0 + 4.0 * mp[1] / 1197.0 + mp[2] / 252.0 - mp[5] / 10.0 -
                        mp[6] / 40.0 - mp[7] / 10.0 - mp[8] / 40.0 - mp[9] / 18.0 -
                        mp[10] / 36.0 + mp[14] / 4.0 - mp[17] / 8.0 + mp[18] / 8.0;
            }
}

/* streaming, 3DRB.F90:957-979 and streamingT, 3DRB.F90:1209-1232 (pull from the haloed
 * post-collision array; ghost reads return the zeros the wall sweeps then overwrite). */
static void streaming3d(orc_state *s, double *dst, const double *src, int q)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i)
                for (int a = 0; a < q; ++a)
                    dst[IDX(s, i, j, k) * q + a] =
                        src[HDX(s, i - EX3[a], j - EY3[a], k - EZ3[a]) * q + a];
}

/* bounceback, 3DRB.F90:987-1058.  Sweep order: x periodic, x no-slip, y no-slip, z periodic,
 * z no-slip.  The periodic copies keep the destination's transverse indices. */
static void bounceback3d(orc_state *s)
{
    double *f = s->f;
    const double *P = s->fpost;
    if (s->c.vel_x == ORC_VEL_PERIODIC) {
        for (int k = 0; k < NZ; ++k)
            for (int j = 0; j < NY; ++j)
                for (int a = 0; a < 19; ++a) {
                    if (EX3[a] == 1) f[IDX(s, 0, j, k) * 19 + a] = P[HDX(s, NX - 1, j, k) * 19 + a];
                    if (EX3[a] == -1) f[IDX(s, NX - 1, j, k) * 19 + a] = P[HDX(s, 0, j, k) * 19 + a];
                }
    } else {
        for (int k = 0; k < NZ; ++k)
            for (int j = 0; j < NY; ++j)
                for (int a = 0; a < 19; ++a) {
                    if (EX3[a] == 1) f[IDX(s, 0, j, k) * 19 + a] = P[HDX(s, 0, j, k) * 19 + OPP3[a]];
                    if (EX3[a] == -1)
                        f[IDX(s, NX - 1, j, k) * 19 + a] = P[HDX(s, NX - 1, j, k) * 19 + OPP3[a]];
                }
    }
    for (int k = 0; k < NZ; ++k)
        for (int i = 0; i < NX; ++i)
            for (int a = 0; a < 19; ++a) {
                if (EY3[a] == 1) f[IDX(s, i, 0, k) * 19 + a] = P[HDX(s, i, 0, k) * 19 + OPP3[a]];
                if (EY3[a] == -1)
                    f[IDX(s, i, NY - 1, k) * 19 + a] = P[HDX(s, i, NY - 1, k) * 19 + OPP3[a]];
            }
    if (s->c.vel_z == ORC_VEL_PERIODIC) {
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i)
                for (int a = 0; a < 19; ++a) {
                    if (EZ3[a] == 1) f[IDX(s, i, j, 0) * 19 + a] = P[HDX(s, i, j, NZ - 1) * 19 + a];
                    if (EZ3[a] == -1) f[IDX(s, i, j, NZ - 1) * 19 + a] = P[HDX(s, i, j, 0) * 19 + a];
                }
    } else {
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i)
                for (int a = 0; a < 19; ++a) {
                    if (EZ3[a] == 1) f[IDX(s, i, j, 0) * 19 + a] = P[HDX(s, i, j, 0) * 19 + OPP3[a]];
                    if (EZ3[a] == -1)
                        f[IDX(s, i, j, NZ - 1) * 19 + a] = P[HDX(s, i, j, NZ - 1) * 19 + OPP3[a]];
                }
    }
}

/* macro, 3DRB.F90:1066-1098: the half force uses the new rho and the old T. */
static void macro3d(orc_state *s)
{
    const double gBeta = s->p.gbeta, Tref = s->p.Tref;
#pragma omp parallel for schedule(static)
    for (long n = 0; n < s->n; ++n) {
        const double *F = s->f + n * 19;
        double rho = F[0] + F[1] + F[2] + F[3] + F[4] + F[5] + F[6] + F[7] + F[8] + F[9] +
                     F[10] + F[11] + F[12] + F[13] + F[14] + F[15] + F[16] + F[17] + F[18];
        s->rho[n] = rho;
        double FyLoc = rho * gBeta * (s->T[n] - Tref);
        s->u[n] = (F[1] - F[2] + F[7] - F[8] + F[9] - F[10] + F[11] - F[12] + F[13] - F[14]) / rho;
        s->v[n] = (F[3] - F[4] + F[7] + F[8] - F[9] - F[10] + F[15] - F[16] + F[17] - F[18] +
                   0.5 * FyLoc) / rho;
        s->w[n] = (F[5] - F[6] + F[11] + F[12] - F[13] - F[14] + F[15] + F[16] - F[17] - F[18]) / rho;
    }
}

/* collisionT, 3DRB.F90:1106-1198 (non-legacy branch). */
static void collisionT3d(orc_state *s)
{
    const double Qk = s->p.Qk, Qnu = s->p.Qnu, SG = s->p.SG;
    const int useG = s->c.use_g;
#pragma omp parallel for collapse(2) schedule(static)
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                const long n = IDX(s, i, j, k);
                const double *G = s->g + n * 7;
                double nm[7], neq[7], q[7], np[7];
                const double Tl = s->T[n];
                const double Bx = s->u[n] * Tl, By = s->v[n] * Tl, Bz = s->w[n] * Tl;
                double dBx = 0.0, dBy = 0.0, dBz = 0.0;
                if (useG) {
                    dBx = Bx - s->bx[n];
                    dBy = By - s->by[n];
                    dBz = Bz - s->bz[n];
                    s->bx[n] = Bx;
                    s->by[n] = By;
                    s->bz[n] = Bz;
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
                neq[4] = s->c.legacy ? s->p.paraA * Tl : -0.75 * Tl;   /* 3DRB.F90:1153-1157 */
                neq[5] = 0.0;
                neq[6] = 0.0;
                q[0] = 0.0; q[1] = Qk; q[2] = Qk; q[3] = Qk; q[4] = Qnu; q[5] = Qnu; q[6] = Qnu;
                np[0] = nm[0] - q[0] * (nm[0] - neq[0]);
                np[1] = nm[1] - q[1] * (nm[1] - neq[1]) + SG * dBx;
                np[2] = nm[2] - q[2] * (nm[2] - neq[2]) + SG * dBy;
                np[3] = nm[3] - q[3] * (nm[3] - neq[3]) + SG * dBz;
                np[4] = nm[4] - q[4] * (nm[4] - neq[4]);
                np[5] = nm[5] - q[5] * (nm[5] - neq[5]);
                np[6] = nm[6] - q[6] * (nm[6] - neq[6]);
                double *P = s->gpost + HDX(s, i, j, k) * 7;
                P[0] = np[0] / 7.0 - np[4] / 7.0;
                P[1] = np[0] / 7.0 + np[1] / 2.0 + np[4] / 42.0 + np[5] / 6.0;
                P[2] = np[0] / 7.0 - np[1] / 2.0 + np[4] / 42.0 + np[5] / 6.0;
                P[3] = np[0] / 7.0 + np[2] / 2.0 + np[4] / 42.0 - np[5] / 12.0 + np[6] / 4.0;
                P[4] = np[0] / 7.0 - np[2] / 2.0 + np[4] / 42.0 - np[5] / 12.0 + np[6] / 4.0;
                P[5] = np[0] / 7.0 + np[3] / 2.0 + np[4] / 42.0 - np[5] / 12.0 - np[6] / 4.0;
                P[6] = np[0] / 7.0 - np[3] / 2.0 + np[4] / 42.0 - np[5] / 12.0 - np[6] / 4.0;
            }
}

/* bouncebackT, 3DRB.F90:1243-1351 (non-legacy branch).  Sweep order: x periodic, x const-T,
 * x adiabatic, y adiabatic, y const-T, z periodic, z adiabatic.  Hot wall at the low index. */
static void bouncebackT3d(orc_state *s)
{
    double *g = s->g;
    const double *P = s->gpost;
    const double *wT = s->omegaT;
    const double Th = s->c.thot, Tc = s->c.tcold;
    for (int k = 0; k < NZ; ++k)
        for (int j = 0; j < NY; ++j)
            for (int a = 0; a < 7; ++a) {
                long lo = IDX(s, 0, j, k) * 7 + a, hi = IDX(s, NX - 1, j, k) * 7 + a;
                if (s->c.temp_x == ORC_T_PERIODIC) {
                    if (EX3[a] == 1) g[lo] = P[HDX(s, NX - 1, j, k) * 7 + a];
                    if (EX3[a] == -1) g[hi] = P[HDX(s, 0, j, k) * 7 + a];
                } else if (s->c.temp_x == ORC_T_CONST) {
                    if (EX3[a] == 1) g[lo] = -P[HDX(s, 0, j, k) * 7 + OPP7[a]] + abb3(s, wT[a], Th);
                    if (EX3[a] == -1) g[hi] = -P[HDX(s, NX - 1, j, k) * 7 + OPP7[a]] + abb3(s, wT[a], Tc);
                } else {
                    if (EX3[a] == 1) g[lo] = P[HDX(s, 0, j, k) * 7 + OPP7[a]];
                    if (EX3[a] == -1) g[hi] = P[HDX(s, NX - 1, j, k) * 7 + OPP7[a]];
                }
            }
    for (int k = 0; k < NZ; ++k)
        for (int i = 0; i < NX; ++i)
            for (int a = 0; a < 7; ++a) {
                long lo = IDX(s, i, 0, k) * 7 + a, hi = IDX(s, i, NY - 1, k) * 7 + a;
                if (s->c.temp_y == ORC_T_ADIABATIC) {
                    if (EY3[a] == 1) g[lo] = P[HDX(s, i, 0, k) * 7 + OPP7[a]];
                    if (EY3[a] == -1) g[hi] = P[HDX(s, i, NY - 1, k) * 7 + OPP7[a]];
                } else if (s->c.temp_y == ORC_T_CONST) {
                    if (EY3[a] == 1) g[lo] = -P[HDX(s, i, 0, k) * 7 + OPP7[a]] + abb3(s, wT[a], Th);
                    if (EY3[a] == -1) g[hi] = -P[HDX(s, i, NY - 1, k) * 7 + OPP7[a]] + abb3(s, wT[a], Tc);
                }
            }
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i)
            for (int a = 0; a < 7; ++a) {
                long lo = IDX(s, i, j, 0) * 7 + a, hi = IDX(s, i, j, NZ - 1) * 7 + a;
                if (s->c.temp_z == ORC_T_PERIODIC) {
                    if (EZ3[a] == 1) g[lo] = P[HDX(s, i, j, NZ - 1) * 7 + a];
                    if (EZ3[a] == -1) g[hi] = P[HDX(s, i, j, 0) * 7 + a];
                } else {
                    if (EZ3[a] == 1) g[lo] = P[HDX(s, i, j, 0) * 7 + OPP7[a]];
                    if (EZ3[a] == -1) g[hi] = P[HDX(s, i, j, NZ - 1) * 7 + OPP7[a]];
                }
            }
}

/* macroT, 3DRB.F90:1362-1381 */
static void macroT3d(orc_state *s)
{
#pragma omp parallel for schedule(static)
    for (long n = 0; n < s->n; ++n) {
        const double *G = s->g + n * 7;
        s->T[n] = G[0] + G[1] + G[2] + G[3] + G[4] + G[5] + G[6];
    }
}

/* ========================================================================================= */
/* 2D: D2Q9 MRT flow + D2Q5 MRT temperature                                                  */
/* ========================================================================================= */

/* collision, 2DRB.F90:675-766.  Stores Fx, Fy (old rho, old T) for macro. */
static void collision2d(orc_state *s)
{
    const orc_params *p = &s->p;
    const double gBeta = p->gbeta, Tref = p->Tref, Snu = p->Snu, Sq = p->Sq;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i) {
            const long n = IDX(s, i, j, 0);
            const double *F = s->f + n * 9;
            double m[9], meq[9], mp[9], sr[9], src[9];
            const double rho = s->rho[n], u = s->u[n], v = s->v[n];
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

            sr[0] = 0.0; sr[1] = Snu; sr[2] = Snu; sr[3] = 0.0; sr[4] = Sq; sr[5] = 0.0;
            sr[6] = Sq; sr[7] = Snu; sr[8] = Snu;

            const double Fx = 0.0;
            const double Fy = rho * gBeta * (s->T[n] - Tref);
            s->Fx[n] = Fx;
            s->Fy[n] = Fy;

            src[0] = 0.0;
            src[1] = (6.0 - 3.0 * sr[1]) * (u * Fx + v * Fy);
            src[2] = -(6.0 - 3.0 * sr[2]) * (u * Fx + v * Fy);
            src[3] = (1.0 - 0.5 * sr[3]) * Fx;
            src[4] = -(1.0 - 0.5 * sr[4]) * Fx;
            src[5] = (1.0 - 0.5 * sr[5]) * Fy;
            src[6] = -(1.0 - 0.5 * sr[6]) * Fy;
            src[7] = (2.0 - sr[7]) * (u * Fx - v * Fy);
            src[8] = (1.0 - 0.5 * sr[8]) * (u * Fy + v * Fx);

            for (int a = 0; a < 9; ++a) mp[a] = m[a] - sr[a] * (m[a] - meq[a]) + src[a];

            double *P = s->fpost + HDX(s, i, j, -1) * 9;
            P[0] = mp[0] / 9.0 - mp[1] / 9.0 + mp[2] / 9.0;
            P[1] = mp[0] / 9.0 - mp[1] / 36.0 - mp[2] / 18.0 + mp[3] / 6.0 - mp[4] / 6.0 + mp[7] / 4.0;
            P[2] = mp[0] / 9.0 - mp[1] / 36.0 - mp[2] / 18.0 + mp[5] / 6.0 - mp[6] / 6.0 - mp[7] / 4.0;
            P[3] = mp[0] / 9.0 - mp[1] / 36.0 - mp[2] / 18.0 - mp[3] / 6.0 + mp[4] / 6.0 + mp[7] / 4.0;
            P[4] = mp[0] / 9.0 - mp[1] / 36.0 - mp[2] / 18.0 - mp[5] / 6.0 + mp[6] / 6.0 - mp[7] / 4.0;
            P[5] = mp[0] / 9.0 + mp[1] / 18.0 + mp[2] / 36.0 + mp[3] / 6.0 + mp[4] / 12.0 +
                   mp[5] / 6.0 + mp[6] / 12.0 + mp[8] / 4.0;
            P[6] = mp[0] / 9.0 + mp[1] / 18.0 + mp[2] / 36.0 - mp[3] / 6.0 - mp[4] / 12.0 +
                   mp[5] / 6.0 + mp[6] / 12.0 - mp[8] / 4.0;
            P[7] = mp[0] / 9.0 + mp[1] / 18.0 + mp[2] / 36.0 - mp[3] / 6.0 - mp[4] / 12.0 -
                   mp[5] / 6.0 - mp[6] / 12.0 + mp[8] / 4.0;
            P[8] = mp[0] / 9.0 + mp[1] / 18.0 + mp[2] / 36.0 + mp[3] / 6.0 + mp[4] / 12.0 -
                   mp[5] / 6.0 - mp[6] / 12.0 - mp[8] / 4.0;
        }
}

/* streaming, 2DRB.F90:777-798; streamingT, 2DRB.F90:993-1014 */
static void streaming2d(orc_state *s, double *dst, const double *src, int q)
{
#pragma omp parallel for schedule(static)
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i)
            for (int a = 0; a < q; ++a)
                dst[IDX(s, i, j, 0) * q + a] = src[HDX(s, i - EX2[a], j - EY2[a], -1) * q + a];
}

/* bounceback, 2DRB.F90:811-866: x periodic (same row) or x no-slip, then y no-slip. */
static void bounceback2d(orc_state *s)
{
    double *f = s->f;
    const double *P = s->fpost;
#define F2(i, j, a) f[IDX(s, (i), (j), 0) * 9 + (a)]
#define P2(i, j, a) P[HDX(s, (i), (j), -1) * 9 + (a)]
    for (int j = 0; j < NY; ++j) {
        if (s->c.vel_x == ORC_VEL_PERIODIC) {
            F2(0, j, 1) = P2(NX - 1, j, 1);
            F2(0, j, 5) = P2(NX - 1, j, 5);
            F2(0, j, 8) = P2(NX - 1, j, 8);
            F2(NX - 1, j, 3) = P2(0, j, 3);
            F2(NX - 1, j, 6) = P2(0, j, 6);
            F2(NX - 1, j, 7) = P2(0, j, 7);
        } else {
            F2(0, j, 1) = P2(0, j, 3);
            F2(0, j, 5) = P2(0, j, 7);
            F2(0, j, 8) = P2(0, j, 6);
            F2(NX - 1, j, 3) = P2(NX - 1, j, 1);
            F2(NX - 1, j, 6) = P2(NX - 1, j, 8);
            F2(NX - 1, j, 7) = P2(NX - 1, j, 5);
        }
    }
    for (int i = 0; i < NX; ++i) {
        F2(i, 0, 2) = P2(i, 0, 4);
        F2(i, 0, 5) = P2(i, 0, 7);
        F2(i, 0, 6) = P2(i, 0, 8);
        F2(i, NY - 1, 4) = P2(i, NY - 1, 2);
        F2(i, NY - 1, 7) = P2(i, NY - 1, 5);
        F2(i, NY - 1, 8) = P2(i, NY - 1, 6);
    }
#undef F2
#undef P2
}

/* macro, 2DRB.F90:878-894: half force from the Fx, Fy stored by collision. */
static void macro2d(orc_state *s)
{
#pragma omp parallel for schedule(static)
    for (long n = 0; n < s->n; ++n) {
        const double *F = s->f + n * 9;
        double rho = F[0] + F[1] + F[2] + F[3] + F[4] + F[5] + F[6] + F[7] + F[8];
        s->rho[n] = rho;
        s->u[n] = (F[1] - F[3] + F[5] - F[6] - F[7] + F[8] + 0.5 * s->Fx[n]) / rho;
        s->v[n] = (F[2] - F[4] + F[5] + F[6] - F[7] - F[8] + 0.5 * s->Fy[n]) / rho;
    }
}

/* collisionT, 2DRB.F90:906-980 (non-legacy branch). */
static void collisionT2d(orc_state *s)
{
    const double Qk = s->p.Qk, Qnu = s->p.Qnu, SG = s->p.SG;
    const int useG = s->c.use_g;
#pragma omp parallel for schedule(static)
    for (int j = 0; j < NY; ++j)
        for (int i = 0; i < NX; ++i) {
            const long n = IDX(s, i, j, 0);
            const double *G = s->g + n * 5;
            double nm[5], neq[5], q[5], np[5];
            const double Tl = s->T[n];
            const double Bx = s->u[n] * Tl, By = s->v[n] * Tl;
            double dBx = 0.0, dBy = 0.0;
            if (useG) {
                dBx = Bx - s->bx[n];
                dBy = By - s->by[n];
                s->bx[n] = Bx;
                s->by[n] = By;
            }
            nm[0] = G[0] + G[1] + G[2] + G[3] + G[4];
            nm[1] = G[1] - G[3];
            nm[2] = G[2] - G[4];
            nm[3] = -4.0 * G[0] + G[1] + G[2] + G[3] + G[4];
            nm[4] = G[1] - G[2] + G[3] - G[4];
            neq[0] = Tl;
            neq[1] = Tl * s->u[n];
            neq[2] = Tl * s->v[n];
            neq[3] = s->c.legacy ? Tl * s->p.paraA : Tl * (-2.0 / 3.0);   /* 2DRB.F90:949-953 */
            neq[4] = 0.0;
            q[0] = 0.0; q[1] = Qk; q[2] = Qk; q[3] = Qnu; q[4] = Qnu;
            np[0] = nm[0] - q[0] * (nm[0] - neq[0]);
            np[1] = nm[1] - q[1] * (nm[1] - neq[1]) + SG * dBx;
            np[2] = nm[2] - q[2] * (nm[2] - neq[2]) + SG * dBy;
            np[3] = nm[3] - q[3] * (nm[3] - neq[3]);
            np[4] = nm[4] - q[4] * (nm[4] - neq[4]);
            double *P = s->gpost + HDX(s, i, j, -1) * 5;
            P[0] = 0.2 * np[0] - 0.2 * np[3];
            P[1] = 0.2 * np[0] + 0.5 * np[1] + 0.05 * np[3] + 0.25 * np[4];
            P[2] = 0.2 * np[0] + 0.5 * np[2] + 0.05 * np[3] - 0.25 * np[4];
            P[3] = 0.2 * np[0] - 0.5 * np[1] + 0.05 * np[3] + 0.25 * np[4];
            P[4] = 0.2 * np[0] - 0.5 * np[2] + 0.05 * np[3] - 0.25 * np[4];
        }
}

/* bouncebackT, 2DRB.F90:1026-1107 (non-legacy branch). */
static void bouncebackT2d(orc_state *s)
{
    double *g = s->g;
    const double *P = s->gpost;
    const double *wT = s->omegaT;
    const double Th = s->c.thot, Tc = s->c.tcold;
#define G2(i, j, a) g[IDX(s, (i), (j), 0) * 5 + (a)]
#define Q2(i, j, a) P[HDX(s, (i), (j), -1) * 5 + (a)]
    for (int j = 0; j < NY; ++j) {
        if (s->c.temp_x == ORC_T_PERIODIC) {
            G2(0, j, 1) = Q2(NX - 1, j, 1);
            G2(NX - 1, j, 3) = Q2(0, j, 3);
        } else if (s->c.temp_x == ORC_T_CONST) {
            G2(0, j, 1) = -Q2(0, j, 3) + abb2(s, wT[3], Th);
            G2(NX - 1, j, 3) = -Q2(NX - 1, j, 1) + abb2(s, wT[1], Tc);
        } else {
            G2(0, j, 1) = Q2(0, j, 3);
            G2(NX - 1, j, 3) = Q2(NX - 1, j, 1);
        }
    }
    for (int i = 0; i < NX; ++i) {
        if (s->c.temp_y == ORC_T_ADIABATIC) {
            G2(i, 0, 2) = Q2(i, 0, 4);
            G2(i, NY - 1, 4) = Q2(i, NY - 1, 2);
        } else if (s->c.temp_y == ORC_T_CONST) {
            G2(i, 0, 2) = -Q2(i, 0, 4) + abb2(s, wT[4], Th);
            G2(i, NY - 1, 4) = -Q2(i, NY - 1, 2) + abb2(s, wT[2], Tc);
        }
    }
#undef G2
#undef Q2
}

/* macroT, 2DRB.F90:1120-1134 */
static void macroT2d(orc_state *s)
{
#pragma omp parallel for schedule(static)
    for (long n = 0; n < s->n; ++n) {
        const double *G = s->g + n * 5;
        s->T[n] = G[0] + G[1] + G[2] + G[3] + G[4];
    }
}

/* One pass of the reference time loop body (3DRB.F90:227-243, 2DRB.F90:238-254). */
void orc_step(orc_state *s, int nsteps)
{
    for (int it = 0; it < nsteps; ++it) {
        s->itc += 1;
        if (s->c.dims == 3) {
            collision3d(s);
            streaming3d(s, s->f, s->fpost, 19);
            bounceback3d(s);
            macro3d(s);
            collisionT3d(s);
            streaming3d(s, s->g, s->gpost, 7);
            bouncebackT3d(s);
            macroT3d(s);
        } else {
            collision2d(s);
            streaming2d(s, s->f, s->fpost, 9);
            bounceback2d(s);
            macro2d(s);
            collisionT2d(s);
            streaming2d(s, s->g, s->gpost, 5);
            bouncebackT2d(s);
            macroT2d(s);
        }
    }
}

void orc_get(const orc_state *s, double *f, double *g, double *rho, double *u, double *v,
             double *w, double *T, double *bx, double *by, double *bz)
{
    size_t n = (size_t)s->n;
    if (f) memcpy(f, s->f, n * s->qf * sizeof(double));
    if (g) memcpy(g, s->g, n * s->qt * sizeof(double));
    if (rho) memcpy(rho, s->rho, n * sizeof(double));
    if (u) memcpy(u, s->u, n * sizeof(double));
    if (v) memcpy(v, s->v, n * sizeof(double));
    if (w) memcpy(w, s->w, n * sizeof(double));
    if (T) memcpy(T, s->T, n * sizeof(double));
    if (bx) memcpy(bx, s->bx, n * sizeof(double));
    if (by) memcpy(by, s->by, n * sizeof(double));
    if (bz) memcpy(bz, s->bz, n * sizeof(double));
}

/* reconstruct_macro_from_fg, 3DRB.F90:1392-1463 and 2DRB.F90:1144-1200: macroT, then rho,
 * momentum and three fixed-point passes of u = (sum e f + F/2)/rho, then B_prev = u T. */
int orc_set_fg(orc_state *s, const double *f, const double *g)
{
    const double gBeta = s->p.gbeta, Tref = s->p.Tref;
    memcpy(s->f, f, (size_t)s->n * s->qf * sizeof(double));
    memcpy(s->g, g, (size_t)s->n * s->qt * sizeof(double));
    if (s->c.dims == 3) macroT3d(s);
    else macroT2d(s);
    int bad = 0;
    for (long n = 0; n < s->n; ++n) {
        const double *F = s->f + n * s->qf;
        double rho, momx, momy, momz = 0.0;
        if (s->c.dims == 3) {
            rho = F[0] + F[1] + F[2] + F[3] + F[4] + F[5] + F[6] + F[7] + F[8] + F[9] + F[10] +
                  F[11] + F[12] + F[13] + F[14] + F[15] + F[16] + F[17] + F[18];
            momx = F[1] - F[2] + F[7] - F[8] + F[9] - F[10] + F[11] - F[12] + F[13] - F[14];
            momy = F[3] - F[4] + F[7] + F[8] - F[9] - F[10] + F[15] - F[16] + F[17] - F[18];
            momz = F[5] - F[6] + F[11] + F[12] - F[13] - F[14] + F[15] + F[16] - F[17] - F[18];
        } else {
            rho = F[0] + F[1] + F[2] + F[3] + F[4] + F[5] + F[6] + F[7] + F[8];
            momx = F[1] - F[3] + F[5] - F[6] - F[7] + F[8];
            momy = F[2] - F[4] + F[5] + F[6] - F[7] - F[8];
        }
        s->rho[n] = rho;
        if (rho > 0.0) {
            s->u[n] = momx / rho;
            s->v[n] = momy / rho;
            s->w[n] = momz / rho;
            for (int iter = 0; iter < 3; ++iter) {
                double FxLoc = 0.0, FyLoc = rho * gBeta * (s->T[n] - Tref), FzLoc = 0.0;
                s->u[n] = (momx + 0.5 * FxLoc) / rho;
                s->v[n] = (momy + 0.5 * FyLoc) / rho;
                s->w[n] = (momz + 0.5 * FzLoc) / rho;
                if (s->c.dims == 2) {
                    s->Fx[n] = FxLoc;
                    s->Fy[n] = FyLoc;
                }
            }
        } else {
            bad = 1;
            s->u[n] = s->v[n] = s->w[n] = 0.0;
            s->Fx[n] = s->Fy[n] = 0.0;
        }
        if (s->c.use_g) {
            s->bx[n] = s->u[n] * s->T[n];
            s->by[n] = s->v[n] * s->T[n];
            s->bz[n] = s->w[n] * s->T[n];
        } else if (s->c.dims == 3) {
            s->bx[n] = s->by[n] = s->bz[n] = 0.0;
        }
    }
    return bad ? -1 : 0;
}

/* check, 3DRB.F90:1475-1528 (3D keeps a 1e-30 guard; 2D, 2DRB.F90:1213-1260, has none). */
void orc_check(orc_state *s, double *error_u, double *error_t)
{
    ksum e1 = {0, 0}, e2 = {0, 0}, e5 = {0, 0}, e6 = {0, 0};
    for (int k = 0; k < NZ; ++k) {
        ksum p1 = {0, 0}, p2 = {0, 0}, p5 = {0, 0}, p6 = {0, 0};
        for (int j = 0; j < NY; ++j)
            for (int i = 0; i < NX; ++i) {
                long n = IDX(s, i, j, k);
                double du = s->u[n] - s->up[n], dv = s->v[n] - s->vp[n], dw = s->w[n] - s->wp[n];
                double a = du * du + dv * dv, b = s->u[n] * s->u[n] + s->v[n] * s->v[n];
                if (s->c.dims == 3) {
                    a = a + dw * dw;
                    b = b + s->w[n] * s->w[n];
                }
                ks_add(&p1, a);
                ks_add(&p2, b);
                ks_add(&p5, fabs(s->T[n] - s->Tp[n]));
                ks_add(&p6, fabs(s->T[n]));
                s->up[n] = s->u[n];
                s->vp[n] = s->v[n];
                s->wp[n] = s->w[n];
                s->Tp[n] = s->T[n];
            }
        ks_merge(&e1, p1);
        ks_merge(&e2, p2);
        ks_merge(&e5, p5);
        ks_merge(&e6, p6);
    }
    double error1 = ks_val(e1), error2 = ks_val(e2), error5 = ks_val(e5), error6 = ks_val(e6);
    if (s->c.dims == 3) {
        *error_u = (error2 > 1.0e-30) ? sqrt(error1) / sqrt(error2) : sqrt(error1);
        *error_t = (error6 > 1.0e-30) ? error5 / error6 : error5;
    } else {
        *error_u = sqrt(error1) / sqrt(error2);
        *error_t = error5 / error6;
    }
}

/* fit_adiabatic_wall_T4, 3DRB.F90:2212-2241: least-squares T = a + b (y - y0)^2, returns a. */
void orc_fit_adiabatic_wall_T4(double y0, const double *y, const double *tt, double *t_wall)
{
    double sq[4], S0 = 4.0, S1 = 0.0, S2 = 0.0, B0 = 0.0, B1 = 0.0;
    for (int k = 0; k < 4; ++k) sq[k] = (y[k] - y0) * (y[k] - y0);
    for (int k = 0; k < 4; ++k) {
        S1 = S1 + sq[k];
        S2 = S2 + sq[k] * sq[k];
        B0 = B0 + tt[k];
        B1 = B1 + tt[k] * sq[k];
    }
    double D = S0 * S2 - S1 * S1;
    *t_wall = (B0 * S2 - B1 * S1) / D;
}

/* fit_parabola_ls5, 3DRB.F90:2252-2334: 5-point least-squares parabola, vertex if interior. */
void orc_fit_parabola_ls5(const double *y, const double *f, int mode, double *fstar,
                          double *ystar)
{
    const double epsD = 1.0e-20, epsA = 1.0e-14;
    int kbest = 0;
    for (int k = 1; k < 5; ++k) {
        if (mode == 1) {
            if (f[k] > f[kbest]) kbest = k;
        } else {
            if (f[k] < f[kbest]) kbest = k;
        }
    }
    double S0 = 0, S1 = 0, S2 = 0, S3 = 0, S4 = 0, F0 = 0, F1 = 0, F2 = 0;
    for (int k = 0; k < 5; ++k) {
        S0 = S0 + 1.0;
        S1 = S1 + y[k];
        S2 = S2 + y[k] * y[k];
        S3 = S3 + y[k] * y[k] * y[k];
        S4 = S4 + y[k] * y[k] * y[k] * y[k];
        F0 = F0 + f[k];
        F1 = F1 + f[k] * y[k];
        F2 = F2 + f[k] * y[k] * y[k];
    }
    double D = S0 * (S2 * S4 - S3 * S3) - S1 * (S1 * S4 - S2 * S3) + S2 * (S1 * S3 - S2 * S2);
    double DA = F0 * (S2 * S4 - S3 * S3) - S1 * (F1 * S4 - S3 * F2) + S2 * (F1 * S3 - S2 * F2);
    double DB = S0 * (F1 * S4 - S3 * F2) - F0 * (S1 * S4 - S2 * S3) + S2 * (S1 * F2 - F1 * S2);
    double DC = S0 * (S2 * F2 - F1 * S3) - S1 * (S1 * F2 - F1 * S2) + F0 * (S1 * S3 - S2 * S2);
    if (fabs(D) <= epsD) {
        *ystar = y[kbest];
        *fstar = f[kbest];
        return;
    }
    double A = DA / D, B = DB / D, C = DC / D;
    double ymin = y[0], ymax = y[0];
    for (int k = 1; k < 5; ++k) {
        if (y[k] < ymin) ymin = y[k];
        if (y[k] > ymax) ymax = y[k];
    }
    if (fabs(C) <= epsA || (mode == 1 && C >= 0.0) || (mode == -1 && C <= 0.0)) {
        *ystar = y[kbest];
        *fstar = f[kbest];
        return;
    }
    double ys = -B / (2.0 * C);
    if (ys < ymin || ys > ymax) {
        *ystar = y[kbest];
        *fstar = f[kbest];
        return;
    }
    *ystar = ys;
    *fstar = A + B * ys + C * ys * ys;
}

/* Side-heated cell (hot wall x = 0, cold wall x = L, L = nx): the SideHeatedCell branch of
 * calNuRe (3DRB.F90:1732-1741, 2DRB.F90:1631-1640; Nu_vol from u T), SideHeatedcalc_Nu_global
 * (3DRB.F90:1789-1828, 2DRB.F90:1686-1734) and SideHeatedcalc_Nu_wall_avg (3DRB.F90:1889-2038,
 * 2DRB.F90:1746-1941): the hot-wall profile runs along y (averaged over z), its corner values
 * come from adiabatic-wall fits at y = 0 and y = Ly, and the minimum's 5-point window is
 * jmin-4 .. jmin (0 .. 4 when jmin < 4). */
static void diagnostics_side_heated(const orc_state *s, orc_diag *d)
{
    const orc_params *p = &s->p;
    const double Th = s->c.thot, Tc = s->c.tcold, Tref = p->Tref, L = p->L;
    const double dx = 1.0 / L, deltaT = Th - Tc, coef = L / p->diffusivity;
    const double N = (double)s->n;
    const int nx = NX, ny = NY, nz = NZ;

    ksum snu = {0, 0}, sre = {0, 0}, sqx = {0, 0}, sm = {0, 0}, sh = {0, 0}, ske = {0, 0};
    for (int k = 0; k < nz; ++k) {
        ksum a = {0, 0}, b = {0, 0}, c = {0, 0}, m = {0, 0}, h = {0, 0}, e = {0, 0};
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long n = IDX(s, i, j, k);
                double u = s->u[n], v = s->v[n], w = s->w[n], T = s->T[n];
                ks_add(&a, u * T);
                double q2 = u * u + v * v;
                if (s->c.dims == 3) q2 = q2 + w * w;
                ks_add(&b, q2);
                double dT;
                if (i == 0) dT = (-3.0 * T - s->T[IDX(s, 1, j, k)] + 4.0 * Th) / (3.0 * dx);
                else if (i == nx - 1)
                    dT = (-4.0 * Tc + 3.0 * T + s->T[IDX(s, nx - 2, j, k)]) / (3.0 * dx);
                else dT = (s->T[IDX(s, i - 1, j, k)] - s->T[IDX(s, i + 1, j, k)]) / (2.0 * dx);
                ks_add(&c, coef * u * (T - Tref) + dT);
                ks_add(&m, s->rho[n]);
                ks_add(&h, T);
                ks_add(&e, 0.5 * s->rho[n] * q2);
            }
        ks_merge(&snu, a);
        ks_merge(&sre, b);
        ks_merge(&sqx, c);
        ks_merge(&sm, m);
        ks_merge(&sh, h);
        ks_merge(&ske, e);
    }
    d->nu_vol = ks_val(snu) / N * L / p->diffusivity + 1.0;
    d->re_vol = sqrt(ks_val(sre) / N) * L / p->viscosity;
    d->nu_global = (ks_val(sqx) / N) / deltaT;
    d->mass = ks_val(sm);
    d->heat = ks_val(sh);
    d->kinetic = ks_val(ske);

    /* Hot-wall profile along y: per-j averages over z of T(1,j,k) and of the wall flux. */
    double *Nu_ext = (double *)calloc((size_t)ny + 2, sizeof(double));
    double *Tleft = (double *)calloc((size_t)ny + 2, sizeof(double));
    ksum shot = {0, 0}, scold = {0, 0}, smid = {0, 0};
    for (int j = 0; j < ny; ++j) {
        ksum tl = {0, 0}, nl = {0, 0};
        for (int k = 0; k < nz; ++k) {
            ks_add(&tl, s->T[IDX(s, 0, j, k)]);
            double qx = (8.0 * Th - 9.0 * s->T[IDX(s, 0, j, k)] + s->T[IDX(s, 1, j, k)]) / (3.0 * dx);
            ks_add(&nl, qx / deltaT);
            double qc = (-8.0 * Tc + 9.0 * s->T[IDX(s, nx - 1, j, k)] - s->T[IDX(s, nx - 2, j, k)]) /
                        (3.0 * dx);
            ks_add(&scold, qc / deltaT);
            double mid;
            if (nx % 2 == 1) {
                int im = (nx + 1) / 2 - 1;
                long nm = IDX(s, im, j, k);
                mid = (coef * s->u[nm] * (s->T[nm] - Tref) +
                       (s->T[IDX(s, im - 1, j, k)] - s->T[IDX(s, im + 1, j, k)]) / (2.0 * dx)) /
                      deltaT;
            } else {
                int il = nx / 2 - 1, ir = il + 1;
                long nl_ = IDX(s, il, j, k), nr_ = IDX(s, ir, j, k);
                mid = (coef * 0.5 * (s->u[nl_] * (s->T[nl_] - Tref) + s->u[nr_] * (s->T[nr_] - Tref)) +
                       (s->T[nl_] - s->T[nr_]) / dx) /
                      deltaT;
            }
            ks_add(&smid, mid);
        }
        Tleft[j + 1] = ks_val(tl) / (double)nz;
        Nu_ext[j + 1] = ks_val(nl) / (double)nz;
        ks_add(&shot, Nu_ext[j + 1]);
    }
    d->nu_hot = ks_val(shot) / (double)ny;
    d->nu_cold = ks_val(scold) / ((double)ny * nz);
    d->nu_middle = ks_val(smid) / ((double)ny * nz);

    const double *yp = s->yp;
    double yfit[4], Tfit[4], Twb, Twt;
    for (int m = 0; m < 4; ++m) {
        yfit[m] = yp[1 + m];
        Tfit[m] = Tleft[1 + m];
    }
    orc_fit_adiabatic_wall_T4(0.0, yfit, Tfit, &Twb);
    Nu_ext[0] = (2.0 * (Th - Twb) / dx) / deltaT;
    for (int m = 0; m < 4; ++m) {
        yfit[m] = yp[ny - 3 + m];
        Tfit[m] = Tleft[ny - 3 + m];
    }
    orc_fit_adiabatic_wall_T4(yp[ny + 1], yfit, Tfit, &Twt);
    Nu_ext[ny + 1] = (2.0 * (Th - Twt) / dx) / deltaT;

    int jmax = 0, jmin = 0;
    double vmax = Nu_ext[0], vmin = Nu_ext[0];
    for (int j = 1; j <= ny + 1; ++j) {
        if (Nu_ext[j] > vmax) { vmax = Nu_ext[j]; jmax = j; }
        if (Nu_ext[j] < vmin) { vmin = Nu_ext[j]; jmin = j; }
    }
    for (int pass = 0; pass < 2; ++pass) {
        int j0;
        if (pass == 0) j0 = jmax <= 2 ? 0 : (jmax >= ny - 1 ? ny - 3 : jmax - 2);
        else j0 = jmin >= 4 ? jmin - 4 : 0;
        double yk[5], fk[5], fs, ys;
        for (int m = 0; m < 5; ++m) {
            yk[m] = yp[j0 + m];
            fk[m] = Nu_ext[j0 + m];
        }
        orc_fit_parabola_ls5(yk, fk, pass == 0 ? 1 : -1, &fs, &ys);
        if (pass == 0) { d->nu_hot_max = fs; d->nu_hot_max_pos = ys; }
        else { d->nu_hot_min = fs; d->nu_hot_min_pos = ys; }
    }
    free(Nu_ext);
    free(Tleft);
}

/* Diagnostics: calNuRe (3DRB.F90:1712-1778, 2DRB.F90:1610-1675), RBcalc_Nu_global
 * (3DRB.F90:1839-1878, 2DRB.F90:2251-2297), RBcalc_Nu_wall_avg (3DRB.F90:2049-2201,
 * 2DRB.F90:2308-2473), plus mass, heat content and kinetic energy; the side-heated cell's
 * formulas when side_heated is set. */
void orc_diagnostics(const orc_state *s, orc_diag *d)
{
    memset(d, 0, sizeof(*d));
    if (s->c.side_heated) {
        diagnostics_side_heated(s, d);
        return;
    }
    const orc_params *p = &s->p;
    const double Th = s->c.thot, Tc = s->c.tcold, Tref = p->Tref, L = p->L;
    const double dy = 1.0 / L, deltaT = Th - Tc, coef = L / p->diffusivity;
    const double N = (double)s->n;
    const int nx = NX, ny = NY, nz = NZ;

    ksum snu = {0, 0}, sre = {0, 0}, sqy = {0, 0}, sm = {0, 0}, sh = {0, 0}, ske = {0, 0};
    for (int k = 0; k < nz; ++k) {
        ksum a = {0, 0}, b = {0, 0}, c = {0, 0}, m = {0, 0}, h = {0, 0}, e = {0, 0};
        for (int j = 0; j < ny; ++j)
            for (int i = 0; i < nx; ++i) {
                long n = IDX(s, i, j, k);
                double u = s->u[n], v = s->v[n], w = s->w[n], T = s->T[n];
                ks_add(&a, v * T);
                double q2 = u * u + v * v;
                if (s->c.dims == 3) q2 = q2 + w * w;
                ks_add(&b, q2);
                double dTdy;
                if (j == 0) dTdy = (3.0 * T + s->T[IDX(s, i, 1, k)] - 4.0 * Th) / (3.0 * dy);
                else if (j == ny - 1)
                    dTdy = (4.0 * Tc - 3.0 * T - s->T[IDX(s, i, ny - 2, k)]) / (3.0 * dy);
                else dTdy = (s->T[IDX(s, i, j + 1, k)] - s->T[IDX(s, i, j - 1, k)]) / (2.0 * dy);
                ks_add(&c, coef * v * (T - Tref) - dTdy);
                ks_add(&m, s->rho[n]);
                ks_add(&h, T);
                ks_add(&e, 0.5 * s->rho[n] * q2);
            }
        ks_merge(&snu, a);
        ks_merge(&sre, b);
        ks_merge(&sqy, c);
        ks_merge(&sm, m);
        ks_merge(&sh, h);
        ks_merge(&ske, e);
    }
    d->nu_vol = ks_val(snu) / N * L / p->diffusivity + 1.0;
    d->re_vol = sqrt(ks_val(sre) / N) * L / p->viscosity;
    d->nu_global = (ks_val(sqy) / N) / deltaT;
    d->mass = ks_val(sm);
    d->heat = ks_val(sh);
    d->kinetic = ks_val(ske);

    /* Hot-wall profile: per-x averages over z of T(i,1,k) and of the wall flux. */
    double *Nu_ext = (double *)calloc((size_t)nx + 2, sizeof(double));
    double *Tbot = (double *)calloc((size_t)nx + 2, sizeof(double));
    ksum shot = {0, 0}, scold = {0, 0}, smid = {0, 0};
    for (int i = 0; i < nx; ++i) {
        ksum tb = {0, 0}, nb = {0, 0};
        for (int k = 0; k < nz; ++k) {
            ks_add(&tb, s->T[IDX(s, i, 0, k)]);
            double qy = (8.0 * Th - 9.0 * s->T[IDX(s, i, 0, k)] + s->T[IDX(s, i, 1, k)]) / (3.0 * dy);
            ks_add(&nb, qy / deltaT);
            double qc = (-8.0 * Tc + 9.0 * s->T[IDX(s, i, ny - 1, k)] - s->T[IDX(s, i, ny - 2, k)]) /
                        (3.0 * dy);
            ks_add(&scold, qc / deltaT);
            double mid;
            if (ny % 2 == 1) {
                int jm = (ny + 1) / 2 - 1;
                long nmid = IDX(s, i, jm, k);
                mid = (coef * s->v[nmid] * (s->T[nmid] - Tref) -
                       (s->T[IDX(s, i, jm + 1, k)] - s->T[IDX(s, i, jm - 1, k)]) / (2.0 * dy)) /
                      deltaT;
            } else {
                int jb = ny / 2 - 1, jt = jb + 1;
                long nb_ = IDX(s, i, jb, k), nt_ = IDX(s, i, jt, k);
                mid = (coef * 0.5 * (s->v[nb_] * (s->T[nb_] - Tref) + s->v[nt_] * (s->T[nt_] - Tref)) +
                       (s->T[nb_] - s->T[nt_]) / dy) /
                      deltaT;
            }
            ks_add(&smid, mid);
        }
        Tbot[i + 1] = ks_val(tb) / (double)nz;
        Nu_ext[i + 1] = ks_val(nb) / (double)nz;
        ks_add(&shot, Nu_ext[i + 1]);
    }
    d->nu_hot = ks_val(shot) / (double)nx;
    d->nu_cold = ks_val(scold) / ((double)nx * nz);
    d->nu_middle = ks_val(smid) / ((double)nx * nz);

    const double *xp = s->xp;
    double xfit[4], Tfit[4], Twl, Twr;
    for (int m = 0; m < 4; ++m) {
        xfit[m] = xp[1 + m];
        Tfit[m] = Tbot[1 + m];
    }
    orc_fit_adiabatic_wall_T4(0.0, xfit, Tfit, &Twl);
    Nu_ext[0] = (2.0 * (Th - Twl) / dy) / deltaT;
    for (int m = 0; m < 4; ++m) {
        xfit[m] = xp[nx - 3 + m];
        Tfit[m] = Tbot[nx - 3 + m];
    }
    orc_fit_adiabatic_wall_T4(xp[nx + 1], xfit, Tfit, &Twr);
    Nu_ext[nx + 1] = (2.0 * (Th - Twr) / dy) / deltaT;

    int imax = 0, imin = 0;
    double vmax = Nu_ext[0], vmin = Nu_ext[0];
    for (int i = 1; i <= nx + 1; ++i) {
        if (Nu_ext[i] > vmax) { vmax = Nu_ext[i]; imax = i; }
        if (Nu_ext[i] < vmin) { vmin = Nu_ext[i]; imin = i; }
    }
    for (int pass = 0; pass < 2; ++pass) {
        int ic = pass == 0 ? imax : imin;
        int ii[5];
        if (ic <= 2) for (int m = 0; m < 5; ++m) ii[m] = m;
        else if (ic >= nx - 1) for (int m = 0; m < 5; ++m) ii[m] = nx - 3 + m;
        else for (int m = 0; m < 5; ++m) ii[m] = ic - 2 + m;
        double xk[5], fk[5], fs, xs;
        for (int m = 0; m < 5; ++m) {
            xk[m] = xp[ii[m]];
            fk[m] = Nu_ext[ii[m]];
        }
        orc_fit_parabola_ls5(xk, fk, pass == 0 ? 1 : -1, &fs, &xs);
        if (pass == 0) { d->nu_hot_max = fs; d->nu_hot_max_pos = xs; }
        else { d->nu_hot_min = fs; d->nu_hot_min_pos = xs; }
    }
    free(Nu_ext);
    free(Tbot);
}

/* ---------------------------------------------------------------------------------------
 * Mid-plane velocity maxima
 * 3D (both cells): calc_umid_max_common / calc_vmid_max_common / calc_wmid_max_common
 *   (3DRB.F90:2399-2518) with find_bracketing_index and interp_scalar_x/y/z (:2525-2600):
 *   the field is interpolated linearly onto the plane at half the domain's extent and the
 *   first grid maximum in loop order (outer index slowest) is reported with its coordinates.
 * 2D Rayleigh-Benard cell: RBcalc_umid_max (2DRB.F90:2484-2564; u on x = 0.5, a fixed
 *   nondimensional coordinate, the fitted position folded into the lower half) and
 *   RBcalc_vmid_max (:2575-2642; v on the middle row).
 * 2D side-heated cell: SideHeatedcalc_umid_max (2DRB.F90:2081-2153; u on the middle column)
 *   and SideHeatedcalc_vmid_max (:2166-2238).
 * The 2D maxima come from the 5-point least-squares parabola around the grid maximum.
 * ------------------------------------------------------------------------------------- */
static void bracket(const double *coord, int n, double target, int *iL, int *iR, double *weight)
{
    int l = 1;
    while (l < n && coord[l + 1] < target) ++l;
    int r = l + 1 < n ? l + 1 : n;
    double w;
    if (fabs(coord[r] - coord[l]) <= 1.0e-14) w = 0.0;
    else w = (target - coord[l]) / (coord[r] - coord[l]);
    w = w < 0.0 ? 0.0 : (w > 1.0 ? 1.0 : w);
    *iL = l;
    *iR = r;
    *weight = w;
}

/* value of field at (a = interpolated index pair along `axis`, other two indices) */
static double interp(const orc_state *s, const double *f, int axis, int L_, int R_, double w,
                     int p, int q)
{
    long nl, nr;
    if (axis == 0) { nl = IDX(s, L_ - 1, p - 1, q - 1); nr = IDX(s, R_ - 1, p - 1, q - 1); }
    else if (axis == 1) { nl = IDX(s, p - 1, L_ - 1, q - 1); nr = IDX(s, p - 1, R_ - 1, q - 1); }
    else { nl = IDX(s, p - 1, q - 1, L_ - 1); nr = IDX(s, p - 1, q - 1, R_ - 1); }
    if (L_ == R_) return f[nl];
    return (1.0 - w) * f[nl] + w * f[nr];
}

/* 5-point window around the grid maximum i0 of a line of n values (1-based), as the 2D
 * routines choose it, then the least-squares parabola through it. */
static void line_fit(const double *coord, const double *line, int n, double *fmax, double *pos)
{
    int i0 = 1;
    double best = line[1];
    for (int i = 2; i <= n; ++i)
        if (line[i] > best) { best = line[i]; i0 = i; }
    int a = i0 <= 2 ? 1 : (i0 >= n - 1 ? n - 4 : i0 - 2);
    double xk[5], fk[5];
    for (int m = 0; m < 5; ++m) { xk[m] = coord[a + m]; fk[m] = line[a + m]; }
    orc_fit_parabola_ls5(xk, fk, 1, fmax, pos);
}

void orc_midplane(const orc_state *s, orc_mid *m)
{
    const orc_params *p = &s->p;
    const double coef = p->L / p->diffusivity;
    const int nx = NX, ny = NY, nz = NZ;
    memset(m, 0, sizeof(*m));
    if (s->c.dims == 3) {
        const double *fields[3] = {s->u, s->v, s->w};
        const double *coords[3] = {s->xp, s->yp, s->zp};
        const int n[3] = {nx, ny, nz};
        double *out[3] = {&m->u_max, &m->v_max, &m->w_max};
        for (int a = 0; a < 3; ++a) {
            /* in-plane axes: (p, q) = (y, z) for u, (x, z) for v, (x, y) for w; q outermost */
            const int pa = a == 0 ? 1 : 0, qa = a == 2 ? 1 : 2;
            const double target = 0.5 * coords[a][n[a] + 1];
            int iL, iR;
            double w;
            bracket(coords[a], n[a], target, &iL, &iR, &w);
            double best = -HUGE_VAL;
            int pb = 1, qb = 1;
            for (int q = 1; q <= n[qa]; ++q)
                for (int pp = 1; pp <= n[pa]; ++pp) {
                    double val = interp(s, fields[a], a, iL, iR, w, pp, q);
                    if (val > best) { best = val; pb = pp; qb = q; }
                }
            out[a][0] = best * coef;
            out[a][1] = coords[pa][pb];
            out[a][2] = coords[qa][qb];
            out[a][3] = target;
        }
        return;
    }
    /* 2D */
    double *uline = (double *)calloc((size_t)ny + 1, sizeof(double));
    double *vline = (double *)calloc((size_t)nx + 1, sizeof(double));
    const double *xp = s->xp, *yp = s->yp;
    double xmid;
    if (s->c.side_heated) {
        if (nx % 2 == 1) {
            int im = (nx + 1) / 2;
            xmid = xp[im];
            for (int j = 1; j <= ny; ++j) uline[j] = s->u[IDX(s, im - 1, j - 1, 0)];
        } else {
            int il = nx / 2, ir = il + 1;
            xmid = 0.5 * (xp[il] + xp[ir]);
            for (int j = 1; j <= ny; ++j)
                uline[j] = 0.5 * (s->u[IDX(s, il - 1, j - 1, 0)] + s->u[IDX(s, ir - 1, j - 1, 0)]);
        }
    } else {
        const double target = 0.5;
        int il = 1;
        while (il < nx && xp[il + 1] < target) ++il;
        int ir = il + 1 < nx ? il + 1 : nx;
        xmid = target;
        if (fabs(xp[il] - target) <= 1.0e-14 || ir == il) {
            for (int j = 1; j <= ny; ++j) uline[j] = s->u[IDX(s, il - 1, j - 1, 0)];
        } else {
            double w = (target - xp[il]) / (xp[ir] - xp[il]);
            for (int j = 1; j <= ny; ++j)
                uline[j] = (1.0 - w) * s->u[IDX(s, il - 1, j - 1, 0)] + w * s->u[IDX(s, ir - 1, j - 1, 0)];
        }
    }
    double fmax, pos;
    line_fit(yp, uline, ny, &fmax, &pos);
    if (!s->c.side_heated) {
        double yLen = yp[ny + 1];
        if (pos > 0.5 * yLen) pos = yLen - pos;
    }
    m->u_max = fmax * coef;
    m->u_pos1 = pos;
    m->u_plane = xmid;

    double ymid;
    if (ny % 2 == 1) {
        int jm = (ny + 1) / 2;
        ymid = yp[jm];
        for (int i = 1; i <= nx; ++i) vline[i] = s->v[IDX(s, i - 1, jm - 1, 0)];
    } else {
        int jb = ny / 2, jt = jb + 1;
        ymid = 0.5 * (yp[jb] + yp[jt]);
        for (int i = 1; i <= nx; ++i)
            vline[i] = 0.5 * (s->v[IDX(s, i - 1, jb - 1, 0)] + s->v[IDX(s, i - 1, jt - 1, 0)]);
    }
    line_fit(xp, vline, nx, &fmax, &pos);
    m->v_max = fmax * coef;
    m->v_pos1 = pos;
    m->v_plane = ymid;
    free(uline);
    free(vline);
}

/* Mid-plane slices: write_midplane_x/y/z (3DRB.F90:2611-2713) and write_midplane_stream_x/y/z
 * (3DRB.F90:2721-3011).  On the plane across `axis` with in-plane axes (a1, a2) = (y, z),
 * (x, z), (x, y): psi integrates the a1 velocity component along a2 from the wall (the first
 * cell by (21 v1 - v2) / 72 dz, then Simpson's rule with quadratic midpoint values), and the
 * vorticity is the plane-normal component, d c2 / d a1 - d c1 / d a2 for the x and z planes and
 * its negative for the y plane (omega_y = du/dz - dw/dx), central differences inside and
 * second-order one-sided differences at the ends. */
static double slice_d(const double *q, int n, long st, int t, double h)
{
    if (n == 1) return 0.0;
    if (t == 0) return n == 2 ? (q[st] - q[0]) / h : (-3.0 * q[0] + 4.0 * q[st] - q[2 * st]) / (2.0 * h);
    if (t == n - 1)
        return n == 2 ? (q[t * st] - q[(t - 1) * st]) / h
                      : (3.0 * q[t * st] - 4.0 * q[(t - 1) * st] + q[(t - 2) * st]) / (2.0 * h);
    return (q[(t + 1) * st] - q[(t - 1) * st]) / (2.0 * h);
}

void orc_midplane_slice(const orc_state *s, int axis, double *uo, double *vo, double *wo,
                        double *To, double *psi, double *vort)
{
    const orc_params *p = &s->p;
    const double coef = p->L / p->diffusivity, d = 1.0 / p->L;
    const int n[3] = {NX, NY, NZ};
    const double *coords[3] = {s->xp, s->yp, s->zp};
    const double *fields[4] = {s->u, s->v, s->w, s->T};
    const int a1 = axis == 0 ? 1 : 0, a2 = axis == 2 ? 1 : 2;
    const int n1 = n[a1], n2 = n[a2];
    int iL, iR;
    double w;
    bracket(coords[axis], n[axis], 0.5 * coords[axis][n[axis] + 1], &iL, &iR, &w);
    double *sl[4];
    for (int c = 0; c < 4; ++c) {
        sl[c] = (double *)malloc(sizeof(double) * (size_t)n1 * n2);
        for (int q2 = 1; q2 <= n2; ++q2)
            for (int q1 = 1; q1 <= n1; ++q1)
                sl[c][(q1 - 1) + (size_t)n1 * (q2 - 1)] = interp(s, fields[c], axis, iL, iR, w, q1, q2);
    }
    double *outs[4] = {uo, vo, wo, To};
    for (int c = 0; c < 4; ++c)
        if (outs[c])
            for (long t = 0; t < (long)n1 * n2; ++t) outs[c][t] = c == 3 ? sl[3][t] : coef * sl[c][t];
    const double *c1 = sl[a1], *c2 = sl[a2];
    if (psi)
        for (int q1 = 0; q1 < n1; ++q1) {
            const double *q = c1 + q1;
            double *o = psi + q1;
            if (n2 == 1) { o[0] = 0.0; continue; }
            o[0] = d * (21.0 * (q[0] * coef) - q[n1] * coef) / 72.0;
            for (int k = 2; k <= n2; ++k) {
                const int m = k - 1;
                double mid;
                if (n2 == 2) mid = 0.5 * (q[0] + q[n1]) * coef;
                else if (m == 1) mid = (3.0 / 8.0 * q[0] + 3.0 / 4.0 * q[n1] - 1.0 / 8.0 * q[2 * n1]) * coef;
                else
                    mid = (-1.0 / 8.0 * q[(long)(m - 2) * n1] + 3.0 / 4.0 * q[(long)(m - 1) * n1] +
                           3.0 / 8.0 * q[(long)m * n1]) * coef;
                o[(long)(k - 1) * n1] = o[(long)(k - 2) * n1] +
                                        d / 6.0 * (q[(long)(k - 2) * n1] * coef + 4.0 * mid + q[(long)(k - 1) * n1] * coef);
            }
        }
    if (vort)
        for (int q2 = 0; q2 < n2; ++q2)
            for (int q1 = 0; q1 < n1; ++q1) {
                const double dc2 = slice_d(c2 + (long)q2 * n1, n1, 1, q1, d);
                const double dc1 = slice_d(c1 + q1, n2, n1, q2, d);
                vort[q1 + (long)n1 * q2] = axis == 1 ? coef * (dc1 - dc2) : coef * (dc2 - dc1);
            }
    for (int c = 0; c < 4; ++c) free(sl[c]);
}

/* ---------------------------------------------------------------------------------------
 * 2D stream function, vorticity and |psi| statistics: calc_psi_vort_and_output
 * (2DRB.F90:2657-2770), output_psi_center_abs (:3085-3160; 4-point Lagrange interpolation in
 * x then y around the centre), calc_psi_absmax_fine_spline (:2873-2989; natural cubic splines
 * through the wall zeros and the node values, along x on every row, then along y on every
 * column of a fine grid of 10000 intervals per unit length), spline_natural / splint
 * (:3000-3074).
 * ------------------------------------------------------------------------------------- */
static void spline_nat(int n, const double *x, const double *y, double *y2, double *u)
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

static double splint_eval(int n, const double *xa, const double *ya, const double *y2a, double x)
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

void orc_psi_vort_2d(const orc_state *s, double *psi_out, double *vort_out, double *stats)
{
    const orc_params *p = &s->p;
    const int nx = NX, ny = NY;
    const double dx = 1.0 / p->L, dy = 1.0 / p->L, coef = p->L / p->diffusivity;
    const double *xp = s->xp, *yp = s->yp;
#define U2(i, j) s->u[IDX(s, (i) - 1, (j) - 1, 0)]
#define V2(i, j) s->v[IDX(s, (i) - 1, (j) - 1, 0)]
    double *psi = (double *)malloc(sizeof(double) * (size_t)nx * ny);
#define PSI(i, j) psi[((i) - 1) + (size_t)nx * ((j) - 1)]
    for (int i = 1; i <= nx; ++i) {
        const double u1 = U2(i, 1) * coef, u2 = U2(i, 2) * coef;
        PSI(i, 1) = dy * (21.0 * u1 - u2) / 72.0;
        for (int j = 2; j <= ny; ++j) {
            const int m = j - 1;
            double um;
            if (m == 1) um = (3.0 / 8.0 * U2(i, 1) + 3.0 / 4.0 * U2(i, 2) - 1.0 / 8.0 * U2(i, 3)) * coef;
            else um = (-1.0 / 8.0 * U2(i, m - 1) + 3.0 / 4.0 * U2(i, m) + 3.0 / 8.0 * U2(i, m + 1)) * coef;
            const double inc = dy / 6.0 * ((U2(i, j - 1) * coef) + 4.0 * um + (U2(i, j) * coef));
            PSI(i, j) = PSI(i, j - 1) + inc;
        }
    }
    if (psi_out) memcpy(psi_out, psi, sizeof(double) * (size_t)nx * ny);
    if (vort_out)
        for (int j = 1; j <= ny; ++j)
            for (int i = 1; i <= nx; ++i) {
                double dvdx, dudy;
                if (i == 1) dvdx = (3.0 * V2(1, j) + V2(2, j) - 4.0 * 0.0) / (3.0 * dx);
                else if (i == nx) dvdx = (-3.0 * V2(nx, j) - V2(nx - 1, j) + 4.0 * 0.0) / (3.0 * dx);
                else dvdx = (V2(i + 1, j) - V2(i - 1, j)) / (2.0 * dx);
                if (j == 1) dudy = (3.0 * U2(i, 1) + U2(i, 2) - 4.0 * 0.0) / (3.0 * dy);
                else if (j == ny) dudy = (-3.0 * U2(i, ny) - U2(i, ny - 1) + 4.0 * 0.0) / (3.0 * dy);
                else dudy = (U2(i, j + 1) - U2(i, j - 1)) / (2.0 * dy);
                vort_out[(i - 1) + (size_t)nx * (j - 1)] = coef * (dvdx - dudy);
            }
    if (stats) {
        /* coarse centre value */
        const double x0 = 0.5 * xp[nx + 1], y0 = 0.5 * yp[ny + 1];
        double pc;
        if (nx < 4 || ny < 4) {
            pc = PSI((nx + 1) / 2, (ny + 1) / 2);
        } else {
            int i0 = 1;
            while (i0 < nx && xp[i0 + 1] <= x0) ++i0;
            i0 = i0 - 1 < nx - 3 ? i0 - 1 : nx - 3;
            if (i0 < 1) i0 = 1;
            int j0 = 1;
            while (j0 < ny && yp[j0 + 1] <= y0) ++j0;
            j0 = j0 - 1 < ny - 3 ? j0 - 1 : ny - 3;
            if (j0 < 1) j0 = 1;
            double x4[4], y4[4], f4[4], gx[4];
            for (int q = 0; q < 4; ++q) { x4[q] = xp[i0 + q]; y4[q] = yp[j0 + q]; }
            for (int q = 0; q < 4; ++q) {
                for (int pp = 0; pp < 4; ++pp) f4[pp] = PSI(i0 + pp, j0 + q);
                gx[q] = lagrange4(x0, x4, f4);
            }
            pc = lagrange4(y0, y4, gx);
        }
        stats[0] = fabs(pc);
        /* fine spline grid */
        const double xLen = xp[nx + 1], yLen = yp[ny + 1];
        const int nfx = (int)fmax(2.0, (double)lround(xLen * 10000.0) + 1.0);
        const int nfy = (int)fmax(2.0, (double)lround(yLen * 10000.0) + 1.0);
        const int nxe = nx + 2, nye = ny + 2;
        double *xe = (double *)malloc(sizeof(double) * nxe), *ye = (double *)malloc(sizeof(double) * nye);
        xe[0] = 0.0; xe[nxe - 1] = xp[nx + 1];
        for (int i = 1; i <= nx; ++i) xe[i] = xp[i];
        ye[0] = 0.0; ye[nye - 1] = yp[ny + 1];
        for (int j = 1; j <= ny; ++j) ye[j] = yp[j];
        const int nmax = nxe > nye ? nxe : nye;
        double *row = (double *)malloc(sizeof(double) * nmax), *y2 = (double *)malloc(sizeof(double) * nmax);
        double *wk = (double *)malloc(sizeof(double) * nmax);
        double *pcx = (double *)malloc(sizeof(double) * ny);
        double *pxf = (double *)malloc(sizeof(double) * (size_t)nfx * ny);
        for (int j = 1; j <= ny; ++j) {
            row[0] = 0.0; row[nxe - 1] = 0.0;
            for (int i = 1; i <= nx; ++i) row[i] = PSI(i, j);
            spline_nat(nxe, xe, row, y2, wk);
            pcx[j - 1] = splint_eval(nxe, xe, row, y2, 0.5 * xLen);
            for (int k = 0; k < nfx; ++k) {
                const double xq = xLen * (double)k / (double)(nfx - 1);
                pxf[(size_t)k * ny + (j - 1)] = splint_eval(nxe, xe, row, y2, xq);
            }
        }
        row[0] = 0.0; row[nye - 1] = 0.0;
        for (int j = 1; j <= ny; ++j) row[j] = pcx[j - 1];
        spline_nat(nye, ye, row, y2, wk);
        stats[1] = fabs(splint_eval(nye, ye, row, y2, 0.5 * yLen));
        double best = -1.0, xb = 0.0, yb = 0.0;
        for (int k = 0; k < nfx; ++k) {
            row[0] = 0.0; row[nye - 1] = 0.0;
            for (int j = 1; j <= ny; ++j) row[j] = pxf[(size_t)k * ny + (j - 1)];
            spline_nat(nye, ye, row, y2, wk);
            for (int l = 0; l < nfy; ++l) {
                const double yq = yLen * (double)l / (double)(nfy - 1);
                const double v = fabs(splint_eval(nye, ye, row, y2, yq));
                if (v > best) {
                    best = v;
                    xb = xLen * (double)k / (double)(nfx - 1);
                    yb = yq;
                }
            }
        }
        stats[2] = best;
        stats[3] = xb;
        stats[4] = yb;
        free(xe); free(ye); free(row); free(y2); free(wk); free(pcx); free(pxf);
    }
    free(psi);
#undef PSI
#undef U2
#undef V2
}

void orc_spline_eval(int n, const double *x, const double *y, int nq, const double *xq, double *out)
{
    double *y2 = (double *)malloc(sizeof(double) * n), *u = (double *)malloc(sizeof(double) * n);
    spline_nat(n, x, y, y2, u);
    for (int q = 0; q < nq; ++q) out[q] = splint_eval(n, x, y, y2, xq[q]);
    free(y2);
    free(u);
}

double orc_lagrange4(double xq, const double *xk, const double *fk) { return lagrange4(xq, xk, fk); }
