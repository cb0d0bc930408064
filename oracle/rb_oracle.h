/*
 * rb_oracle.h -- CPU FP64 restatement of the reference's Rayleigh-Benard MRT-LBM hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker: only tests/, the smoke check
 * in __graft_entry__.py and the cpu_baseline leg of bench.py may load it.  The product path
 * (the HIP library behind include/lbm_rb.h) never links, loads or calls it.
 *
 * It follows 2DRB/3DRB.F90 (D3Q19 + D3Q7) and 2DRB/2DRB.F90 (D2Q9 + D2Q5) routine by routine,
 * keeping the reference's haloed array-of-structures layout (population index fastest,
 * f_post with a one-cell halo whose ghosts stay zero), its sweep order, expression order and
 * boundary quirks.  Global sums use Neumaier-compensated FP64 accumulation so that the result
 * does not depend on the OpenMP thread count.
 */
#ifndef RB_ORACLE_H
#define RB_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Wall-condition codes, mirroring the reference's cpp switches (2DRB/3DRB.F90:13-37). */
#define ORC_VEL_NOSLIP 0
#define ORC_VEL_PERIODIC 1
#define ORC_T_CONST 0
#define ORC_T_ADIABATIC 1
#define ORC_T_PERIODIC 2

typedef struct orc_cfg {
    int dims;                 /* 2 or 3 */
    int nx, ny, nz;           /* interior nodes; nz is 1 in 2D */
    double rayleigh, prandtl, mach, thot, tcold;
    int side_heated;          /* 0: Rayleigh-Benard cell (L = ny); 1: side-heated (L = nx) */
    int vel_x, vel_z;         /* ORC_VEL_*; y walls are always no-slip */
    int temp_x, temp_y, temp_z; /* ORC_T_* */
    int use_g;                /* EnableUseG: M1G flux correction */
    int legacy;               /* EnableLegacyThermalScheme: fixed Qk, Qnu, weights from paraA */
} orc_cfg;

typedef struct orc_params {
    double L, Tref, tauf, viscosity, diffusivity, gbeta, time_unit, velocity_unit;
    double Se, Seps, Snu, Spi, Sq, Sm;      /* D3Q19 rates (2D uses Snu, Sq) */
    double taug, Qk, Qnu, SG, heat_flux_scale;
    double paraA;             /* legacy thermal scheme: 3D 42 sqrt3 kappa - 6, 2D 20 sqrt3 kappa - 4 */
} orc_params;

typedef struct orc_diag {
    double nu_vol, re_vol;                   /* calNuRe */
    double nu_global;                        /* RBcalc_Nu_global */
    double nu_hot, nu_cold, nu_middle;       /* RBcalc_Nu_wall_avg */
    double nu_hot_max, nu_hot_max_pos, nu_hot_min, nu_hot_min_pos;
    double mass, heat, kinetic;              /* sum rho, sum T, 0.5 sum rho |u|^2 */
} orc_diag;

/* Mid-plane velocity maxima, scaled by velocityScaleCompare = L / kappa as the reference
 * reports them.  For each component: the maximum, its position (two in-plane coordinates in
 * 3D; one in 2D, the second is 0) and the plane's coordinate. */
typedef struct orc_mid {
    double u_max, u_pos1, u_pos2, u_plane;
    double v_max, v_pos1, v_pos2, v_plane;
    double w_max, w_pos1, w_pos2, w_plane;   /* 3D only */
} orc_mid;

typedef struct orc_state orc_state;

int orc_params_compute(const orc_cfg *cfg, orc_params *p);
orc_state *orc_create(const orc_cfg *cfg);
void orc_destroy(orc_state *s);

/* Exact-initialisation branch of `initial`; T0 (Fortran order, interior) replaces the
 * reference temperature profile when non-NULL. */
void orc_init_exact(orc_state *s, const double *T0);
/* Exact initialisation plus seeded uniform noise amp*(Thot-Tcold)*[-1,1) on every node. */
void orc_init_seeded(orc_state *s, uint64_t seed, uint64_t instance, double amp);
double orc_noise(uint64_t seed, uint64_t instance, uint64_t idx);

void orc_step(orc_state *s, int nsteps);
long orc_itc(const orc_state *s);
int orc_set_threads(int n);   /* returns the thread count now in effect */

/* Copy-out of the reference state (Fortran order; f/g population index fastest). */
void orc_get(const orc_state *s, double *f, double *g, double *rho, double *u, double *v,
             double *w, double *T, double *bx, double *by, double *bz);
/* Restart path: load f/g then reconstruct_macro_from_fg; returns -1 on rho <= 0. */
int orc_set_fg(orc_state *s, const double *f, const double *g);

void orc_check(orc_state *s, double *error_u, double *error_t);
void orc_diagnostics(const orc_state *s, orc_diag *d);
void orc_midplane(const orc_state *s, orc_mid *m);
/* 3D mid-plane slice across axis (0 x, 1 y, 2 z): u, v, w scaled by L/kappa, T, stream function
 * and vorticity; n1*n2 values each, first in-plane axis fastest (any pointer may be NULL). */
/* 2D stream function and vorticity (nx*ny, x fastest) and the |psi| statistics
 * stats = {|psi| at the centre from 4x4 Lagrange interpolation, |psi| at the centre from natural
 * splines, max |psi| on the fine spline grid, its x, its y}; any pointer may be NULL. */
void orc_psi_vort_2d(const orc_state *s, double *psi, double *vort, double *stats);
/* the natural cubic spline through (x, y) evaluated at xq (nq points), and the 4-point
 * Lagrange interpolant at xq, as used by orc_psi_vort_2d (exposed for their tests) */
void orc_spline_eval(int n, const double *x, const double *y, int nq, const double *xq, double *out);
double orc_lagrange4(double xq, const double *xk, const double *fk);
void orc_midplane_slice(const orc_state *s, int axis, double *u, double *v, double *w, double *T,
                        double *psi, double *vort);

void orc_fit_adiabatic_wall_T4(double y0, const double *y, const double *tt, double *t_wall);
void orc_fit_parabola_ls5(const double *y, const double *f, int mode, double *fstar,
                          double *ystar);

#ifdef __cplusplus
}
#endif
#endif
