// lbm_internal.h -- host-side declarations shared by the C-ABI layer and the kernel launchers.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lbm_lattice.h"

namespace rblbm {

struct InitParams {
    int use_T0;           // T already in mac[4N..] (host-supplied profile)
    int rb_perturb;       // 0 none, 3 = 3D cos*sin (Ra <= 1e4), 2 = 2D sin*sin (Ra < 1e4)
    double L, pi, perturb_amp;
    uint64_t seed;
    double noise_scale;   // amp * (thot - tcold); 0 disables the noise
};

hipError_t launch_prologue(const KParams &kp, int bx, hipStream_t s, const double *Fin,
                           double *Fout, const double *mac, double *fy_out, bool three_d);
hipError_t launch_step3(const KParams &kp, int bx, hipStream_t s, bool first, const double *Fcur,
                        double *Fnext, const double *Gin, double *Gout, double *Bp, double *aux,
                        const double *mac);
hipError_t launch_step2(const KParams &kp, int bx, hipStream_t s, bool first, const double *Fcur,
                        double *Fnext, const double *Gin, double *Gout, double *Bp,
                        const double *fy_cur, double *fy_next, const double *mac);
hipError_t launch_state(const KParams &kp, int bx, hipStream_t s, bool three_d, const double *Fprev,
                        const double *Gpost, const double *aux_or_fy, double *mac, double *fo,
                        double *go, int k0, int nk);
hipError_t launch_init(const KParams &kp, int bx, hipStream_t s, const InitParams &ip, double *F,
                       double *G, double *Bp, double *mac, int qf, int qt);
hipError_t launch_transpose(bool to_soa, hipStream_t s, const double *in, double *out, int q,
                            uint64_t N, uint64_t c0, uint64_t n);
hipError_t launch_reconstruct(const KParams &kp, hipStream_t s, const double *F, const double *G,
                              double *Bp, double *mac, int qf, int qt, int *bad);
hipError_t launch_reduce_volume(const KParams &kp, hipStream_t s, const double *mac,
                                double L_over_kappa, double dy, double *out, unsigned blocks);
hipError_t launch_reduce_walls(const KParams &kp, hipStream_t s, const double *mac,
                               double L_over_kappa, double dy, double *tbot, double *nubot,
                               double *out);
hipError_t launch_reduce_check(const KParams &kp, hipStream_t s, const double *mac, double *prev,
                               double *out, unsigned blocks);

} // namespace rblbm
