// gt_substeps.h -- launchers of the per-sub-step HALFSPLIT kernels and their
// transposes (gt_substeps.hip), TF = double (the reference arithmetic) or
// float (the batched FP32 arithmetic).  Tracer sets hold ntr fields of
// (kmax+1)*NY*NX values, tracer t at + t*ts.
#pragma once

#include "gt_device.h"

namespace gt {

// the layer heights after each of the five HALFSPLIT sub-steps, D[0..4] =
// D1..D5 from ho (FP64, the sub-steps' expressions and update ranges)
hipError_t launch_heights(hipStream_t st, Dims d, const uint8_t* flags, const double* ho,
                          const double* uu, const double* vv, const double* ww,
                          double* const D[5], double dt, double dx, double dy, double ard1,
                          int scheme, int kshift);
// one horizontal half sub-step (dir 0: u, 1: v), Dio -> Dn; fout must not
// overlap fin
template <typename TF>
hipError_t launch_sub_h(hipStream_t st, Dims d, int dir, int ntr, long ts, const TF* fin, TF* fout,
                        const double* Dio, const double* Dn, const double* vel, const double* Dv,
                        const uint8_t* flags, double dt, double dx, double dy, double ard1,
                        int scheme, double AH);
// the vertical sub-step (adv_ver_iterations = 1), D2 -> D3
template <typename TF>
hipError_t launch_sub_w(hipStream_t st, Dims d, int ntr, long ts, const TF* fin, TF* fout,
                        const double* D2, const double* D3, const double* ww,
                        const uint8_t* flags, double dt, int scheme, int kshift);
// transposes: f is the state before the sub-step, lam_in the adjoint of its
// output, lam_out (distinct from lam_in) receives the adjoint of its input
template <typename TF>
hipError_t launch_sub_h_adj(hipStream_t st, Dims d, int dir, int ntr, long ts, const TF* f,
                            const TF* lam_in, TF* lam_out, const double* Dio, const double* Dn,
                            const double* vel, const double* Dv, const uint8_t* flags, double dt,
                            double dx, double dy, double ard1, int scheme, double AH);
template <typename TF>
hipError_t launch_sub_w_adj(hipStream_t st, Dims d, int ntr, long ts, const TF* f,
                            const TF* lam_in, TF* lam_out, const double* D2, const double* D3,
                            const double* ww, const uint8_t* flags, double dt, int scheme,
                            int kshift);

}  // namespace gt
