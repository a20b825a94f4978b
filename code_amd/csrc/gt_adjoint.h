// gt_adjoint.h -- launchers of the reverse-mode kernels (gt_adjoint.hip).
#pragma once

#include "gt_device.h"

namespace gt {

// dir 0: u sub-step (faces along i), dir 1: v sub-step (faces along j).
// dlen_flux: the length in the CFL and diffusion terms (dx for u, dy for v);
// dlen_face: the face length in the divergence (dy for u, dx for v).
hipError_t launch_split_uv_adj(hipStream_t st, Dims d, int k0, int nslice, int dir,
                               const double* f, const double* lam_in, double* lam_out,
                               const double* Dio, const double* Dn, const double* vel,
                               const double* Dvel, const uint8_t* mflux, const uint8_t* mupd,
                               double dt, double splitfac, int scheme, double AH,
                               double dlen_flux, double dlen_face, double ard1);
hipError_t launch_split_w_adj(hipStream_t st, Dims d, const double* f, const double* lam_in,
                              double* lam_out, const double* hio, const double* hn,
                              const double* ww, const int32_t* az, double dt, double splitfac,
                              int scheme, int kshift);
hipError_t launch_bm_adj(hipStream_t st, Dims d, const int32_t* au, const int32_t* av,
                         const double* zz, const double* hun, const double* hvn,
                         const double* ixbar, const double* iybar, double* bbar, double dx,
                         double dy);
hipError_t launch_eos_linear_adj(hipStream_t st, Dims d, const double* bbar, double* Tbar,
                                 double* Sbar, double dtr0, double dsr0);

}  // namespace gt
