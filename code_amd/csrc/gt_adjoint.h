// gt_adjoint.h -- launchers of the pressure and EOS reverse-mode kernels
// (gt_adjoint.hip).
#pragma once

#include "gt_device.h"

namespace gt {

hipError_t launch_bm_adj(hipStream_t st, Dims d, const int32_t* au, const int32_t* av,
                         const double* zz, const double* hun, const double* hvn,
                         const double* ixbar, const double* iybar, double* bbar, double dx,
                         double dy);
// transpose of ip_shchepetkin_mcwilliams with respect to buoy at the forward
// point (zz from the forward call); bdir and Pbar are work fields (N3 each)
hipError_t launch_sm_adj(hipStream_t st, Dims d, const int32_t* az, const int32_t* au,
                         const int32_t* av, const double* buoy, const double* zz, const double* hn,
                         const double* hun, const double* hvn, const double* sseo,
                         const double* ixbar, const double* iybar, double* bdir, double* Pbar,
                         double* bbar, double dx, double dy);
hipError_t launch_eos_linear_adj(hipStream_t st, Dims d, const double* bbar, double* Tbar,
                                 double* Sbar, double dtr0, double dsr0);

}  // namespace gt
