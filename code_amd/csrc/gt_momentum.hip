// gt_momentum.hip -- the averaging of uv_advect_3d (src/3d/uv_advect_3d.F90),
// a caller of the transport path: the velocity of the U (V) control volumes
// and the transports, vertical velocity and layer heights averaged onto those
// volumes, which do_advection_3d then advects with tag U_TAG (V_TAG).
// One thread per cell, i fastest; HBM-bound.
#include "gt_kernels.h"

namespace gt {

namespace {

// U part (uv_advect_3d.F90:70-92) over every cell of the work fields: cells
// outside the reference's loops (level 0, the last column) are set to zero,
// so the V part starts from the same state on every call
__global__ void k_uv_prep_u(Dims d, const double* __restrict__ uu, const double* __restrict__ vv,
                            const double* __restrict__ ww, const double* __restrict__ ho,
                            const double* __restrict__ hn, const double* __restrict__ hun,
                            const double* __restrict__ hvn, double* __restrict__ fadv,
                            double* __restrict__ uuadv, double* __restrict__ vvadv,
                            double* __restrict__ wwadv, double* __restrict__ huadv,
                            double* __restrict__ hvadv) {
  const long c = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long N3 = d.NXY * (long)(d.kmax + 1);
  if (c >= N3) return;
  const int ii = (int)(c % d.NX);
  if (c >= d.NXY && ii <= d.NX - 2) {
    fadv[c] = uu[c] / hun[c];
    uuadv[c] = 0.5 * (uu[c] + uu[c + 1]);
    vvadv[c] = 0.5 * (vv[c] + vv[c + 1]);
    wwadv[c] = 0.5 * (ww[c] + ww[c + 1]);
    huadv[c] = 0.5 * (ho[c + 1] + hn[c + 1]);
    hvadv[c] = 0.5 * (hvn[c] + hvn[c + 1]);
  } else {
    fadv[c] = 0.0;
    uuadv[c] = 0.0;
    vvadv[c] = 0.0;
    wwadv[c] = 0.0;  // wwadv(:,:,0) = 0, :65
    huadv[c] = 0.0;
    hvadv[c] = 0.0;
  }
}

// V part (uv_advect_3d.F90:238-256): k = 1..kmax, j up to jmax+HALO-1
__global__ void k_uv_prep_v(Dims d, const double* __restrict__ uu, const double* __restrict__ vv,
                            const double* __restrict__ ww, const double* __restrict__ ho,
                            const double* __restrict__ hn, const double* __restrict__ hun,
                            const double* __restrict__ hvn, double* __restrict__ fadv,
                            double* __restrict__ uuadv, double* __restrict__ vvadv,
                            double* __restrict__ wwadv, double* __restrict__ huadv,
                            double* __restrict__ hvadv) {
  const long q = blockIdx.x * (long)blockDim.x + threadIdx.x;
  const long nrow = (long)d.NX * (d.NY - 1);  // cells of rows 0..NY-2 of one level
  if (q >= nrow * d.kmax) return;
  const long k = 1 + q / nrow;
  const long c = q % nrow + d.NXY * k;
  const long n = c + d.NX;
  fadv[c] = vv[c] / hvn[c];
  uuadv[c] = 0.5 * (uu[c] + uu[n]);
  vvadv[c] = 0.5 * (vv[c] + vv[n]);
  wwadv[c] = 0.5 * (ww[c] + ww[n]);
  huadv[c] = 0.5 * (hun[c] + hun[n]);
  hvadv[c] = 0.5 * (ho[n] + hn[n]);
}

}  // namespace

hipError_t launch_uv_prep(hipStream_t st, Dims d, bool upart, const double* uu, const double* vv,
                          const double* ww, const double* ho, const double* hn,
                          const double* hun, const double* hvn, double* work) {
  const long N3 = d.NXY * (long)(d.kmax + 1);
  double* w[6];
  for (int q = 0; q < 6; q++) w[q] = work + q * N3;
  if (upart) {
    hipLaunchKernelGGL(k_uv_prep_u, dim3((unsigned)((N3 + 255) / 256)), dim3(256), 0, st, d, uu,
                       vv, ww, ho, hn, hun, hvn, w[0], w[1], w[2], w[3], w[4], w[5]);
  } else {
    const long n = (long)d.NX * (d.NY - 1) * d.kmax;
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_uv_prep_v, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, d, uu,
                       vv, ww, ho, hn, hun, hvn, w[0], w[1], w[2], w[3], w[4], w[5]);
  }
  return hipGetLastError();
}

}  // namespace gt
