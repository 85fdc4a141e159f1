// plg_prep.hip -- propagator preparation in place: the twisted-mass rotation
// to the physical basis and the antiperiodic boundary sign (counterparts of
// rotateToPhysicalBase_kernel and apply_boundaries_kernel,
// lib/kernels/PLEGMA_propagator_utils.cuh:64-124).  Both are HBM-bound
// element-wise passes; the contractions can apply the same transforms to the
// tiles they load instead (plg_prop.rotate_sign / bc_t0).
#include "plg_internal.h"

namespace plg {

// one thread per (plane, colour pair, spin quadrant (m, n) in {0,1}^2, site):
// the four entries S[m][n], S[m][n+2], S[m+2][n], S[m+2][n+2]
template <typename T>
__global__ void k_prop_rotate(cx<T> *__restrict__ d, int64_t cs, int nt, int64_t V3, int sign) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nt * 36 * V3) return;
  const int64_t x = e % V3;
  const int64_t r = e / V3;
  const int plane = (int)(r / 36), cq = (int)(r % 36);
  const int cp = cq % 9, q = cq / 9, m = q >> 1, n = q & 1;
  cx<T> *b = d + (int64_t)plane * V3 + x;
  auto at = [&](int mu, int nu) -> cx<T> & { return b[(int64_t)((mu * 4 + nu) * 9 + cp) * cs]; };
  cx<T> a = at(m, n), bb = at(m, n + 2), c = at(m + 2, n), dd = at(m + 2, n + 2);
  rot_physical(a, bb, c, dd, sign);
  at(m, n) = a;
  at(m, n + 2) = bb;
  at(m + 2, n) = c;
  at(m + 2, n + 2) = dd;
}

// negate the 144 components of the listed planes
template <typename T>
__global__ void k_prop_negate(cx<T> *__restrict__ d, int64_t cs, int64_t V3, const int *__restrict__ planes,
                              int nplanes) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nplanes * 144 * V3) return;
  const int64_t x = e % V3;
  const int64_t r = e / V3;
  const int k = (int)(r / 144), comp = (int)(r % 144);
  cx<T> &v = d[(int64_t)comp * cs + (int64_t)planes[k] * V3 + x];
  v.re = -v.re;
  v.im = -v.im;
}

int launch_prop_rotate(const plg_prop &p, int64_t V3, int sign, hipStream_t st) {
  const int64_t n = (int64_t)p.nt * 36 * V3;
  dim3 block(256), grid((unsigned)((n + 255) / 256));
  if (p.precision == PLG_F64)
    k_prop_rotate<double><<<grid, block, 0, st>>>((cx<double> *)p.data, p.comp_stride, p.nt, V3, sign);
  else
    k_prop_rotate<float><<<grid, block, 0, st>>>((cx<float> *)p.data, p.comp_stride, p.nt, V3, sign);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

int launch_prop_negate(const plg_prop &p, int64_t V3, const int *d_planes, int nplanes, hipStream_t st) {
  if (nplanes <= 0) return PLG_OK;
  const int64_t n = (int64_t)nplanes * 144 * V3;
  dim3 block(256), grid((unsigned)((n + 255) / 256));
  if (p.precision == PLG_F64)
    k_prop_negate<double><<<grid, block, 0, st>>>((cx<double> *)p.data, p.comp_stride, V3, d_planes, nplanes);
  else
    k_prop_negate<float><<<grid, block, 0, st>>>((cx<float> *)p.data, p.comp_stride, V3, d_planes, nplanes);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

}  // namespace plg
