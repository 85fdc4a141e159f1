// plg_synth.hip -- seeded synthetic propagators on the GPU.
//
// Entry (component i, time t, site x) = N(0,1) + i N(0,1) from Philox4x32-10
// (Salmon et al., SC'11) with counter (x + V3*t, i, stream, tag) and key = seed,
// Box-Muller on two 53-bit uniforms, scaled by exp(-lambda |t - t0|) (no
// periodic wrap).  Perturbed form: (base + eps * pert) * scale with two
// streams.  The synthetic inputs stand in for QUDA point-to-all propagators
// (plegma/Calc_2pt.cpp:85-114), which this path does not compute.
#include "plg_internal.h"

#include <algorithm>

namespace plg {

__device__ __forceinline__ void philox10(uint32_t c[4], uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; r++) {
    const uint32_t lo0 = M0 * c[0], hi0 = __umulhi(M0, c[0]);
    const uint32_t lo1 = M1 * c[2], hi1 = __umulhi(M1, c[2]);
    const uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

__device__ __forceinline__ void normal_pair(uint64_t seed, uint32_t stream, uint32_t site, uint32_t comp,
                                            double &re, double &im) {
  uint32_t c[4] = {site, comp, stream, 0x504C4731u};
  philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const uint64_t a = ((((uint64_t)c[0]) << 32) | c[1]) >> 11;
  const uint64_t b = ((((uint64_t)c[2]) << 32) | c[3]) >> 11;
  const double u1 = ((double)a + 0.5) * (1.0 / 9007199254740992.0);
  const double u2 = ((double)b + 0.5) * (1.0 / 9007199254740992.0);
  const double rad = sqrt(-2.0 * log(u1));
  double sn, cs;
  sincos(2.0 * 3.141592653589793 * u2, &sn, &cs);
  re = rad * cs;
  im = rad * sn;
}

// plane exponent of the analytic envelope exp(-lambda |t - t0|)
__device__ __forceinline__ int envelope_exp(int t, int t0, double lambda) {
  return (int)rint(-lambda * fabs((double)(t - t0)) / 0.6931471805599453);
}

template <typename T>
__global__ void k_synth(cx<T> *__restrict__ data, int64_t cs, int t_first, int nt, int64_t V3, int T_lat,
                        uint64_t seed, uint32_t sa, uint32_t sb, double eps, int t0, double lambda,
                        int *__restrict__ plane_exp) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int pl = blockIdx.y;
  const int comp = blockIdx.z;
  if (x >= V3) return;
  const int t = (t_first + pl) % T_lat;
  double scale = exp(-lambda * fabs((double)(t - t0)));
  if (plane_exp) {
    const int e = envelope_exp(t, t0, lambda);
    if (x == 0 && comp == 0) plane_exp[pl] = e;
    // exp(-lambda|dt|) * 2^-e, evaluated without forming the tiny factor
    scale = exp(-lambda * fabs((double)(t - t0)) - (double)e * 0.6931471805599453);
  }
  const uint32_t site = (uint32_t)(x + V3 * t);
  double re, im;
  normal_pair(seed, sa, site, (uint32_t)comp, re, im);
  if (eps != 0.0) {
    double pr, pi;
    normal_pair(seed, sb, site, (uint32_t)comp, pr, pi);
    re += eps * pr;
    im += eps * pi;
  }
  cx<T> v;
  v.re = (T)(re * scale);
  v.im = (T)(im * scale);
  data[(int64_t)comp * cs + (int64_t)pl * V3 + x] = v;
}

int launch_synth(const plg_prop &dst, const int L[4], uint64_t seed, uint32_t sa, uint32_t sb, double eps,
                 int t0, double lambda, hipStream_t st) {
  const int64_t V3 = (int64_t)L[0] * L[1] * L[2];
  dim3 block(256), grid((unsigned)((V3 + 255) / 256), (unsigned)dst.nt, 144);
  int *pe = const_cast<int *>(dst.plane_exp);
  if (dst.precision == PLG_F64)
    k_synth<double><<<grid, block, 0, st>>>((cx<double> *)dst.data, dst.comp_stride, dst.t_first, dst.nt, V3,
                                            L[3], seed, sa, sb, eps, t0, lambda, pe);
  else
    k_synth<float><<<grid, block, 0, st>>>((cx<float> *)dst.data, dst.comp_stride, dst.t_first, dst.nt, V3,
                                           L[3], seed, sa, sb, eps, t0, lambda, pe);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

// per-plane exponent floor(log2 max|entry|) of an FP64 view's planes (the
// planes of dst), by a max-reduction over components and sites
__global__ void k_fill_int(int *p, int n, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
// largest binary exponent of the entries of each plane: a grid of (blocks,
// planes); every block strides over all 144 components of its plane and
// reduces in registers, across the wavefront and through LDS, so each block
// issues one atomic
__global__ __launch_bounds__(256) void k_plane_exp(int *__restrict__ pexp, int dtf,
                                                   const cx<double> *__restrict__ src, int64_t scs, int stf,
                                                   int64_t V3, int T_lat) {
  __shared__ int red[4];
  const int pl = blockIdx.y;
  const int t = (dtf + pl) % T_lat;
  int spl = t - stf;
  spl += spl < 0 ? T_lat : 0;
  int e = INT_MIN;
  for (int comp = 0; comp < 144; comp++) {
    const cx<double> *b = src + (int64_t)comp * scs + (int64_t)spl * V3;
    for (int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; x < V3; x += (int64_t)gridDim.x * blockDim.x) {
      const cx<double> v = b[x];
      const double a = fmax(fabs(v.re), fabs(v.im));
      if (a > 0) e = max(e, ilogb(a));
    }
  }
  for (int off = 32; off > 0; off >>= 1) e = max(e, __shfl_xor(e, off));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = e;
  __syncthreads();
  if (threadIdx.x == 0) {
    e = max(max(red[0], red[1]), max(red[2], red[3]));
    if (e != INT_MIN) atomicMax(&pexp[pl], e);
  }
}

// dst (F32) <- src (F64), round to nearest, over dst's planes
__global__ void k_cast(cx<float> *__restrict__ dst, int64_t dcs, int dtf, const cx<double> *__restrict__ src,
                       int64_t scs, int stf, int64_t V3, int T_lat, int *__restrict__ pexp) {
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int pl = blockIdx.y, comp = blockIdx.z;
  if (x >= V3) return;
  const int t = (dtf + pl) % T_lat;
  int spl = t - stf;
  spl += spl < 0 ? T_lat : 0;
  cx<double> v = src[(int64_t)comp * scs + (int64_t)spl * V3 + x];
  if (pexp) {
    int e = pexp[pl];
    if (e == INT_MIN) e = 0;          // all-zero plane
    v.re = ldexp(v.re, -e);           // exact power-of-two rescaling
    v.im = ldexp(v.im, -e);
  }
  dst[(int64_t)comp * dcs + (int64_t)pl * V3 + x] = cmake<float>((float)v.re, (float)v.im);
}

__global__ void k_fix_zero_exp(int *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n && p[i] == INT_MIN) p[i] = 0;
}

int launch_cast(const plg_prop &dst, const plg_prop &src, int64_t V3, int T_lat, hipStream_t st) {
  dim3 block(256), grid((unsigned)((V3 + 255) / 256), (unsigned)dst.nt, 144);
  int *pe = const_cast<int *>(dst.plane_exp);
  if (pe) {
    k_fill_int<<<(dst.nt + 255) / 256, 256, 0, st>>>(pe, dst.nt, INT_MIN);
    dim3 pgrid((unsigned)std::min<int64_t>(64, (V3 + 255) / 256), (unsigned)dst.nt);
    k_plane_exp<<<pgrid, block, 0, st>>>(pe, dst.t_first, (const cx<double> *)src.data, src.comp_stride,
                                         src.t_first, V3, T_lat);
  }
  k_cast<<<grid, block, 0, st>>>((cx<float> *)dst.data, dst.comp_stride, dst.t_first, (const cx<double> *)src.data,
                                 src.comp_stride, src.t_first, V3, T_lat, pe);
  if (pe) k_fix_zero_exp<<<(dst.nt + 255) / 256, 256, 0, st>>>(pe, dst.nt);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

}  // namespace plg
