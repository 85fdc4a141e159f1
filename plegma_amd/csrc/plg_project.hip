// plg_project.hip -- momentum projection as a dense phase-matrix GEMM on MFMA
// (FP64 path), and the finalize / position-space scatter kernels.
//
// Replaces the fused naive DFT of lib/kernels/PLEGMA_kernel_utils.cuh:820-860
// (per site and momentum a sincos, an LDS tree reduction and a host sum of
// per-block partials, PLEGMA_mesons.cuh:113-120) by
//
//   part[s][n][m] = sum_{x in split s} exp(-i 2 pi n_m . x / L) c[n][x]
//
// with n = (row, channel) column of the site correlator, computed on the
// (n, -n) momentum pairs as a real GEMM (k_project_f64 below).  The phase
// rows are generated per K tile from the exact integer phase index
//   k = sum_i n_i x_i (M / L_i) mod M,  M = lcm(Lx, Ly, Lz),
// and the FP64 table cos/sin(2 pi k / M); the source offset is applied later
// in FP64 as a per-momentum constant phase (the finalize kernel).
//
//  k_project_f64   v_mfma_f64_16x16x4_f64, FP64 operands and accumulation.
//  (the split-bf16 GEMM of the mixed path is in plg_project_bf16.hip; the
//  separable sums that both paths use where the momenta suit them, in
//  plg_project_sep.hip)
#include "plg_internal.h"

namespace plg {

typedef double d4 __attribute__((ext_vector_type(4)));

constexpr int PJ_THREADS = 256;
constexpr int PJ_NT = 128;  // columns per workgroup (4 waves x 32)

struct SiteXYZ {
  int x, y, z;
};
__device__ __forceinline__ SiteXYZ site_xyz(int64_t s, const ProjGeom &g) {
  SiteXYZ r;
  r.x = (int)(s % g.L[0]);
  int64_t q = s / g.L[0];
  r.y = (int)(q % g.L[1]);
  r.z = (int)(q / g.L[1]);
  return r;
}
__device__ __forceinline__ int phase_index(const int *ms, int m, SiteXYZ p, int M) {
  int k = ms[3 * m] * p.x + ms[3 * m + 1] * p.y + ms[3 * m + 2] * p.z;
  k %= M;
  return k < 0 ? k + M : k;
}

// sites [x_begin, x_end) of split s
__device__ __forceinline__ void split_range(const ProjGeom &g, int s, int64_t &b, int64_t &e) {
  const int64_t per = (g.V3 + g.nsplit - 1) / g.nsplit;
  b = (int64_t)s * per;
  e = b + per < g.V3 ? b + per : g.V3;
}

// Paired form (the same as the mixed path's, plg_project_bf16.hip): a
// createMom list holds n and -n, and with
//   Cr + i Ci = sum_x cos(theta_n x) c(x),  Sr + i Si = sum_x sin(theta_n x) c(x)
// P(+-n) = (Cr +- Si) + i (Ci -+ Sr).  The GEMM has the real rows {cos, sin}
// of one representative momentum per pair (n, -n), K = sites and the real
// columns {Re c, Im c}: part[split][2n + {re, im}][2j + {cos, sin}], half the
// MFMA work of the interleaved complex GEMM over every momentum.  The zero
// momentum is an ordinary pair here (its sin row is zero).
constexpr int F64_KT = 32;              // sites per K tile
constexpr int F64_LD = F64_KT + 1;      // padded LDS row (doubles)

template <int MT>
__global__ __launch_bounds__(PJ_THREADS) void k_project_f64(ProjGeom g, const cx<double> *__restrict__ c,
                                                            double *__restrict__ part) {
  constexpr int MB = MT / 16;  // M blocks per wave
  __shared__ double sB[PJ_NT][F64_LD];
  __shared__ double sA[MT][F64_LD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t nc0 = (int64_t)blockIdx.x * (PJ_NT / 2);  // first complex column
  const int mr0 = blockIdx.y * MT;                        // first real row (2j + cos/sin)
  const int split = blockIdx.z;
  int64_t xb, xe;
  split_range(g, split, xb, xe);
  const int R = 2 * g.npairs;
  const int64_t nreal = 2 * (int64_t)g.ncols;

  d4 acc[MB][2];
#pragma unroll
  for (int i = 0; i < MB; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) acc[i][j] = d4{0, 0, 0, 0};

  for (int64_t xt = xb; xt < xe; xt += F64_KT) {
    // stage B: 64 complex columns x 32 sites as 128 real columns {re, im}
    {
      const int s = tid & 31;
      const int64_t x = xt + s;
#pragma unroll
      for (int i = 0; i < (PJ_NT / 2) / 8; i++) {
        const int col = (tid >> 5) + 8 * i;
        const int64_t n = nc0 + col;
        cx<double> v = cmake<double>(0, 0);
        if (x < xe && n < g.ncols) v = c[cidx(g, n, x)];
        sB[2 * col][s] = v.re;
        sB[2 * col + 1][s] = v.im;
      }
    }
    // stage A: MT/2 pairs x 32 sites, rows {cos, sin} of one phase index
#pragma unroll
    for (int i = 0; i < (MT / 2 * F64_KT) / PJ_THREADS; i++) {
      const int e = tid + PJ_THREADS * i;
      const int jl = e / F64_KT, s = e % F64_KT;
      const int j = (mr0 >> 1) + jl;
      const int64_t x = xt + s;
      double cs = 0, sn = 0;
      if (j < g.npairs && x < xe) {
        const int k = phase_index(g.pair_ms, j, site_xyz(x, g), g.M);
        cs = g.cos_tab[k];
        sn = g.sin_tab[k];
      }
      sA[2 * jl][s] = cs;
      sA[2 * jl + 1][s] = sn;
    }
    __syncthreads();
#pragma unroll
    for (int ks = 0; ks < F64_KT / 4; ks++) {
      const int kk = ks * 4 + (lane >> 4);
      double bf[2];
#pragma unroll
      for (int j = 0; j < 2; j++) bf[j] = sB[wave * 32 + j * 16 + (lane & 15)][kk];
#pragma unroll
      for (int i = 0; i < MB; i++) {
        const double af = sA[i * 16 + (lane & 15)][kk];
#pragma unroll
        for (int j = 0; j < 2; j++)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af, bf[j], acc[i][j], 0, 0, 0);
      }
    }
    __syncthreads();
  }
  // D layout of 16x16x4 f64 on gfx950: lane l, value v -> row l/16 + 4v, column l%16
#pragma unroll
  for (int i = 0; i < MB; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int mr = mr0 + i * 16 + (lane >> 4) + 4 * v;
        const int64_t nr = 2 * nc0 + wave * 32 + j * 16 + (lane & 15);
        if (mr < R && nr < nreal) part[((int64_t)split * nreal + nr) * R + mr] = acc[i][j][v];
      }
}

int launch_project_f64(const ProjGeom &g, const cx<double> *c, cx<double> *part, int mtile,
                       hipStream_t st) {
  if (g.npairs <= 0) return PLG_ERR_INVALID;
  const int R = 2 * g.npairs;
  dim3 block(PJ_THREADS);
  const unsigned gx = (unsigned)((g.ncols + PJ_NT / 2 - 1) / (PJ_NT / 2));
  if (mtile == 32) {
    dim3 grid(gx, (unsigned)((R + 31) / 32), (unsigned)g.nsplit);
    k_project_f64<32><<<grid, block, 0, st>>>(g, c, reinterpret_cast<double *>(part));
  } else {
    dim3 grid(gx, (unsigned)((R + 63) / 64), (unsigned)g.nsplit);
    k_project_f64<64><<<grid, block, 0, st>>>(g, c, reinterpret_cast<double *>(part));
  }
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

// ------------------------------------------------------------------ finalize
// out[base + ch_off[k] + m*m_stride] = scale * exp(+i 2pi n_m . x0 / L) * sum_s part[s][n][m],
// n = crow*nch_c + ch_base + csel[k].  With the GEMM's exp(-i 2pi n.x/L) this
// is exp(-i 2pi n.(x - x0)/L) (PLEGMA_kernel_utils.cuh:833-838, sign -1).
template <typename TO>
__global__ void k_finalize(ProjGeom g, const double *__restrict__ part, const OutRowDev *__restrict__ orows,
                           int nch_c, int nout, const int *__restrict__ csel, const int64_t *__restrict__ ch_off,
                           int64_t m_stride, cx<TO> *__restrict__ out) {
  const OutRowDev R = orows[blockIdx.y];
  const int nr = R.nch_row > 0 ? R.nch_row : nout;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)g.nmom * nr) return;
  const int m = (int)(e / nr), k = (int)(e % nr);
  const int64_t n = (int64_t)R.crow * nch_c + R.ch_base + csel[k];
  const cx<double> v = part_value(g, part, n, m);
  const double re = v.re, im = v.im;
  int ph = g.mom_scaled[3 * m] * R.sx + g.mom_scaled[3 * m + 1] * R.sy + g.mom_scaled[3 * m + 2] * R.sz;
  ph %= g.M;
  if (ph < 0) ph += g.M;
  const double sc = row_scale(R);
  const double cs = g.cos_tab[ph] * sc, sn = g.sin_tab[ph] * sc;
  cx<TO> o;
  o.re = (TO)(re * cs - im * sn);
  o.im = (TO)(re * sn + im * cs);
  out[R.base + ch_off[k] + (int64_t)m * m_stride] = o;
}

int launch_finalize(const ProjGeom &g, const double *part, const OutRowDev *orows, int norows,
                    int nch_c, int nout, const int *csel, const int64_t *ch_off, int64_t m_stride,
                    void *out, int out_prec, hipStream_t st) {
  const int64_t ne = (int64_t)g.nmom * nout;
  dim3 block(256), grid((unsigned)((ne + 255) / 256), (unsigned)norows);
  if (out_prec == PLG_F64)
    k_finalize<double><<<grid, block, 0, st>>>(g, part, orows, nch_c, nout, csel, ch_off, m_stride,
                                               (cx<double> *)out);
  else
    k_finalize<float><<<grid, block, 0, st>>>(g, part, orows, nch_c, nout, csel, ch_off, m_stride,
                                              (cx<float> *)out);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

// ------------------------------------------------------------------ standalone FT output
// column col = i * nt + t of the GEMM (field component i, time slice t) ->
// out[t][i][m] (the buffer layout of lib/kernels/PLEGMA_FT.cuh:95-97), added
// to the previous contents when accumulate is set
template <typename TO>
__global__ void k_ft_out(ProjGeom g, const double *__restrict__ part, int dof, int nt, int accumulate,
                         cx<TO> *__restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)g.ncols * g.nmom) return;
  const int col = (int)(e / g.nmom), m = (int)(e % g.nmom);
  const cx<double> v = part_value(g, part, col, m);
  const double re = v.re, im = v.im;
  const int i = col / nt, t = col % nt;
  cx<TO> &o = out[((int64_t)t * dof + i) * g.nmom + m];
  if (accumulate) {
    o.re = (TO)((double)o.re + re);
    o.im = (TO)((double)o.im + im);
  } else {
    o.re = (TO)re;
    o.im = (TO)im;
  }
}

int launch_ft_out(const ProjGeom &g, const double *part, int dof, int nt, int accumulate, void *out,
                  int out_prec, hipStream_t st) {
  const int64_t ne = (int64_t)g.ncols * g.nmom;
  dim3 block(256), grid((unsigned)((ne + 255) / 256));
  if (out_prec == PLG_F64)
    k_ft_out<double><<<grid, block, 0, st>>>(g, part, dof, nt, accumulate, (cx<double> *)out);
  else
    k_ft_out<float><<<grid, block, 0, st>>>(g, part, dof, nt, accumulate, (cx<float> *)out);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

// ------------------------------------------------------------------ position space
template <typename TI, typename TO>
__global__ void k_scatter_pos(const cx<TI> *__restrict__ c, int64_t V3, int64_t ncols,
                              const OutRowDev *__restrict__ orows, int nch_c, int nout, const int *__restrict__ csel,
                              const int64_t *__restrict__ ch_off, int64_t x_stride, cx<TO> *__restrict__ out) {
  const OutRowDev R = orows[blockIdx.y];
  const int nr = R.nch_row > 0 ? R.nch_row : nout;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= V3 * nr) return;
  const int64_t x = e / nr;
  const int k = (int)(e % nr);
  const cx<TI> v = c[ctile_idx((int64_t)R.crow * nch_c + R.ch_base + csel[k], x, ncols)];
  const double sc = row_scale(R);
  cx<TO> o;
  o.re = (TO)(v.re * sc);
  o.im = (TO)(v.im * sc);
  out[R.base + ch_off[k] + x * x_stride] = o;
}

template <typename TI, typename TO>
static void scatter_t(const void *c, int64_t V3, int64_t ncols, const OutRowDev *orows, int norows, int nch_c,
                      int nout, const int *csel, const int64_t *ch_off, int64_t x_stride, void *out,
                      hipStream_t st) {
  const int64_t ne = V3 * nout;
  dim3 block(256), grid((unsigned)((ne + 255) / 256), (unsigned)norows);
  k_scatter_pos<TI, TO><<<grid, block, 0, st>>>((const cx<TI> *)c, V3, ncols, orows, nch_c, nout, csel, ch_off,
                                                x_stride, (cx<TO> *)out);
}

int launch_scatter_pos(const void *c, int c_prec, int64_t V3, int64_t ncols, const OutRowDev *orows, int norows,
                       int nch_c, int nout, const int *csel, const int64_t *ch_off, int64_t x_stride,
                       void *out, int out_prec, hipStream_t st) {
  if (c_prec == PLG_F64) {
    if (out_prec == PLG_F64)
      scatter_t<double, double>(c, V3, ncols, orows, norows, nch_c, nout, csel, ch_off, x_stride, out, st);
    else scatter_t<double, float>(c, V3, ncols, orows, norows, nch_c, nout, csel, ch_off, x_stride, out, st);
  } else {
    if (out_prec == PLG_F64)
      scatter_t<float, double>(c, V3, ncols, orows, norows, nch_c, nout, csel, ch_off, x_stride, out, st);
    else scatter_t<float, float>(c, V3, ncols, orows, norows, nch_c, nout, csel, ch_off, x_stride, out, st);
  }
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

}  // namespace plg
