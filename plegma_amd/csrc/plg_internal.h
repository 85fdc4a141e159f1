// plg_internal.h -- declarations shared by the translation units of
// libplegma_amd.so (not installed).
#pragma once
#include "plegma_amd.h"
#include "plg_common.h"

#include <string>
#include <vector>

namespace plg {

// channel sets of the meson kernel
enum {
  SET_M16 = 0,       // 16 GAMMAS, one propagator
  SET_M10 = 1,       // contractMesons order, one propagator
  SET_M10_PAIR = 2,  // contractMesonsNew, S1 x conj(S2)
  SET_MALL = 3       // contractMesonsAll, 34 channels
};

template <typename T> struct PropDev {
  const cx<T> *data;
  int64_t cs;  // component stride (complex)
  int t_first;
  int nt;
};

// one contraction row: a time slice of a job and its propagators
struct RowDev {
  int t;
  int pa, pb;
  int job;
};

// a finalize row: computed row, source, output base (complex elements).
// Output channel k of the row reads computed channel ch_base + csel[k] and
// lands at base + ch_off[k] + m * m_stride.
struct OutRowDev {
  int crow;      // row of the site correlator / partial sums
  int ch_base;
  int sx, sy, sz;
  int nch_row;   // channels of this row (0: the launch-wide count)
  int64_t base;  // H offset of (m = 0, k = 0)
  // result scale 2^(ka*ea[pla] + kb*eb[plb]) undoes the propagators' plane
  // exponents (ea/eb NULL: no scaling)
  const int *ea, *eb;
  int pla, plb, ka, kb;
};
__device__ __forceinline__ double row_scale(const OutRowDev &R) {
  int e = 0;
  if (R.ea) e += R.ka * R.ea[R.pla];
  if (R.eb) e += R.kb * R.eb[R.plb];
  return ldexp(1.0, e);
}

template <typename T>
int launch_meson(int set, const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out,
                 int64_t V3, int T_lat, int nch_total, int ch0, hipStream_t st);
template <typename T>
int launch_mesmes(const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out, int64_t V3, int T_lat,
                  int nch_total, int ch0, int tiles_per_block, hipStream_t st);
template <typename T>
int launch_baryon(bool two, bool mes, int gmask, const PropDev<T> *props, const RowDev *rows,
                  int nrows, cx<T> *out, int64_t V3, int T_lat, int nch_total, int ch0,
                  int tiles_per_block, hipStream_t st);

// projection: partial sums part[split][n][m] (complex double) of
// sum_x exp(-i 2pi n.x / L) c[n][x] over the split's site range
struct ProjGeom {
  int64_t V3;
  int L[3];
  int M;          // common phase denominator lcm(Lx,Ly,Lz)
  int nmom;
  int ncols;      // n = row*nch + ch
  int nsplit;
  const int *mom_scaled;   // device, nmom*3: n_k * (M / L_k) mod M
  const double *cos_tab;   // device, M entries cos(2 pi k / M)
  const double *sin_tab;
  // layout of the K x N operand: 1 = site correlator, tile-major (below);
  // 0 = a field, column-major (column n, site x at n*V3 + x)
  int ctile;
};
// The site correlator is stored tile-major: the 16 sites of a site tile are
// contiguous per column, and the columns of a tile are contiguous, so that a
// projection step (16 sites x many columns) reads one contiguous block and a
// contraction workgroup writes one.  Element (column n, site x):
__host__ __device__ __forceinline__ int64_t ctile_idx(int64_t n, int64_t x, int64_t ncols) {
  return ((x >> 4) * ncols + n) * 16 + (x & 15);
}
__host__ __device__ __forceinline__ int64_t cidx(const ProjGeom &g, int64_t n, int64_t x) {
  return g.ctile ? ctile_idx(n, x, g.ncols) : n * g.V3 + x;
}
// sites of a site-correlator column rounded up to whole tiles
__host__ __device__ __forceinline__ int64_t ctile_sites(int64_t V3) { return (V3 + 15) / 16 * 16; }
int launch_project_f64(const ProjGeom &g, const cx<double> *c, cx<double> *part, int mtile,
                       hipStream_t st);
// wtab: M x 6 packed words (cos_p | sin_p << 16, -sin_p | cos_p << 16) for
// the three bf16 pieces p; mbw: 16-row blocks per wave (1, 2, 3, 5, 9)
int64_t proj_workgroups(int code, int64_t ncols, int nmom);
int launch_ft_out(const ProjGeom &g, const cx<double> *part, int dof, int nt, int accumulate, void *out,
                  int out_prec, hipStream_t st);
int launch_project_bf16(const ProjGeom &g, const cx<float> *c, cx<double> *part, int split_terms,
                        const uint32_t *wtab, int mbw, hipStream_t st);
int pick_mbw(int nmom);
// projection tile codes: < 100 per-wave A (MBW row blocks per wave);
// 100 + 10*RB + NBW: A tile shared by the workgroup (RB row blocks, 4*NBW
// column blocks)
int pick_proj(int nmom);
std::vector<int> proj_candidates(int nmom);
// sum over splits, source phase, scatter into the H layout
// nch_c: channels per computed row; nout: output channels per out row
int launch_finalize(const ProjGeom &g, const cx<double> *part, const OutRowDev *orows, int norows,
                    int nch_c, int nout, const int *csel, const int64_t *ch_off, int64_t m_stride,
                    void *out, int out_prec, hipStream_t st);
// position space: c[crow][ch][x] -> out[base + ch_off[k] + x*x_stride] (H position layout)
int launch_scatter_pos(const void *c, int c_prec, int64_t V3, int64_t ncols, const OutRowDev *orows, int norows,
                       int nch_c, int nout, const int *csel, const int64_t *ch_off, int64_t x_stride,
                       void *out, int out_prec, hipStream_t st);

// inputs
int launch_synth(const plg_prop &dst, const int L[4], uint64_t seed, uint32_t sa, uint32_t sb,
                 double eps, int t0, double lambda, hipStream_t st);
int launch_cast(const plg_prop &dst, const plg_prop &src, int64_t V3, int T_lat, hipStream_t st);

// error reporting
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

}  // namespace plg
