// plg_contract.hip -- per-site contractions of the 2pt path (gfx950).
//
// Two kernel families write the site correlator c[row][ch][x] (complex T,
// x contiguous) that the projection GEMM consumes as its K-contiguous operand.
//
//  k_meson<T,SET>   one thread per site, HBM-bound.  Replaces the meson
//                   bodies of lib/kernels/PLEGMA_mesons.cuh:25-52,
//                   PLEGMA_mesonsNew.cuh:25-48, PLEGMA_mesonsAll.cuh:27-83.
//                   Single-propagator sets use the colour-traced spin Gram
//                   matrix G[p][q] = sum_ab S[p][ab] conj(S[q][ab]) restricted
//                   to the xor classes the channels touch (16 real + 24
//                   complex entries); two-propagator sets accumulate the
//                   channels directly.  Loads are coalesced SoA reads: lane i
//                   reads site x0+i of one spin-colour component.
//
//  The baryon contractions (and the fused meson + baryon kernel of the mixed
//  path) are in plg_baryon.hip.
#include "plg_internal.h"

namespace plg {

// ------------------------------------------------------------------ tables
struct MTerm {
  int p, q, ph;  // coefficient i^ph on  S1[p] . conj(S2[q]) (colour traced)
};
struct ChanTable {
  int nch;
  MTerm t[34][16];
};

// Tr[G S1 g5 G S2^dag g5] = sum X[d][a] X[b][g] S1[a][b] conj(S2[d][g]), X = g5 G
constexpr void meson_terms(int op, MTerm *t) {
  SPerm X = sp_mul(G_5, gamma_op(op));
  int c = sp_class(X);
  int n = 0;
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) {
      int d = a ^ c, g = b ^ c;
      t[n++] = MTerm{a * 4 + b, d * 4 + g, (X.ph[d] + X.ph[b]) & 3};
    }
}
// Tr[Gm S1 Gn S2^dag] = sum Gm[d][a] Gn[b][z] S1[a][b] conj(S2[d][z])
constexpr void pair_terms(int opm, int opn, MTerm *t) {
  SPerm M = gamma_op(opm), N = gamma_op(opn);
  int cm = sp_class(M), cn = sp_class(N);
  int n = 0;
  for (int a = 0; a < 4; a++)
    for (int b = 0; b < 4; b++) {
      int d = a ^ cm, z = b ^ cn;
      t[n++] = MTerm{a * 4 + b, d * 4 + z, (M.ph[d] + N.ph[b]) & 3};
    }
}

constexpr ChanTable make_table(int set) {
  ChanTable T{};
  if (set == SET_M16) {
    T.nch = 16;
    for (int op = 0; op < 16; op++) meson_terms(op, T.t[op]);
  } else if (set == SET_M10 || set == SET_M10_PAIR) {
    T.nch = 10;
    for (int i = 0; i < 10; i++) meson_terms(MESON10_OPS[i], T.t[i]);
  } else if (set == SET_MALL) {
    T.nch = 34;
    meson_terms(5, T.t[0]);
    meson_terms(0, T.t[1]);
    for (int p = 0; p < 2; p++)
      for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++)
          pair_terms(1 + i + 5 * p, 1 + j + 5 * p, T.t[2 + i * 4 + j + 16 * p]);
  }
  return T;
}

template <int SET> struct TableOf {
  static constexpr ChanTable v = make_table(SET);
};

__host__ __device__ constexpr int dmask(int d) { return (d << 2) | d; }
// index of p among the 8 representatives p < p^dmask(d)
constexpr int herm_rank(int d, int p) {
  int r = 0;
  for (int q = 0; q < p; q++)
    if (q < (q ^ dmask(d))) r++;
  return r;
}
constexpr int herm_rep(int d, int k) {
  int r = 0;
  for (int p = 0; p < 16; p++)
    if (p < (p ^ dmask(d))) {
      if (r == k) return p;
      r++;
    }
  return -1;
}

template <typename T>
__device__ __forceinline__ const cx<T> *site_base(const PropDev<T> &P, int t, int T_lat, int64_t V3,
                                                  int64_t x) {
  int plane = t - P.t_first;
  plane += plane < 0 ? T_lat : 0;
  return P.data + (int64_t)plane * V3 + x;
}

template <typename T> __device__ __forceinline__ cx<T> ldg(const cx<T> *p) { return gload(p); }

// ------------------------------------------------------------------ mesons
// Single propagator, Hermitian Gram form.  SET_M16 or SET_M10.
template <typename T, int SET>
__global__ __launch_bounds__(256) void k_meson_herm(const PropDev<T> *__restrict__ props,
                                                    const RowDev *__restrict__ rows,
                                                    cx<T> *__restrict__ out, int64_t V3,
                                                    int T_lat, int nch_total, int ch0) {
  const int r = blockIdx.y;
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= V3) return;
  const RowDev row = rows[r];
  const PropDev<T> P = props[row.pa];
  const cx<T> *base = site_base(P, row.t, T_lat, V3, x);
  const int64_t cs = P.cs;

  T n0[16];
  cx<T> gx[3][8];
#pragma unroll
  for (int p = 0; p < 16; p++) n0[p] = 0;
#pragma unroll
  for (int d = 0; d < 3; d++)
#pragma unroll
    for (int k = 0; k < 8; k++) gx[d][k] = cmake<T>(0, 0);

#pragma unroll 1
  for (int c = 0; c < 9; c++) {
    cx<T> s[16];
#pragma unroll
    for (int p = 0; p < 16; p++) s[p] = ldg(base + (int64_t)(p * 9 + c) * cs);
    // rotation to the physical basis (the boundary sign cancels in S x S^dag)
    if (P.rot) rot_spins(s, P.rot);
#pragma unroll
    for (int p = 0; p < 16; p++) n0[p] = fma(s[p].re, s[p].re, fma(s[p].im, s[p].im, n0[p]));
    static_for<1, 4>([&](auto d) {
      static_for<0, 8>([&](auto k) {
        constexpr int p = herm_rep(decltype(d)::value, decltype(k)::value);
        constexpr int q = p ^ dmask(decltype(d)::value);
        cfma_conj(gx[decltype(d)::value - 1][decltype(k)::value], s[p], s[q]);
      });
    });
  }
  const int64_t ncols = (int64_t)gridDim.y * nch_total, col0 = (int64_t)r * nch_total + ch0;
  static_for<0, TableOf<SET>::v.nch>([&](auto ch) {
    cx<T> acc = cmake<T>(0, 0);
    static_for<0, 16>([&](auto i) {
      constexpr MTerm tm = TableOf<SET>::v.t[decltype(ch)::value][decltype(i)::value];
      constexpr int d = (tm.p >> 2) ^ (tm.q >> 2);
      cx<T> v;
      if constexpr (d == 0) v = cmake<T>(n0[tm.p], 0);
      else if constexpr (tm.p < tm.q) v = gx[d - 1][herm_rank(d, tm.p)];
      else v = cconj(gx[d - 1][herm_rank(d, tm.q)]);
      acc = acc + cmul_iph(v, tm.ph);
    });
    gstore(out + ctile_idx(col0 + decltype(ch)::value, x, ncols), acc);
  });
}

// Two propagators, channels accumulated directly: SET_M10_PAIR, SET_MALL.
// Channels [C0, C1) of the set.
template <typename T, int SET, int C0, int C1>
__global__ __launch_bounds__(256) void k_meson_pair(const PropDev<T> *__restrict__ props,
                                                    const RowDev *__restrict__ rows,
                                                    cx<T> *__restrict__ out, int64_t V3,
                                                    int T_lat, int nch_total, int ch0) {
  const int r = blockIdx.y;
  const int64_t x = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (x >= V3) return;
  const RowDev row = rows[r];
  const PropDev<T> P1 = props[row.pa], P2 = props[row.pb];
  const cx<T> *b1 = site_base(P1, row.t, T_lat, V3, x);
  const cx<T> *b2 = site_base(P2, row.t, T_lat, V3, x);
  cx<T> acc[C1 - C0];
#pragma unroll
  for (int i = 0; i < C1 - C0; i++) acc[i] = cmake<T>(0, 0);
#pragma unroll 1
  for (int c = 0; c < 9; c++) {
    cx<T> s1[16], s2[16];
#pragma unroll
    for (int p = 0; p < 16; p++) {
      s1[p] = ldg(b1 + (int64_t)(p * 9 + c) * P1.cs);
      s2[p] = ldg(b2 + (int64_t)(p * 9 + c) * P2.cs);
    }
    if (P1.rot) rot_spins(s1, P1.rot);
    if (P2.rot) rot_spins(s2, P2.rot);
    static_for<C0, C1>([&](auto ch) {
      static_for<0, 16>([&](auto i) {
        constexpr MTerm tm = TableOf<SET>::v.t[decltype(ch)::value][decltype(i)::value];
        cx<T> prod = cmake<T>(0, 0);
        cfma_conj(prod, s1[tm.p], s2[tm.q]);
        acc[decltype(ch)::value - C0] = acc[decltype(ch)::value - C0] + cmul_iph(prod, tm.ph);
      });
    });
  }
  const int64_t ncols = (int64_t)gridDim.y * nch_total, col0 = (int64_t)r * nch_total + ch0;
  // boundary signs: the channels are bilinear in S1 and S2
  const T sg = prop_sign(P1, row.t) * prop_sign(P2, row.t);
#pragma unroll
  for (int ch = C0; ch < C1; ch++)
    gstore(out + ctile_idx(col0 + ch, x, ncols), cmake<T>(sg * acc[ch - C0].re, sg * acc[ch - C0].im));
}

// ------------------------------------------------------------------ launchers
template <typename T>
int launch_meson(int set, const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out,
                 int64_t V3, int T_lat, int nch_total, int ch0, hipStream_t st) {
  dim3 block(256), grid((unsigned)((V3 + 255) / 256), (unsigned)nrows);
  switch (set) {
  case SET_M16:
    k_meson_herm<T, SET_M16><<<grid, block, 0, st>>>(props, rows, out, V3, T_lat, nch_total, ch0);
    break;
  case SET_M10:
    k_meson_herm<T, SET_M10><<<grid, block, 0, st>>>(props, rows, out, V3, T_lat, nch_total, ch0);
    break;
  case SET_M10_PAIR:
    k_meson_pair<T, SET_M10_PAIR, 0, 10><<<grid, block, 0, st>>>(props, rows, out, V3, T_lat, nch_total, ch0);
    break;
  case SET_MALL:
    k_meson_pair<T, SET_MALL, 0, 17><<<grid, block, 0, st>>>(props, rows, out, V3, T_lat, nch_total, ch0);
    k_meson_pair<T, SET_MALL, 17, 34><<<grid, block, 0, st>>>(props, rows, out, V3, T_lat, nch_total, ch0);
    break;
  default:
    return PLG_ERR_INVALID;
  }
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

template int launch_meson<double>(int, const PropDev<double> *, const RowDev *, int, cx<double> *,
                                  int64_t, int, int, int, hipStream_t);
template int launch_meson<float>(int, const PropDev<float> *, const RowDev *, int, cx<float> *,
                                 int64_t, int, int, int, hipStream_t);

}  // namespace plg
