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
//  k_baryon<T,...>  16 lanes per site, 16 sites per 256-thread workgroup,
//                   the site blocks staged in LDS.  Replaces
//                   PLEGMA_baryons_NtoN.cu:6-46 and PLEGMA_baryons_deltas.cu:
//                   61-110 with the factorised forms
//                     proton = D(Cg5) - X(Cg5)
//                     Delta_k = 2 D(Cg_k) - 4 X(Cg_k)     (uuu, Cg_k symmetric)
//                   D = sum_cc' Q^cc' P^cc'_{mu mu'},
//                   Q^cc' = eps eps' sum_{nu nu'} P^aa'_{nu nu'} St^bb'_{nu nu'},
//                   X = eps eps' sum P^ca'_{mu nu'} St^bb'_{nu nu'} P^ac'_{nu mu'},
//                   St = A Qm Abar.  Per site and block 3600 complex MACs
//                   (reference body: 27648 per dataset for the proton).
//                   Optionally also the 16 meson channels (fused mixed path).
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

template <typename T> __device__ __forceinline__ cx<T> ldg(const cx<T> *p) { return *p; }

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
  cx<T> *o = out + ((int64_t)r * nch_total + ch0) * V3 + x;
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
    o[(int64_t)decltype(ch)::value * V3] = acc;
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
    static_for<C0, C1>([&](auto ch) {
      static_for<0, 16>([&](auto i) {
        constexpr MTerm tm = TableOf<SET>::v.t[decltype(ch)::value][decltype(i)::value];
        cx<T> prod = cmake<T>(0, 0);
        cfma_conj(prod, s1[tm.p], s2[tm.q]);
        acc[decltype(ch)::value - C0] = acc[decltype(ch)::value - C0] + cmul_iph(prod, tm.ph);
      });
    });
  }
  cx<T> *o = out + ((int64_t)r * nch_total + ch0) * V3 + x;
#pragma unroll
  for (int ch = C0; ch < C1; ch++) o[(int64_t)ch * V3] = acc[ch - C0];
}

// ------------------------------------------------------------------ baryons
constexpr int BT_SITES = 16;   // sites per workgroup tile
constexpr int BT_THREADS = 256;

// LDS of one workgroup.  w holds the exchange intermediate W of a block,
// the meson Gram rows before the blocks, and the output stage after them.
// s2 is last so the one-propagator variant allocates without it.
template <typename T> struct BaryonShared {
  cx<T> s1[144][BT_SITES];
  cx<T> w[BT_SITES][16][9];
  cx<T> s2[144][BT_SITES];
};

// per meson channel op (GAMMAS order): xor class and phases of X = g5 G
struct MesonX {
  int cls[16];
  int ph[16][4];
};
constexpr MesonX make_mesonx() {
  MesonX m{};
  for (int op = 0; op < 16; op++) {
    SPerm X = sp_mul(G_5, gamma_op(op));
    m.cls[op] = sp_class(X);
    for (int i = 0; i < 4; i++) m.ph[op][i] = X.ph[i];
  }
  return m;
}
__constant__ MesonX c_mesonx = make_mesonx();

template <typename T> __device__ __forceinline__ cx<T> shfl_xor16(cx<T> v, int m) {
  return cmake<T>(__shfl_xor(v.re, m, 16), __shfl_xor(v.im, m, 16));
}

// one factorised block for diquark `W`: returns D and X for lane (mu,mu')
template <typename T, int WQ>
__device__ __forceinline__ void baryon_block(const cx<T> (*P)[BT_SITES], const cx<T> (*Qm)[BT_SITES],
                                             cx<T> (*Wsh)[9], int s, int lane, cx<T> &D,
                                             cx<T> &X) {
  constexpr SPerm A = diquark(WQ);
  constexpr SPerm Ab = diquark_bar(WQ);
  constexpr int cAb = sp_class(Ab);
  const int hi = lane >> 2, lo = lane & 3;
  // ---- step 1: lane = (nu, nu'): partial Q, then D on lane = (mu, mu')
  cx<T> pa[9], st[9];
  {
    const int nu = hi, nup = lo;
    const int lam = A.col[nu];
    const int lamp = nup ^ cAb;  // Abar[lam'][nu'] != 0
    const int ph = A.ph[nu] + Ab.ph[lamp];
#pragma unroll
    for (int i = 0; i < 9; i++) {
      pa[i] = P[pidx(nu, nup, 0, 0) + i][s];
      st[i] = cmul_iph(Qm[pidx(lam, lamp, 0, 0) + i][s], ph);
    }
  }
  cx<T> Q[9];
#pragma unroll
  for (int i = 0; i < 9; i++) Q[i] = cmake<T>(0, 0);
#pragma unroll
  for (int e1 = 0; e1 < 6; e1++)
#pragma unroll
    for (int e2 = 0; e2 < 6; e2++) {
      const int a = EPS[e1][0], b = EPS[e1][1], c = EPS[e1][2];
      const int ap = EPS[e2][0], bp = EPS[e2][1], cp = EPS[e2][2];
      cx<T> prod = pa[a * 3 + ap] * st[b * 3 + bp];
      if (SGN_EPS[e1] * SGN_EPS[e2] > 0) Q[c * 3 + cp] = Q[c * 3 + cp] + prod;
      else Q[c * 3 + cp] = Q[c * 3 + cp] - prod;
    }
#pragma unroll
  for (int m = 1; m < 16; m <<= 1)
#pragma unroll
    for (int i = 0; i < 9; i++) Q[i] = Q[i] + shfl_xor16(Q[i], m);
  D = cmake<T>(0, 0);
#pragma unroll
  for (int i = 0; i < 9; i++) cfma(D, Q[i], pa[i]);  // pa holds P[mu][mu'] on this lane

  // ---- step 2: lane = (nu', mu'): W^{c a'}_{nu' mu'}
  {
    const int nup = hi, mup = lo;
    const int lamp = nup ^ cAb;
    cx<T> Wl[9];
#pragma unroll
    for (int i = 0; i < 9; i++) Wl[i] = cmake<T>(0, 0);
#pragma unroll 1
    for (int nu = 0; nu < 4; nu++) {
      const int lam = A.col[nu];
      const int ph = A.ph[nu] + Ab.ph[lamp];
      cx<T> sq[9], pp[9];
#pragma unroll
      for (int i = 0; i < 9; i++) {
        sq[i] = cmul_iph(Qm[pidx(lam, lamp, 0, 0) + i][s], ph);
        pp[i] = P[pidx(nu, mup, 0, 0) + i][s];
      }
#pragma unroll
      for (int e1 = 0; e1 < 6; e1++)
#pragma unroll
        for (int e2 = 0; e2 < 6; e2++) {
          // eps_{abc} eps_{a'b'c'} St^{bb'} P^{ac'} -> W[c][a']
          const int a = EPS[e1][0], b = EPS[e1][1], c = EPS[e1][2];
          const int ap = EPS[e2][0], bp = EPS[e2][1], cp = EPS[e2][2];
          cx<T> prod = sq[b * 3 + bp] * pp[a * 3 + cp];
          if (SGN_EPS[e1] * SGN_EPS[e2] > 0) Wl[c * 3 + ap] = Wl[c * 3 + ap] + prod;
          else Wl[c * 3 + ap] = Wl[c * 3 + ap] - prod;
        }
    }
#pragma unroll
    for (int i = 0; i < 9; i++) Wsh[lane][i] = Wl[i];
  }
  __syncthreads();
  // ---- step 3: lane = (mu, mu'): X = sum_{c a' nu'} P^{c a'}_{mu nu'} W^{c a'}_{nu' mu'}
  {
    const int mu = hi, mup = lo;
    X = cmake<T>(0, 0);
#pragma unroll 1
    for (int nup = 0; nup < 4; nup++)
#pragma unroll
      for (int i = 0; i < 9; i++) cfma(X, P[pidx(mu, nup, 0, 0) + i][s], Wsh[nup * 4 + mup][i]);
  }
  __syncthreads();
}

// Output channels per row, in order:
//   [mesons16 of S1 if MES] then for each dataset d (nd = 2 if TWO else 1),
//   for each selected group g: 16 spin components (mu*4 + mu').
// Groups: bit 0 proton (nucl_nucl), bits 7..9 Delta++ k=1..3.
template <typename T, bool TWO, bool MES, int GMASK>
__global__ __launch_bounds__(BT_THREADS) void k_baryon(const PropDev<T> *__restrict__ props,
                                                       const RowDev *__restrict__ rows,
                                                       cx<T> *__restrict__ out, int64_t V3,
                                                       int T_lat, int nch_total, int ch0, int ntiles) {
  constexpr int NG = ((GMASK >> 0) & 1) + ((GMASK >> 7) & 1) + ((GMASK >> 8) & 1) + ((GMASK >> 9) & 1);
  constexpr int ND = TWO ? 2 : 1;
  constexpr int NOUT = (MES ? 1 : 0) + ND * NG;  // outputs per lane
  static_assert(NOUT <= 9, "output stage holds 9 values per lane");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  BaryonShared<T> &sh = *reinterpret_cast<BaryonShared<T> *>(smem_raw);

  const int r = blockIdx.y;
  const RowDev row = rows[r];
  const PropDev<T> P1 = props[row.pa];
  const PropDev<T> P2 = props[TWO ? row.pb : row.pa];
  const int tid = threadIdx.x;
  const int s = tid >> 4, lane = tid & 15;

  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t x0 = (int64_t)tile * BT_SITES;
    // ---- stage the 16 site blocks (component-major, coalesced along x)
    {
      const int ls = tid & 15;
      const int64_t x = x0 + ls;
      const bool ok = x < V3;
      const cx<T> *b1 = site_base(P1, row.t, T_lat, V3, ok ? x : 0);
      const cx<T> *b2 = site_base(P2, row.t, T_lat, V3, ok ? x : 0);
      for (int i = tid >> 4; i < 144; i += BT_THREADS / BT_SITES) {
        sh.s1[i][ls] = ok ? ldg(b1 + (int64_t)i * P1.cs) : cmake<T>(0, 0);
        if (TWO) sh.s2[i][ls] = ok ? ldg(b2 + (int64_t)i * P2.cs) : cmake<T>(0, 0);
      }
    }
    __syncthreads();
    const cx<T>(*S1)[BT_SITES] = sh.s1;
    const cx<T>(*S2)[BT_SITES] = TWO ? sh.s2 : sh.s1;

    cx<T> res[9];
    int nr = 0;
    // ---- mesons: lane p = (alpha,beta) forms G[p][p^d], lane ch sums 16 terms
    if (MES) {
      cx<T>(*g)[16] = reinterpret_cast<cx<T>(*)[16]>(&sh.w[s][0][0]);  // [4][16]
#pragma unroll
      for (int d = 0; d < 4; d++) {
        cx<T> gv = cmake<T>(0, 0);
        const int q = lane ^ dmask(d);
#pragma unroll
        for (int c = 0; c < 9; c++) cfma_conj(gv, S1[lane * 9 + c][s], S1[q * 9 + c][s]);
        g[d][lane] = gv;
      }
      __syncthreads();
      // lane = channel op (GAMMAS order): C = sum i^(ph[a^c] + ph[b]) G[(a,b)][(a^c,b^c)]
      const int cX = c_mesonx.cls[lane];
      cx<T> acc = cmake<T>(0, 0);
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
          acc = acc + cmul_iph(g[cX][a * 4 + b], c_mesonx.ph[lane][a ^ cX] + c_mesonx.ph[lane][b]);
      res[nr++] = acc;
      __syncthreads();
    }
    // ---- baryon blocks
#pragma unroll
    for (int d = 0; d < ND; d++) {
      const cx<T>(*P)[BT_SITES] = d == 0 ? S1 : S2;
      const cx<T>(*Qo)[BT_SITES] = d == 0 ? S2 : S1;
      cx<T> D, X;
      if (GMASK & 1) {
        baryon_block<T, 0>(P, Qo, sh.w[s], s, lane, D, X);
        res[nr++] = D - X;
      }
      if (GMASK & (1 << 7)) {
        baryon_block<T, 1>(P, P, sh.w[s], s, lane, D, X);
        res[nr++] = cmake<T>(2, 0) * D - cmake<T>(4, 0) * X;
      }
      if (GMASK & (1 << 8)) {
        baryon_block<T, 2>(P, P, sh.w[s], s, lane, D, X);
        res[nr++] = cmake<T>(2, 0) * D - cmake<T>(4, 0) * X;
      }
      if (GMASK & (1 << 9)) {
        baryon_block<T, 3>(P, P, sh.w[s], s, lane, D, X);
        res[nr++] = cmake<T>(2, 0) * D - cmake<T>(4, 0) * X;
      }
    }
    // ---- stage outputs [k][lane] per site in the w buffer, then coalesced stores
#pragma unroll
    for (int k = 0; k < NOUT; k++) sh.w[s][lane][k] = res[k];
    __syncthreads();
    {
      const int ls = tid & 15;
      const int64_t x = x0 + ls;
      if (x < V3) {
        cx<T> *o = out + ((int64_t)r * nch_total + ch0) * V3 + x;
        for (int ch = tid >> 4; ch < NOUT * 16; ch += BT_THREADS / BT_SITES) {
          const int k = ch >> 4, l = ch & 15;
          o[(int64_t)ch * V3] = sh.w[ls][l][k];
        }
      }
    }
    __syncthreads();
  }
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

template <typename T, bool TWO, bool MES, int GM>
static int launch_baryon_t(const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out,
                           int64_t V3, int T_lat, int nch_total, int ch0, int tiles_per_block,
                           hipStream_t st) {
  const int ntiles = (int)((V3 + BT_SITES - 1) / BT_SITES);
  int gx = (ntiles + tiles_per_block - 1) / tiles_per_block;
  dim3 grid((unsigned)gx, (unsigned)nrows), block(BT_THREADS);
  size_t shm = sizeof(BaryonShared<T>);
  if (!TWO) shm -= sizeof(cx<T>) * 144 * BT_SITES;  // s2 is the last member
  static bool attr_set = false;
  if (!attr_set) {
    if (hipFuncSetAttribute((const void *)k_baryon<T, TWO, MES, GM>,
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) != hipSuccess)
      return PLG_ERR_HIP;
    attr_set = true;
  }
  k_baryon<T, TWO, MES, GM><<<grid, block, shm, st>>>(props, rows, out, V3, T_lat, nch_total, ch0, ntiles);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

template <typename T>
int launch_baryon(bool two, bool mes, int gmask, const PropDev<T> *props, const RowDev *rows,
                  int nrows, cx<T> *out, int64_t V3, int T_lat, int nch_total, int ch0,
                  int tiles_per_block,
                  hipStream_t st) {
#define PLG_B(TW, ME, GMV)                                                                         \
  if (two == TW && mes == ME && gmask == GMV)                                                      \
    return launch_baryon_t<T, TW, ME, GMV>(props, rows, nrows, out, V3, T_lat, nch_total, ch0,    \
                                           tiles_per_block, st);
  constexpr int NUC = 1, DEL = (1 << 7) | (1 << 8) | (1 << 9), ALL = NUC | DEL;
  PLG_B(true, false, NUC)
  PLG_B(true, false, DEL)
  PLG_B(true, false, ALL)
  PLG_B(false, false, NUC)
  PLG_B(false, false, DEL)
  PLG_B(false, false, ALL)
  PLG_B(false, true, 0)
  PLG_B(false, true, NUC)
  PLG_B(false, true, DEL)
  PLG_B(false, true, ALL)
#undef PLG_B
  return PLG_ERR_INVALID;
}

template int launch_meson<double>(int, const PropDev<double> *, const RowDev *, int, cx<double> *,
                                  int64_t, int, int, int, hipStream_t);
template int launch_meson<float>(int, const PropDev<float> *, const RowDev *, int, cx<float> *,
                                 int64_t, int, int, int, hipStream_t);
template int launch_baryon<double>(bool, bool, int, const PropDev<double> *, const RowDev *, int,
                                   cx<double> *, int64_t, int, int, int, int, hipStream_t);
template int launch_baryon<float>(bool, bool, int, const PropDev<float> *, const RowDev *, int,
                                  cx<float> *, int64_t, int, int, int, int, hipStream_t);

}  // namespace plg
