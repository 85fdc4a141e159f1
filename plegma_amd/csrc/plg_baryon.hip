// plg_baryon.hip -- baryon two-point contractions (and the fused meson +
// baryon kernel of the mixed path) on gfx950.
//
// Replaces PLEGMA_baryons_NtoN.cu:6-46 (proton, Cg5 x Cg5) and
// PLEGMA_baryons_deltas.cu:61-110 (Delta iso-3/2, Cg_k x g4 (Cg_k)^dag g4)
// with factorised forms.  For a diquark matrix A (Abar = g4 A^dag g4) and
// propagators P (quarks outside the diquark's partner) and Qm:
//   St^{bb'}_{nu nu'} = A_{nu lam} Qm^{bb'}_{lam lam'} Abar_{lam' nu'}
//   Q^{cc'}  = eps_abc eps_a'b'c' sum_{nu nu'} P^{aa'}_{nu nu'} St^{bb'}_{nu nu'}
//   D_{mu mu'} = sum_{cc'} Q^{cc'} P^{cc'}_{mu mu'}
//   X_{mu mu'} = eps eps' sum_{nu nu'} P^{ca'}_{mu nu'} St^{bb'}_{nu nu'} P^{ac'}_{nu mu'}
//   proton  (Qm = other flavour)      = D - X
//   Delta_k (uuu, Cg_k symmetric)     = 2 D - 4 X
// The reference bodies cost 27648 complex MACs per proton dataset and about
// 99000 per Delta component; this form costs 2304 per block (all four blocks
// of one dataset: 9216), with 16 lanes of a wave per lattice site:
//
//   step 1, lane (nu, nu'):  Q partials, DPP row sum over the 16 lanes, D
//   step 2, lane (nu', mu'): W^{c a'}_{nu' mu'} = sum_nu eps eps' St P   -> LDS
//   step 3, lane (mu, mu'):  X = sum_{nu' c a'} P^{c a'}_{mu nu'} W_{nu' mu'}
//
// The four diquarks Cg5, Cg1, Cg2, Cg3 are signed permutations of two xor
// classes; blocks of one class that share Qm (proton and Delta_3, Delta_1 and
// Delta_2 on one propagator) differ only by phases, so their step-2 colour
// contraction T(nu) is computed once and accumulated with both blocks' row
// phases (compile-time); lane phases are applied once per block.
#include "plg_tile.h"

namespace plg {

// LDS of one workgroup.  w holds the exchange intermediates of a block pair,
// the meson Gram rows before the blocks and the output stage after them.
// s2 is last so the one-propagator variant is allocated without it.
template <typename T> struct BaryonShared {
  cx<T> s1[144][BT_SP];
  cx<T> w[BT_SITES][BT_WS];
  cx<T> s2[144][BT_SP];
};

__constant__ MesonX c_mesonx = make_mesonx();

// eps eps' contraction of two colour blocks x[a][a'] y[b][b'] into out[c][c'] (+=),
// every term one complex FMA
template <typename T> __device__ __forceinline__ void eps2(const cx<T> *x, const cx<T> *y, cx<T> *out) {
  static_for<0, 6>([&](auto e1) {
    static_for<0, 6>([&](auto e2) {
      constexpr int a = EPS[e1.value][0], b = EPS[e1.value][1], c = EPS[e1.value][2];
      constexpr int ap = EPS[e2.value][0], bp = EPS[e2.value][1], cp = EPS[e2.value][2];
      constexpr int ph = SGN_EPS[e1.value] * SGN_EPS[e2.value] > 0 ? 0 : 2;
      cfma_iph<ph>(out[c * 3 + cp], x[a * 3 + ap], y[b * 3 + bp]);
    });
  });
}
// exchange form: out[c][a'] += i^PH eps eps' st[b][b'] p[a][c']
template <int PH, typename T>
__device__ __forceinline__ void eps2x(const cx<T> *st, const cx<T> *p, cx<T> *out) {
  static_for<0, 6>([&](auto e1) {
    static_for<0, 6>([&](auto e2) {
      constexpr int a = EPS[e1.value][0], b = EPS[e1.value][1], c = EPS[e1.value][2];
      constexpr int ap = EPS[e2.value][0], bp = EPS[e2.value][1], cp = EPS[e2.value][2];
      constexpr int ph = PH + (SGN_EPS[e1.value] * SGN_EPS[e2.value] > 0 ? 0 : 2);
      cfma_iph<ph>(out[c * 3 + ap], st[b * 3 + bp], p[a * 3 + cp]);
    });
  });
}

// D and X of diquark W0 and, if W1 >= 0, of W1 (same spin class, same Qm).
// pa: this lane's P[hi][lo] colour block (lane = (nu, nu') = (mu, mu')).
template <typename T, int W0, int W1>
__device__ __forceinline__ void baryon_group(const cx<T> (*P)[BT_SP], const cx<T> (*Qm)[BT_SP],
                                             cx<T> (*Wsh)[16][9], int s, int lane, const cx<T> *pa,
                                             cx<T> *D, cx<T> *X) {
  constexpr int NB = W1 >= 0 ? 2 : 1;
  constexpr SPerm A0 = diquark(W0), Ab0 = diquark_bar(W0);
  constexpr SPerm A1 = diquark(W1 >= 0 ? W1 : W0), Ab1 = diquark_bar(W1 >= 0 ? W1 : W0);
  constexpr int c = sp_class(A0);
  static_assert(sp_class(A1) == c && sp_class(Ab0) == c && sp_class(Ab1) == c, "one spin class per group");
  const int hi = lane >> 2, lo = lane & 3;

  // ---- step 1: lane (nu, nu') -> Q partial, row sum, D on lane (mu, mu')
  {
    cx<T> q[9], Qr[9];
#pragma unroll
    for (int i = 0; i < 9; i++) {
      q[i] = Qm[pidx(hi ^ c, lo ^ c, 0, 0) + i][s];
      Qr[i] = cmake<T>(0, 0);
    }
    eps2(pa, q, Qr);
    static_for<0, NB>([&](auto k) {
      constexpr SPerm A = k.value == 0 ? A0 : A1;
      constexpr SPerm Ab = k.value == 0 ? Ab0 : Ab1;
      int ph = 0;
      // lane phase A[nu][nu^c] * Abar[nu'^c][nu']
#pragma unroll
      for (int i = 0; i < 4; i++) ph += (hi == i ? A.ph[i] : 0) + (lo == i ? Ab.ph[i ^ c] : 0);
      cx<T> d = cmake<T>(0, 0);
#pragma unroll
      for (int i = 0; i < 9; i++) {
        cx<T> v = cmul_iph_rt(Qr[i], ph);
        v.re = rowsum16(v.re);
        v.im = rowsum16(v.im);
        cfma(d, v, pa[i]);
      }
      D[k.value] = d;
    });
  }
  // ---- step 2: lane (nu', mu') -> W = sum_nu A-phase(nu) T(nu); the
  // Abar-phase of row nu' is applied by the readers in step 3
  // The second block's row phases are A1.ph = A0.ph + RA or + RA + 2 (mod 4),
  // so U_0 and U_1 (the A0-phased T(nu) of the two kinds of rows) are
  // accumulated by FMA and W_0 = U_0 + U_1, W_1 = i^RA (U_0 - U_1).
  {
    constexpr int RA = (A1.ph[0] - A0.ph[0]) & 3;
    cx<T> W[NB][9];
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int i = 0; i < 9; i++) W[b][i] = cmake<T>(0, 0);
    static_for<0, 4>([&](auto nu) {
      constexpr int r = (A1.ph[nu.value] - A0.ph[nu.value] - RA) & 3;
      static_assert(r == 0 || r == 2, "row phases of a class pair differ by +-i^RA");
      constexpr int acc = (NB == 2 && r == 2) ? 1 : 0;
      cx<T> sq[9], pp[9];
#pragma unroll
      for (int i = 0; i < 9; i++) {
        sq[i] = Qm[pidx(nu.value ^ c, hi ^ c, 0, 0) + i][s];
        pp[i] = P[pidx(nu.value, lo, 0, 0) + i][s];
      }
      eps2x<A0.ph[nu.value]>(sq, pp, W[acc]);
    });
    if constexpr (NB == 2) {
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const cx<T> u0 = W[0][i], u1 = W[1][i];
        W[0][i] = u0 + u1;
        W[1][i] = cmul_iph(u0 - u1, RA);
      }
    }
    // shared without this lane's Abar phase: the reader knows the row
    // (nu' = its loop index) and folds the phase into its FMAs
#pragma unroll
    for (int b = 0; b < NB; b++)
#pragma unroll
      for (int i = 0; i < 9; i++) Wsh[b][lane][i] = W[b][i];
  }
  wave_lds_sync();
  // ---- step 3: lane (mu, mu') -> X = sum_{nu'} P[mu][nu'] . W[nu'][mu']
  {
    cx<T> x[NB];
#pragma unroll
    for (int b = 0; b < NB; b++) x[b] = cmake<T>(0, 0);
    static_for<0, 4>([&](auto nu) {
      constexpr int nup = nu.value;
      // the Abar phase of the writing lane (row nu')
      constexpr int ph0 = Ab0.ph[nup ^ c], ph1 = Ab1.ph[nup ^ c];
#pragma unroll
      for (int i = 0; i < 9; i++) {
        const cx<T> p = P[pidx(hi, nup, 0, 0) + i][s];
        cfma_iph<ph0>(x[0], p, Wsh[0][nup * 4 + lo][i]);
        if constexpr (NB == 2) cfma_iph<ph1>(x[1], p, Wsh[1][nup * 4 + lo][i]);
      }
    });
#pragma unroll
    for (int b = 0; b < NB; b++) X[b] = x[b];
  }
  wave_lds_sync();  // W is overwritten by the next group of this wavefront
}

// Output channels per row (from ch0): [16 mesons of S1 if MES], then for each
// dataset d (2 if TWO) and each selected group g in the reference order
// (nucl_nucl, deltapp_deltamm_11, _22, _33): 16 components mu*4 + mu'.
// FP32 tiles fit two workgroups per CU in LDS, so their register budget is
// capped at two waves per SIMD.
template <typename T, bool TWO, bool MES, int GMASK>
__global__ __launch_bounds__(BT_THREADS) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
void k_baryon(const PropDev<T> *__restrict__ props,
                                                       const RowDev *__restrict__ rows,
                                                       cx<T> *__restrict__ out, int64_t V3, int T_lat,
                                                       int nch_total, int ch0, int ntiles) {
  constexpr bool H0 = GMASK & 1, H1 = GMASK & (1 << 7), H2 = GMASK & (1 << 8), H3 = GMASK & (1 << 9);
  constexpr int NG = (int)H0 + (int)H1 + (int)H2 + (int)H3;
  constexpr int R0 = 0, R1 = (int)H0, R2 = (int)H0 + (int)H1, R3 = (int)H0 + (int)H1 + (int)H2;
  constexpr int ND = TWO ? 2 : 1;
  constexpr int MOFF = MES ? 1 : 0;
  constexpr int NOUT = MOFF + ND * NG;
  static_assert(NOUT <= 9, "output stage holds 9 values per lane");
  extern __shared__ __align__(16) unsigned char smem_raw[];
  BaryonShared<T> &sh = *reinterpret_cast<BaryonShared<T> *>(smem_raw);

  const RowDev row = rows[blockIdx.y];
  const PropDev<T> P1 = props[row.pa];
  const PropDev<T> P2 = props[TWO ? row.pb : row.pa];
  const int tid = threadIdx.x;
  const int s = tid >> 4, lane = tid & 15;
  const int hi = lane >> 2, lo = lane & 3;
  const int64_t plane1 = (int64_t)((row.t - P1.t_first + T_lat) % T_lat) * V3;
  const int64_t plane2 = (int64_t)((row.t - P2.t_first + T_lat) % T_lat) * V3;
  // preparation of the loaded entries (boundary sign, rotation)
  const T sg1 = prop_sign(P1, row.t), sg2 = prop_sign(P2, row.t);
  const bool prep_rot = P1.rot != 0 || (TWO && P2.rot != 0);

  // tile staging: thread -> site tid%16, components tid/16 + 16 j (j < 9).
  // The components of tile t+1 are fetched into registers while tile t is
  // contracted (coalesced: 16 consecutive sites per component).
  constexpr int NF = 144 / (BT_THREADS / BT_SITES);
  cx<T> pf1[NF], pf2[TWO ? NF : 1];
  // one pointer per propagator, advanced by 16 components per load
  const cx<T> *f1 = P1.data + plane1 + (int64_t)(tid >> 4) * P1.cs + (tid & 15);
  const cx<T> *f2 = P2.data + plane2 + (int64_t)(tid >> 4) * P2.cs + (tid & 15);
  const int64_t st1 = 16 * P1.cs, st2 = 16 * P2.cs;
  auto fetch = [&](int tl) {
    const int64_t x0 = (int64_t)tl * BT_SITES;
    const cx<T> *b1 = f1 + x0, *b2 = f2 + x0;
    if (x0 + BT_SITES <= V3) {
#pragma unroll
      for (int j = 0; j < NF; j++) {
        pf1[j] = gload(b1);
        if constexpr (TWO) pf2[j] = gload(b2);
        b1 += st1;
        b2 += st2;
      }
    } else {
      const bool ok = x0 + (tid & 15) < V3;
#pragma unroll
      for (int j = 0; j < NF; j++) {
        pf1[j] = ok ? gload(b1 + j * st1) : cmake<T>(0, 0);
        if constexpr (TWO) pf2[j] = ok ? gload(b2 + j * st2) : cmake<T>(0, 0);
      }
    }
  };
  if (blockIdx.x < ntiles) fetch(blockIdx.x);
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t x0 = (int64_t)tile * BT_SITES;
#pragma unroll
    for (int j = 0; j < NF; j++) {
      sh.s1[(tid >> 4) + 16 * j][tid & 15] = cmake<T>(sg1 * pf1[j].re, sg1 * pf1[j].im);
      if constexpr (TWO) sh.s2[(tid >> 4) + 16 * j][tid & 15] = cmake<T>(sg2 * pf2[j].re, sg2 * pf2[j].im);
    }
    __syncthreads();
    if (prep_rot) {  // rotation to the physical basis on the staged tiles
      if (P1.rot) rot_tile(sh.s1, P1.rot);
      if (TWO && P2.rot) rot_tile(sh.s2, P2.rot);
      __syncthreads();
    }
    if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
    const cx<T>(*S1)[BT_SP] = sh.s1;
    const cx<T>(*S2)[BT_SP] = TWO ? sh.s2 : sh.s1;

    cx<T> res[NOUT];
    cx<T> pa1[9];
#pragma unroll
    for (int i = 0; i < 9; i++) pa1[i] = S1[pidx(hi, lo, 0, 0) + i][s];
    // ---- mesons: lane p = (alpha, beta) forms G[p][p ^ (d,d)], lane op sums
    if constexpr (MES) {
      cx<T>(*g)[16] = reinterpret_cast<cx<T>(*)[16]>(&sh.w[s][0]);
      static_for<0, 4>([&](auto d) {
        const int q = lane ^ ((d.value << 2) | d.value);
        cx<T> gv = cmake<T>(0, 0);
#pragma unroll
        for (int cc = 0; cc < 9; cc++) cfma_conj(gv, pa1[cc], d.value == 0 ? pa1[cc] : S1[q * 9 + cc][s]);
        g[d.value][lane] = gv;
      });
      wave_lds_sync();
      const int cX = c_mesonx.cls[lane];
      cx<T> acc = cmake<T>(0, 0);
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
          acc = acc + cmul_iph_rt(g[cX][a * 4 + b], c_mesonx.ph[lane][a ^ cX] + c_mesonx.ph[lane][b]);
      res[0] = acc;
      wave_lds_sync();
    }
    // ---- baryon blocks per dataset
    static_for<0, ND>([&](auto dd) {
      constexpr int d = dd.value;
      const cx<T>(*P)[BT_SP] = d == 0 ? S1 : S2;
      const cx<T>(*Qo)[BT_SP] = d == 0 ? S2 : S1;  // the proton's other flavour
      cx<T> pa[9];
#pragma unroll
      for (int i = 0; i < 9; i++) pa[i] = d == 0 ? pa1[i] : P[pidx(hi, lo, 0, 0) + i][s];
      constexpr int base = MOFF + d * NG;
      cx<T> D[2], X[2];
      // class of Cg5 and Cg3: proton with Delta_3 when both use this propagator
      if constexpr (H0 && H3 && !TWO) {
        baryon_group<T, 0, 3>(P, P, reinterpret_cast<cx<T>(*)[16][9]>(&sh.w[s][0]), s, lane, pa, D, X);
        res[base + R0] = D[0] - X[0];
        res[base + R3] = cmake<T>(2, 0) * D[1] - cmake<T>(4, 0) * X[1];
      } else {
        if constexpr (H0) {
          baryon_group<T, 0, -1>(P, Qo, reinterpret_cast<cx<T>(*)[16][9]>(&sh.w[s][0]), s, lane, pa, D, X);
          res[base + R0] = D[0] - X[0];
        }
        if constexpr (H3) {
          baryon_group<T, 3, -1>(P, P, reinterpret_cast<cx<T>(*)[16][9]>(&sh.w[s][0]), s, lane, pa, D, X);
          res[base + R3] = cmake<T>(2, 0) * D[0] - cmake<T>(4, 0) * X[0];
        }
      }
      // class of Cg1 and Cg2
      if constexpr (H1 && H2) {
        baryon_group<T, 1, 2>(P, P, reinterpret_cast<cx<T>(*)[16][9]>(&sh.w[s][0]), s, lane, pa, D, X);
        res[base + R1] = cmake<T>(2, 0) * D[0] - cmake<T>(4, 0) * X[0];
        res[base + R2] = cmake<T>(2, 0) * D[1] - cmake<T>(4, 0) * X[1];
      } else {
        if constexpr (H1) {
          baryon_group<T, 1, -1>(P, P, reinterpret_cast<cx<T>(*)[16][9]>(&sh.w[s][0]), s, lane, pa, D, X);
          res[base + R1] = cmake<T>(2, 0) * D[0] - cmake<T>(4, 0) * X[0];
        }
        if constexpr (H2) {
          baryon_group<T, 2, -1>(P, P, reinterpret_cast<cx<T>(*)[16][9]>(&sh.w[s][0]), s, lane, pa, D, X);
          res[base + R2] = cmake<T>(2, 0) * D[0] - cmake<T>(4, 0) * X[0];
        }
      }
    });
    // ---- stage [k][lane] per site in w, then coalesced stores along x
    cx<T> *stage = &sh.w[0][0];
    constexpr int SST = BT_WS;  // complex per site in w
#pragma unroll
    for (int k = 0; k < NOUT; k++) stage[s * SST + lane * 9 + k] = res[k];
    __syncthreads();
    {
      const int ls = tid & 15;
      const int64_t x = x0 + ls;
      if (x < V3) {
        const int64_t ncols = (int64_t)gridDim.y * nch_total, col0 = (int64_t)blockIdx.y * nch_total + ch0;
        for (int ch = tid >> 4; ch < NOUT * 16; ch += BT_THREADS / BT_SITES)
          gstore(out + ctile_idx(col0 + ch, x, ncols), stage[ls * SST + (ch & 15) * 9 + (ch >> 4)]);
      }
    }
    __syncthreads();
  }
}

template <typename T, bool TWO, bool MES, int GM>
static int launch_baryon_t(const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out, int64_t V3,
                           int T_lat, int nch_total, int ch0, int tiles_per_block, hipStream_t st) {
  const int ntiles = (int)((V3 + BT_SITES - 1) / BT_SITES);
  const int gx = (ntiles + tiles_per_block - 1) / tiles_per_block;
  dim3 grid((unsigned)gx, (unsigned)nrows), block(BT_THREADS);
  size_t shm = sizeof(BaryonShared<T>);
  if (!TWO) shm -= sizeof(cx<T>) * 144 * BT_SP;  // s2 is the last member
  static bool attr_set = false;
  if (!attr_set) {
    if (hipFuncSetAttribute((const void *)k_baryon<T, TWO, MES, GM>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)shm) != hipSuccess)
      return PLG_ERR_HIP;
    attr_set = true;
  }
  k_baryon<T, TWO, MES, GM><<<grid, block, shm, st>>>(props, rows, out, V3, T_lat, nch_total, ch0, ntiles);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

template <typename T>
int launch_baryon(bool two, bool mes, int gmask, const PropDev<T> *props, const RowDev *rows, int nrows,
                  cx<T> *out, int64_t V3, int T_lat, int nch_total, int ch0, int tiles_per_block,
                  hipStream_t st) {
#define PLG_B(TW, ME, GMV)                                                                       \
  if (two == TW && mes == ME && gmask == GMV)                                                    \
    return launch_baryon_t<T, TW, ME, GMV>(props, rows, nrows, out, V3, T_lat, nch_total, ch0,  \
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

template int launch_baryon<double>(bool, bool, int, const PropDev<double> *, const RowDev *, int, cx<double> *,
                                   int64_t, int, int, int, int, hipStream_t);
template int launch_baryon<float>(bool, bool, int, const PropDev<float> *, const RowDev *, int, cx<float> *,
                                  int64_t, int, int, int, int, hipStream_t);

}  // namespace plg
