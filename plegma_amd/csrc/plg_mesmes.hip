// plg_mesmes.hip -- heavy-light mesons and the four-propagator meson-meson
// (tetraquark-like) contraction on gfx950.
//
// Counterpart of lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:5-61,163-212
// (a table-driven sum over spin/colour entries of S1 S2 conj(S3) conj(S4)),
// in the factorised form of two colour-traced meson blocks plus the crossed
// term.  On 12x12 spin-colour matrices, with X = g5 G (a signed permutation):
//   A_G = G S_L X S_H^dag g5 = G M_G G^-1,   M_G = S_L X S_H^dag X
//   meson(G)  = Tr M_G                          (16 GAMMAS, enum order)
//   mesmes(G) = (Tr M_G)^2 - Tr[M_G M_G]        (G = g5, g1, g2, g3)
// Lane (mu, nu) of a site holds the colour block
//   M'^{ab}_{mu nu} = sum_rho phi(rho) sum_c S_L^{ac}_{mu rho} conj(S_H^{bc}_{nu^x rho^x})
// (M = M' times the lane phase phi(nu^x)); Tr[M M] pairs lane (mu, nu) with
// the transposed lane (nu, mu) by lane permutes.  g5/g3 share X's xor class 0 and
// g1/g2 class 1; within a class the row phases differ by +-i^RA, so one pass
// over rho fills both operators (the same split as the baryon blocks).
// Cost per site: 16 lanes x (2 x 108 + 36 + 36) complex MACs.
#include "plg_tile.h"

namespace plg {

__constant__ MesonX c_mesonx_mm = make_mesonx();

// w holds the meson Gram rows (4 x 16 per site) and then the output stage
// (20 per site); 68 = 64 + 4 keeps the sites' rows on different banks.  The
// crossed terms exchange their blocks by lane permutes, not through LDS: a
// workgroup needs 47.9 KB (FP32), and the kernel is held at two waves per
// SIMD by its registers (the next tile's prefetch among them), which measured
// faster than three waves with a few spilled registers.
constexpr int MM_WS = 68;
template <typename T> struct MesMesShared {
  cx<T> s1[144][BT_SP];
  cx<T> s2[144][BT_SP];
  cx<T> w[BT_SITES][MM_WS];
};

// mesmes of operators OPA, OPB (one xor class of X = g5 G); tr_a / tr_b are
// Tr M of the two operators, held by every lane of the site
template <typename T, int OPA, int OPB>
__device__ __forceinline__ void mesmes_pair(const cx<T> (*SL)[BT_SP], const cx<T> (*SH)[BT_SP], int s, int lane,
                                            cx<T> &mm_a, cx<T> &mm_b) {
  constexpr SPerm XA = sp_mul(G_5, gamma_op(OPA)), XB = sp_mul(G_5, gamma_op(OPB));
  constexpr int x = sp_class(XA);
  static_assert(sp_class(XB) == x, "one xor class per pair");
  constexpr int RA = (XB.ph[0] - XA.ph[0]) & 3;
  const int hi = lane >> 2, lo = lane & 3;
  cx<T> U[2][9];
#pragma unroll
  for (int k = 0; k < 2; k++)
#pragma unroll
    for (int i = 0; i < 9; i++) U[k][i] = cmake<T>(0, 0);
  static_for<0, 4>([&](auto rho) {
    constexpr int r = (XB.ph[rho.value] - XA.ph[rho.value] - RA) & 3;
    static_assert(r == 0 || r == 2, "row phases of a class pair differ by +-i^RA");
    constexpr int acc = r == 2 ? 1 : 0;
    cx<T> a[9], b[9];
#pragma unroll
    for (int i = 0; i < 9; i++) {
      a[i] = SL[pidx(hi, rho.value, 0, 0) + i][s];
      b[i] = SH[pidx(lo ^ x, rho.value ^ x, 0, 0) + i][s];
    }
    static_for<0, 3>([&](auto ca) {
      static_for<0, 3>([&](auto cb) {
        static_for<0, 3>([&](auto cc) {
          cfma_conj_iph<XA.ph[rho.value]>(U[acc][ca.value * 3 + cb.value], a[ca.value * 3 + cc.value],
                                          b[cb.value * 3 + cc.value]);
        });
      });
    });
  });
  // M'_A = U0 + U1, M'_B = i^RA (U0 - U1)
#pragma unroll
  for (int i = 0; i < 9; i++) {
    const cx<T> u0 = U[0][i], u1 = U[1][i];
    U[0][i] = u0 + u1;
    U[1][i] = cmul_iph(u0 - u1, RA);
  }
  // Tr[M M] = sum_lanes phi(nu^x) phi(mu^x) sum_ab M'^{ab}_{mu nu} M'^{ba}_{nu mu}:
  // the transposed lane (nu, mu) of the same site's row sends its blocks
  const int partner = (int)(threadIdx.x & 48) + lo * 4 + hi;
  cx<T> sa = cmake<T>(0, 0), sb = cmake<T>(0, 0);
#pragma unroll
  for (int ca = 0; ca < 3; ca++)
#pragma unroll
    for (int cb = 0; cb < 3; cb++) {
      const cx<T> ta = cmake<T>(__shfl(U[0][cb * 3 + ca].re, partner, 64), __shfl(U[0][cb * 3 + ca].im, partner, 64));
      const cx<T> tb = cmake<T>(__shfl(U[1][cb * 3 + ca].re, partner, 64), __shfl(U[1][cb * 3 + ca].im, partner, 64));
      cfma(sa, U[0][ca * 3 + cb], ta);
      cfma(sb, U[1][ca * 3 + cb], tb);
    }
  const int pha = pick4(XA.ph, lo ^ x) + pick4(XA.ph, hi ^ x);
  const int phb = pick4(XB.ph, lo ^ x) + pick4(XB.ph, hi ^ x);
  sa = cmul_iph_rt(sa, pha);
  sb = cmul_iph_rt(sb, phb);
  mm_a = cmake<T>(rowsum16(sa.re), rowsum16(sa.im));
  mm_b = cmake<T>(rowsum16(sb.re), rowsum16(sb.im));
}

// Output channels per row (from ch0): 16 mesons Tr[G S_L g5 G S_H^dag g5]
// (GAMMAS order), then mesmes for G = g5, g1, g2, g3.
template <typename T>
__global__ __launch_bounds__(BT_THREADS) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 4 ? 2 : 1)))
void k_mesmes(const PropDev<T> *__restrict__ props, const RowDev *__restrict__ rows, cx<T> *__restrict__ out,
              int64_t V3, int T_lat, int nch_total, int ch0, int ntiles) {
  constexpr int NCH = 20;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  MesMesShared<T> &sh = *reinterpret_cast<MesMesShared<T> *>(smem_raw);

  const RowDev row = rows[blockIdx.y];
  const PropDev<T> P1 = props[row.pa];
  const PropDev<T> P2 = props[row.pb];
  const int tid = threadIdx.x;
  const int s = tid >> 4, lane = tid & 15;
  const int hi = lane >> 2, lo = lane & 3;
  const int64_t plane1 = (int64_t)((row.t - P1.t_first + T_lat) % T_lat) * V3;
  const int64_t plane2 = (int64_t)((row.t - P2.t_first + T_lat) % T_lat) * V3;
  // preparation of the loaded entries (boundary sign, rotation)
  const T sg1 = prop_sign(P1, row.t), sg2 = prop_sign(P2, row.t);

  constexpr int NF = 144 / (BT_THREADS / BT_SITES);
  cx<T> pf1[NF], pf2[NF];
  // thread -> site tid % 16 of the tile, components tid / 16 + 16 j: one
  // pointer per propagator, advanced by 16 components per load
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
        pf2[j] = gload(b2);
        b1 += st1;
        b2 += st2;
      }
    } else {
      const bool ok = x0 + (tid & 15) < V3;
#pragma unroll
      for (int j = 0; j < NF; j++) {
        pf1[j] = ok ? gload(b1 + j * st1) : cmake<T>(0, 0);
        pf2[j] = ok ? gload(b2 + j * st2) : cmake<T>(0, 0);
      }
    }
  };
  if (blockIdx.x < ntiles) fetch(blockIdx.x);
  for (int tile = blockIdx.x; tile < ntiles; tile += gridDim.x) {
    const int64_t x0 = (int64_t)tile * BT_SITES;
#pragma unroll
    for (int j = 0; j < NF; j++) {
      sh.s1[(tid >> 4) + 16 * j][tid & 15] = cmake<T>(sg1 * pf1[j].re, sg1 * pf1[j].im);
      sh.s2[(tid >> 4) + 16 * j][tid & 15] = cmake<T>(sg2 * pf2[j].re, sg2 * pf2[j].im);
    }
    __syncthreads();
    if (P1.rot != 0 || P2.rot != 0) {  // rotation to the physical basis on the staged tiles
      if (P1.rot) rot_tile(sh.s1, P1.rot);
      if (P2.rot) rot_tile(sh.s2, P2.rot);
      __syncthreads();
    }
    if (tile + (int)gridDim.x < ntiles) fetch(tile + gridDim.x);
    const cx<T>(*SL)[BT_SP] = sh.s1;
    const cx<T>(*SH)[BT_SP] = sh.s2;

    // ---- 16 meson traces: lane p = (alpha, beta) forms the colour-traced
    // Gram entries S_L[p] . conj(S_H[p ^ (d,d)]), lane op sums them with phases
    cx<T> tr;
    {
      cx<T> pa[9];
#pragma unroll
      for (int i = 0; i < 9; i++) pa[i] = SL[pidx(hi, lo, 0, 0) + i][s];
      cx<T>(*g)[16] = reinterpret_cast<cx<T>(*)[16]>(&sh.w[s][0]);
      static_for<0, 4>([&](auto d) {
        const int q = lane ^ ((d.value << 2) | d.value);
        cx<T> gv = cmake<T>(0, 0);
#pragma unroll
        for (int cc = 0; cc < 9; cc++) cfma_conj(gv, pa[cc], SH[q * 9 + cc][s]);
        g[d.value][lane] = gv;
      });
      wave_lds_sync();
      const int cX = c_mesonx_mm.cls[lane];
      tr = cmake<T>(0, 0);
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
          tr = tr + cmul_iph_rt(g[cX][a * 4 + b], c_mesonx_mm.ph[lane][a ^ cX] + c_mesonx_mm.ph[lane][b]);
      wave_lds_sync();
    }
    // ---- crossed terms of the two class pairs
    cx<T> mm5, mm3, mm1, mm2;
    mesmes_pair<T, 5, 3>(SL, SH, s, lane, mm5, mm3);
    mesmes_pair<T, 1, 2>(SL, SH, s, lane, mm1, mm2);
    // mesmes(G) on lane k of the site (k = 0..3 for G = g5, g1, g2, g3)
    cx<T> mm;
    {
      const int rowbase = threadIdx.x & 48;  // first lane of this site's row in the wavefront
      const int opk = lane == 0 ? 5 : lane == 1 ? 1 : lane == 2 ? 2 : 3;
      cx<T> t;
      t.re = __shfl(tr.re, rowbase + opk, 64);
      t.im = __shfl(tr.im, rowbase + opk, 64);
      const cx<T> crossed = lane == 0 ? mm5 : lane == 1 ? mm1 : lane == 2 ? mm2 : mm3;
      mm = t * t - crossed;
    }
    // ---- stage [k][lane] per site in w, then coalesced stores along x
    cx<T> *stage = &sh.w[0][0];
    stage[s * MM_WS + lane * 2 + 0] = tr;
    stage[s * MM_WS + lane * 2 + 1] = mm;
    __syncthreads();
    {
      const int ls = tid & 15;
      const int64_t x = x0 + ls;
      if (x < V3) {
        const int64_t ncols = (int64_t)gridDim.y * nch_total, col0 = (int64_t)blockIdx.y * nch_total + ch0;
        for (int ch = tid >> 4; ch < NCH; ch += BT_THREADS / BT_SITES)
          gstore(out + ctile_idx(col0 + ch, x, ncols), stage[ls * MM_WS + (ch & 15) * 2 + (ch >> 4)]);
      }
    }
    __syncthreads();
  }
}

template <typename T>
int launch_mesmes(const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out, int64_t V3, int T_lat,
                  int nch_total, int ch0, int tiles_per_block, hipStream_t st) {
  const int ntiles = (int)((V3 + BT_SITES - 1) / BT_SITES);
  const int gx = (ntiles + tiles_per_block - 1) / tiles_per_block;
  dim3 grid((unsigned)gx, (unsigned)nrows), block(BT_THREADS);
  const size_t shm = sizeof(MesMesShared<T>);
  static bool attr_set = false;
  if (!attr_set) {
    if (hipFuncSetAttribute((const void *)k_mesmes<T>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm) !=
        hipSuccess)
      return PLG_ERR_HIP;
    attr_set = true;
  }
  k_mesmes<T><<<grid, block, shm, st>>>(props, rows, out, V3, T_lat, nch_total, ch0, ntiles);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

template int launch_mesmes<double>(const PropDev<double> *, const RowDev *, int, cx<double> *, int64_t, int, int,
                                   int, int, hipStream_t);
template int launch_mesmes<float>(const PropDev<float> *, const RowDev *, int, cx<float> *, int64_t, int, int,
                                  int, int, hipStream_t);

}  // namespace plg
