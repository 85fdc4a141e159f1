// plg_project_bf16.hip -- mixed-precision momentum projection on MFMA.
//
// Momenta come in pairs (n, -n) (a createMom list is symmetric).  With
//   Cr + i Ci = sum_x cos(theta) c(x),   Sr + i Si = sum_x sin(theta) c(x),
//   theta = 2 pi n.x / L,
// the projections exp(-i theta) of both members of a pair are
//   P(n)  = (Cr + Si) + i (Ci - Sr),   P(-n) = (Cr - Si) + i (Ci + Sr),
// so the projection is the real GEMM
//   part[col][row] = sum_x A[row][x] B[x][col]
// with rows {cos, sin} (2j, 2j + 1) of the representative momentum of pair j
// and real columns {Re c, Im c} (2n, 2n + 1) of every site-correlator column:
// half the arithmetic of projecting every momentum on its own.  A momentum
// without a partner in the list is a pair of its own.
//
// Split-bf16 operands on v_mfma_f32_16x16x32_bf16 (K = 32 sites per step):
//   a = a0 + a1 + a2, b = b0 + b1 + b2 (bf16 pieces, ~2^-24 relative)
//   3 products  a0b0 + a0b1 + a1b0               (split = 3)
//   6 products  + a0b2 + a1b1 + a2b0             (split = 6)
// small products first into an FP32 accumulator that is added into FP64
// registers every FLUSH steps (128 sites); FP64 partials per K split.
//
// Workgroup: 256 threads, RB row blocks (16 rows) x 64*NBW real columns over
// one K range.  The phase tile of a step (RB*16 rows x 32 sites, NP pieces)
// is built once in LDS from the phase table and read by the four waves,
// which split the columns (wave w: real column blocks w*NBW + j).
#include <map>

#include "plg_internal.h"

namespace plg {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr int PQ_THREADS = 256;
constexpr int PQ_KT = 32;              // sites per step (the MFMA K)
constexpr int PQ_LD = PQ_KT + 8;       // bf16 per LDS tile row (80 B, 16-byte aligned)
constexpr int PQ_MMAX = 512;           // phase-table entries (10-bit packed momenta)
constexpr int PQ_FLUSH = 4;

// k mod M for 0 <= k < 2^24 with a float reciprocal (exact quotient +-1, fixed up)
__device__ __forceinline__ int mod_m(int k, int M, float invM) {
  int q = (int)((float)k * invM);
  int r = k - q * M;
  r += r < 0 ? M : 0;
  r -= r >= M ? M : 0;
  return r;
}

// two floats -> three packed bf16 pairs p0 + p1 + p2, each rounded to nearest
// even by v_cvt_pk_bf16_f32 (one instruction per pair)
typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 b2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split3(float a, float b, uint32_t *w) {
  const f2v x = {a, b};
  const b2v h0 = __builtin_convertvector(x, b2v);
  const f2v r1 = x - __builtin_convertvector(h0, f2v);
  const b2v h1 = __builtin_convertvector(r1, b2v);
  const f2v r2 = r1 - __builtin_convertvector(h1, f2v);
  const b2v h2 = __builtin_convertvector(r2, b2v);
  w[0] = __builtin_bit_cast(uint32_t, h0);
  w[1] = __builtin_bit_cast(uint32_t, h1);
  w[2] = __builtin_bit_cast(uint32_t, h2);
}

template <int RB, int NBW, int TERMS>
__global__ __launch_bounds__(PQ_THREADS, 2) void k_project_pair(ProjGeom g, const cx<float> *__restrict__ c,
                                                                const uint32_t *__restrict__ wtab,
                                                                double *__restrict__ part) {
  constexpr int NP = TERMS == 3 ? 2 : 3;
  constexpr int NR = 4 * NBW * 16;        // real columns per workgroup
  constexpr int NCX = NR / 2;             // complex columns per workgroup
  constexpr int NPR = RB * 8;             // pairs per row group
  constexpr int AIT = NPR * (PQ_KT / 2);  // phase items (pair, site pair) per step
  constexpr int API = (AIT + PQ_THREADS - 1) / PQ_THREADS;
  constexpr int BIT = NCX * (PQ_KT / 8);  // column items (complex column, 8 sites) per step
  constexpr int BPI = (BIT + PQ_THREADS - 1) / PQ_THREADS;
  __shared__ __align__(16) uint16_t sA[NP][RB * 16][PQ_LD];
  __shared__ __align__(16) uint16_t sB[NP][NR][PQ_LD];
  extern __shared__ __align__(16) uint32_t sW[];  // M x 4 words: cos_p | sin_p << 16 (p < 3), 0

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t cx0 = (int64_t)blockIdx.x * NCX;  // first complex column
  const int split = blockIdx.y;
  const int row0 = blockIdx.z * RB * 16;
  const int R = 2 * g.npairs;
  const int M = g.M;
  const float invM = 1.0f / (float)M;
  // K range of the split, whole steps of 32 sites
  int64_t xb, xe;
  {
    const int64_t per = ((g.V3 + g.nsplit - 1) / g.nsplit + PQ_KT - 1) / PQ_KT * PQ_KT;
    xb = (int64_t)split * per;
    xe = xb + per < g.V3 ? xb + per : g.V3;
  }
  for (int i = tid; i < M; i += PQ_THREADS) {
#pragma unroll
    for (int p = 0; p < 3; p++) sW[4 * i + p] = wtab[6 * i + p];
    sW[4 * i + 3] = 0;
  }
  // phase items of this thread: pair pj[q] of the row group, site pair
  // tid & 15 (the same for every q); packed scaled momentum, -1 = padding row
  int ppk[API], pj[API];
#pragma unroll
  for (int q = 0; q < API; q++) {
    const int it = tid + PQ_THREADS * q;
    pj[q] = it >> 4;
    const int pair = row0 / 2 + pj[q];
    ppk[q] = (it < AIT && pair < g.npairs)
                 ? (g.pair_ms[3 * pair] | (g.pair_ms[3 * pair + 1] << 10) | (g.pair_ms[3 * pair + 2] << 20))
                 : -1;
  }
  const int sp = tid & 15;
  // coordinates of site xt + 2 sp, advanced by one step per iteration
  int ux, uy, uz;
  {
    const int x = (int)(xb + 2 * sp);
    ux = x % g.L[0];
    const int qd = x / g.L[0];
    uy = qd % g.L[1];
    uz = qd / g.L[1];
  }

  f4 acc[RB][NBW];
  double acc64[RB][NBW][4];
#pragma unroll
  for (int i = 0; i < RB; i++)
#pragma unroll
    for (int j = 0; j < NBW; j++) {
      acc[i][j] = f4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; v++) acc64[i][j][v] = 0;
    }

  // site-correlator values of this thread's column items (8 sites each)
  cx<float> bv[BPI][8];
  auto load = [&](int64_t xt) {
#pragma unroll
    for (int q = 0; q < BPI; q++) {
      const int it = tid + PQ_THREADS * q;
      const int64_t n = cx0 + (it >> 2);
      const int64_t x0 = xt + 8 * (it & 3);
      const bool live = it < BIT && n < g.ncols;
      if (live && g.ctile && x0 + 8 <= xe) {
        // 8 sites of one column inside a 16-site tile: 64 contiguous bytes
        const f4 *src = reinterpret_cast<const f4 *>(c + ctile_idx(n, x0, g.ncols));
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const f4 v = src[s];
          bv[q][2 * s] = cmake<float>(v[0], v[1]);
          bv[q][2 * s + 1] = cmake<float>(v[2], v[3]);
        }
      } else {
#pragma unroll
        for (int s = 0; s < 8; s++)
          bv[q][s] = (live && x0 + s < xe) ? c[cidx(g, n, x0 + s)] : cmake<float>(0, 0);
      }
    }
  };

  int step = 0;
  load(xb);
  __syncthreads();  // phase table
  for (int64_t xt = xb; xt < xe; xt += PQ_KT) {
    // ---- column tile: real columns 2n (Re) and 2n + 1 (Im), 8 sites per item
#pragma unroll
    for (int q = 0; q < BPI; q++) {
      const int it = tid + PQ_THREADS * q;
      if (BPI * PQ_THREADS == BIT || it < BIT) {
        const int nl = it >> 2, o = it & 3;
        u4 wr[3], wi[3];
#pragma unroll
        for (int s = 0; s < 4; s++) {
          uint32_t a[3], b[3];
          split3(bv[q][2 * s].re, bv[q][2 * s + 1].re, a);
          split3(bv[q][2 * s].im, bv[q][2 * s + 1].im, b);
#pragma unroll
          for (int p = 0; p < 3; p++) {
            wr[p][s] = a[p];
            wi[p][s] = b[p];
          }
        }
#pragma unroll
        for (int p = 0; p < NP; p++) {
          *reinterpret_cast<u4 *>(&sB[p][2 * nl][8 * o]) = wr[p];
          *reinterpret_cast<u4 *>(&sB[p][2 * nl + 1][8 * o]) = wi[p];
        }
      }
    }
    // ---- phase tile: rows 2j (cos) and 2j + 1 (sin) of pair j at sites
    // xt + 2 sp and xt + 2 sp + 1
    {
      int vx = ux + 1, vy = uy, vz = uz;
      if (vx == g.L[0]) {
        vx = 0;
        if (++vy == g.L[1]) {
          vy = 0;
          ++vz;
        }
      }
      const bool in0 = xt + 2 * sp < xe, in1 = xt + 2 * sp + 1 < xe;
#pragma unroll
      for (int q = 0; q < API; q++) {
        const int it = tid + PQ_THREADS * q;
        if (API * PQ_THREADS == AIT || it < AIT) {
          const int pk = ppk[q];
          const int mx = pk & 1023, my = (pk >> 10) & 1023, mz = (pk >> 20) & 1023;
          const bool ok0 = pk >= 0 && in0, ok1 = pk >= 0 && in1;
          const int k0 = ok0 ? mod_m(mx * ux + my * uy + mz * uz, M, invM) : 0;
          const int k1 = ok1 ? mod_m(mx * vx + my * vy + mz * vz, M, invM) : 0;
          const u4 e0 = *reinterpret_cast<const u4 *>(&sW[4 * k0]);
          const u4 e1 = *reinterpret_cast<const u4 *>(&sW[4 * k1]);
#pragma unroll
          for (int p = 0; p < NP; p++) {
            const uint32_t w0 = ok0 ? e0[p] : 0u, w1 = ok1 ? e1[p] : 0u;
            *reinterpret_cast<uint32_t *>(&sA[p][2 * pj[q]][2 * sp]) = (w0 & 0xFFFFu) | (w1 << 16);
            *reinterpret_cast<uint32_t *>(&sA[p][2 * pj[q] + 1][2 * sp]) = (w0 >> 16) | (w1 & 0xFFFF0000u);
          }
        }
      }
      ux += PQ_KT;
      while (ux >= g.L[0]) {
        ux -= g.L[0];
        if (++uy == g.L[1]) {
          uy = 0;
          ++uz;
        }
      }
    }
    if (xt + PQ_KT < xe) load(xt + PQ_KT);
    __syncthreads();
    // ---- MFMAs, one A piece at a time (RB*NBW independent accumulators
    // between two products into the same one)
    {
      const int kk = 8 * (lane >> 4);
      bf8 bf[NBW][NP];
#pragma unroll
      for (int j = 0; j < NBW; j++)
#pragma unroll
        for (int p = 0; p < NP; p++)
          bf[j][p] = *reinterpret_cast<const bf8 *>(&sB[p][(wave * NBW + j) * 16 + (lane & 15)][kk]);
#pragma unroll
      for (int pa = NP - 1; pa >= 0; pa--) {
        bf8 af[RB];
#pragma unroll
        for (int i = 0; i < RB; i++) af[i] = *reinterpret_cast<const bf8 *>(&sA[pa][i * 16 + (lane & 15)][kk]);
        // B pieces that pair with A piece pa: pa + pb <= 2 (split 6), <= 1 (split 3)
#pragma unroll
        for (int pb = NP - 1 - pa; pb >= 0; pb--)
#pragma unroll
          for (int i = 0; i < RB; i++)
#pragma unroll
            for (int j = 0; j < NBW; j++)
              acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j][pb], acc[i][j], 0, 0, 0);
      }
    }
    if (++step == PQ_FLUSH) {
      step = 0;
#pragma unroll
      for (int i = 0; i < RB; i++)
#pragma unroll
        for (int j = 0; j < NBW; j++) {
#pragma unroll
          for (int v = 0; v < 4; v++) acc64[i][j][v] += (double)acc[i][j][v];
          acc[i][j] = f4{0, 0, 0, 0};
        }
    }
    __syncthreads();
  }
  // D of 16x16x32 (f32): lane l, value v -> row 4*(l/16) + v, column l%16
  const int64_t nreal = 2 * (int64_t)g.ncols;
#pragma unroll
  for (int i = 0; i < RB; i++)
#pragma unroll
    for (int j = 0; j < NBW; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = row0 + i * 16 + 4 * (lane >> 4) + v;
        const int64_t n = 2 * cx0 + (wave * NBW + j) * 16 + (lane & 15);
        if (r < R && n < nreal) part[((int64_t)split * nreal + n) * R + r] = acc64[i][j][v] + (double)acc[i][j][v];
      }
}

// tile codes 100 + 10 RB + NBW (RB row blocks of 16 rows, 64 NBW real columns)
#define PLG_PQ_CODES(X) X(9, 1) X(6, 1) X(5, 1) X(4, 1) X(4, 2) X(3, 2) X(2, 2) X(1, 2)

int launch_project_bf16(const ProjGeom &g, const cx<float> *c, double *part, int terms, const uint32_t *wtab,
                        int code, hipStream_t st) {
  if (g.M > PQ_MMAX) return PLG_ERR_UNSUPPORTED;
  if (g.npairs <= 0 || !g.pair_ms) return PLG_ERR_INVALID;
  const int rb = (code - 100) / 10, nbw = code % 10;
  if (code < 100 || rb <= 0 || nbw <= 0) return PLG_ERR_INVALID;
  const int nblk = (2 * g.npairs + 15) / 16;
  dim3 block(PQ_THREADS);
  dim3 grid((unsigned)((g.ncols + 32 * nbw - 1) / (32 * nbw)), (unsigned)g.nsplit, (unsigned)((nblk + rb - 1) / rb));
  const size_t dyn = (size_t)g.M * 4 * sizeof(uint32_t);
#define PLG_PQ(RB, NBW)                                                                    \
  if (rb == RB && nbw == NBW) {                                                            \
    if (terms == 6) k_project_pair<RB, NBW, 6><<<grid, block, dyn, st>>>(g, c, wtab, part); \
    else k_project_pair<RB, NBW, 3><<<grid, block, dyn, st>>>(g, c, wtab, part);            \
    return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;                         \
  }
  PLG_PQ_CODES(PLG_PQ)
#undef PLG_PQ
  return PLG_ERR_INVALID;
}

#define PLG_PQ_CODE(RB, NBW) 100 + 10 * RB + NBW,
static const int k_pq_codes[] = {PLG_PQ_CODES(PLG_PQ_CODE)};
#undef PLG_PQ_CODE

// tile codes the autotuner tries for a pair count (no row tile taller than the rows)
std::vector<int> proj_candidates(int npairs) {
  const int nblk = (2 * npairs + 15) / 16;
  std::vector<int> v;
  for (int code : k_pq_codes)
    if ((code - 100) / 10 <= std::max(nblk, 1)) v.push_back(code);
  return v;
}

// workgroups of one K split for a tile code
int64_t proj_workgroups(int code, int64_t ncols, int npairs) {
  const int rb = (code - 100) / 10, nbw = code % 10;
  const int nblk = (2 * npairs + 15) / 16;
  return ((ncols + 32 * nbw - 1) / (32 * nbw)) * ((nblk + rb - 1) / rb);
}

// default tile: the fewest padded row blocks, then the tallest row tile
int pick_proj(int npairs) {
  const int nblk = (2 * npairs + 15) / 16;
  int best = 112, bestw = 1 << 30, bestrb = 0;
  for (int code : k_pq_codes) {
    const int rb = (code - 100) / 10;
    if (rb > std::max(nblk, 1)) continue;
    const int waste = ((nblk + rb - 1) / rb) * rb - nblk;
    if (waste < bestw || (waste == bestw && rb > bestrb)) {
      best = code;
      bestw = waste;
      bestrb = rb;
    }
  }
  return best;
}

// Pairs (n, -n) of a momentum list, on the scaled momenta ms (nmom x 3,
// n_k (M / L_k) mod M).  pair_ms: the representative of each pair;
// mom_pair[m] = 2 j + 1 when m is the -n member of pair j, else 2 j.  A
// momentum equal to its own negative (mod M: zero, or components L/2) has a
// zero sine row and stays alone, as does one whose partner is absent or
// already taken.
int make_pairs(const std::vector<int> &ms, int M, std::vector<int> &pair_ms, std::vector<int> &mom_pair) {
  const int nmom = (int)ms.size() / 3;
  pair_ms.clear();
  mom_pair.assign(nmom, -1);
  std::map<std::array<int, 3>, std::vector<int>> open;  // representatives waiting for a partner
  for (int m = 0; m < nmom; m++) {
    const std::array<int, 3> k = {ms[3 * m], ms[3 * m + 1], ms[3 * m + 2]};
    const std::array<int, 3> nk = {(M - k[0]) % M, (M - k[1]) % M, (M - k[2]) % M};
    auto it = open.find(nk);
    if (nk != k && it != open.end() && !it->second.empty()) {
      const int rep = it->second.back();
      it->second.pop_back();
      mom_pair[m] = mom_pair[rep] + 1;
      continue;
    }
    const int j = (int)pair_ms.size() / 3;
    pair_ms.insert(pair_ms.end(), k.begin(), k.end());
    mom_pair[m] = 2 * j;
    if (nk != k) open[k].push_back(m);
  }
  return (int)pair_ms.size() / 3;
}

}  // namespace plg
