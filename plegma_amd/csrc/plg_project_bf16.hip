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
// without a partner in the list is a pair of its own.  The zero momentum (all
// cos 1, all sin 0) is not a GEMM row: when the list holds it, it is the last
// pair and the producer waves of the first row group sum the columns
// directly.
//
// Split-bf16 operands on v_mfma_f32_16x16x32_bf16 (K = 32 sites per step):
//   a = a0 + a1 + a2, b = b0 + b1 + b2 (bf16 pieces, ~2^-24 relative)
//   3 products  a0b0 + a0b1 + a1b0               (split = 3)
//   6 products  + a0b2 + a1b1 + a2b0             (split = 6)
// small products first into an FP32 accumulator that is added into FP64
// registers every FLUSH steps (128 sites); FP64 partials per K split.
//
// The phase matrix does not depend on the column (sources and time slices
// take their source phase in the finalize step).  Its integer phase indices
// k = n_j . x (M / L) mod M are built once per momentum list
// (k_phase_index), step-major,
//   pidx[(kstep * PP + j) * 32 + s],  PP = pairs rounded to 8,
// and the workgroups expand them through an LDS table of the bf16 pieces of
// cos/sin(2 pi k / M): 2 bytes read per phase instead of 12.
// Workgroup: 256 threads, RB row blocks (16 rows) x 64*NBW real columns over
// one K range.  Per step it stages the phase tile (RB*16 rows x 32 sites, NP
// pieces) and the split column tile in LDS; the four waves split the columns
// (wave w: real column blocks w*NBW + j).  Global loads of step s + 1 are in
// flight during the MFMAs of step s.
#include <map>
#include <type_traits>

#include "plg_internal.h"

namespace plg {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr int PQ_THREADS = 256;
constexpr int PQ_KT = 32;              // sites per step (the MFMA K)

// LDS tiles hold 64-byte rows (32 bf16 sites) without padding; the four
// 16-byte quarters q of row r sit at quarter q ^ ((r >> 1) & 3).  With that
// swizzle the MFMA fragment access (lane l: row l % 16, quarter l / 16) and
// the tile stores, which use the same lane pattern, are free of LDS bank
// conflicts on gfx950 (measured: the plain and row-padded layouts take 2-3x
// the cycles).
__device__ __forceinline__ int lds_q(int r, int q) { return r * PQ_KT + 8 * (q ^ ((r >> 1) & 3)); }
constexpr int PQ_FLUSH = 4;
constexpr int PQ_INFLIGHT = 3;         // register sets of steps loaded ahead by a producer

constexpr int PQ_MMAX = 512;           // phase-table entries in LDS

// pair rows of the phase index table (the phase rows 2j, 2j + 1 rounded to 16)
__host__ __device__ __forceinline__ int64_t pidx_pairs(int npairs) { return (npairs + 7) / 8 * 8; }

size_t phase_index_bytes(int64_t V3, int npairs) {
  return (size_t)((V3 + PQ_KT - 1) / PQ_KT) * pidx_pairs(npairs) * PQ_KT * sizeof(uint16_t);
}

// one thread per (step, pair, site of the step): k = n_j . x (M / L) mod M,
// exact in integers; 0 for padding pairs and sites past the volume
__global__ void k_phase_index(ProjGeom g, uint16_t *__restrict__ tab) {
  const int64_t PP = pidx_pairs(g.npairs);
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nk = (g.V3 + PQ_KT - 1) / PQ_KT;
  if (e >= nk * PP * PQ_KT) return;
  const int s = (int)(e % PQ_KT);
  const int j = (int)((e / PQ_KT) % PP);
  const int64_t x = (e / (PQ_KT * PP)) * PQ_KT + s;
  int64_t k = 0;
  if (x < g.V3 && j < g.npairs) {
    const int64_t ux = x % g.L[0], uy = (x / g.L[0]) % g.L[1], uz = x / ((int64_t)g.L[0] * g.L[1]);
    k = (g.pair_ms[3 * j] * ux + g.pair_ms[3 * j + 1] * uy + g.pair_ms[3 * j + 2] * uz) % g.M;
  }
  tab[e] = (uint16_t)k;
}

int launch_phase_index(const ProjGeom &g, uint16_t *tab, hipStream_t st) {
  if (g.npairs <= 0 || !g.pair_ms) return PLG_ERR_INVALID;
  const int64_t n = (g.V3 + PQ_KT - 1) / PQ_KT * pidx_pairs(g.npairs) * PQ_KT;
  k_phase_index<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(g, tab);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

// column-major field f[n][x] (n < ncols, x < V3) -> tile-major site
// correlator layout (ctile_idx), zero at the padding sites of the last tile
__global__ void k_to_tiles(const cx<float> *__restrict__ f, int64_t ncols, int64_t V3, cx<float> *__restrict__ t) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t S = ctile_sites(V3);
  if (e >= ncols * S) return;
  const int64_t n = e / S, x = e % S;
  t[ctile_idx(n, x, ncols)] = x < V3 ? f[n * V3 + x] : cmake<float>(0, 0);
}

int launch_to_tiles(const cx<float> *f, int64_t ncols, int64_t V3, cx<float> *t, hipStream_t st) {
  const int64_t n = ncols * ctile_sites(V3);
  k_to_tiles<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st>>>(f, ncols, V3, t);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
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

// Per-column shift a[n] of the mixed projection.  Any value is exact (the
// phases of n != 0 sum to zero over the volume); a value near the column's
// mean keeps the FP32 accumulators from carrying its coherent part, which
// otherwise leaves an error of order 2^-24 |C(0)| at momenta with some
// n_k = 0 (a large mean: the pion's sum of |S|^2).  From 256 sites (16 site
// tiles spread through the volume), per real part: mean m and spread s in
// FP64; a = 0 unless |m| > s / 4 (four standard errors), else m rounded to 8
// significant bits, so that c - a is exact in FP32 for every value within
// 2^16 |a| of a (a rounding of c - a that depends on a's low bits would be
// the same at every site, a coherent error of V3 ulp).
__device__ __forceinline__ float col_shift_part(double m, double s2) {
  if (m * m * 16.0 <= s2) return 0.0f;
  uint32_t u = __float_as_uint((float)m);
  u = (u + 0x7FFFu + ((u >> 16) & 1u)) & 0xFFFF0000u;
  return __uint_as_float(u);
}
// 256 threads per 16 columns: thread (column j, site s of a tile) reads site
// s of 16 site tiles spread through the volume (the 16 columns of one tile
// are 2 KB contiguous, read by the workgroup at once); the 16 sites of a
// column are combined by shuffles within the wave.
constexpr int CS_TILES = 16, CS_COLS = 16;
__global__ __launch_bounds__(256) void k_col_shift(const cx<float> *__restrict__ c, int64_t ncols, int64_t V3,
                                                   cx<float> *__restrict__ a) {
  const int s = threadIdx.x & 15, j = threadIdx.x >> 4;
  const int64_t n = (int64_t)blockIdx.x * CS_COLS + j;
  const int64_t ntiles = V3 / 16;  // whole tiles only
  double re = 0, im = 0, re2 = 0, im2 = 0;
  int64_t cnt = 0;
  if (n < ncols) {
    if (ntiles > 0) {
      const int nt = ntiles < CS_TILES ? (int)ntiles : CS_TILES;
      for (int i = 0; i < nt; i++) {
        const cx<float> v = c[ctile_idx(n, (int64_t)i * ntiles / nt * 16 + s, ncols)];
        re += v.re;
        im += v.im;
        re2 += (double)v.re * v.re;
        im2 += (double)v.im * v.im;
      }
      cnt = (int64_t)nt * 16;
    } else if (s < V3) {  // fewer than 16 sites
      const cx<float> v = c[ctile_idx(n, s, ncols)];
      re = v.re;
      im = v.im;
      re2 = (double)v.re * v.re;
      im2 = (double)v.im * v.im;
      cnt = V3;
    }
  }
#pragma unroll
  for (int off = 8; off >= 1; off >>= 1) {  // the 16 lanes of a column
    re += __shfl_xor(re, off, 16);
    im += __shfl_xor(im, off, 16);
    re2 += __shfl_xor(re2, off, 16);
    im2 += __shfl_xor(im2, off, 16);
  }
  if (s == 0 && n < ncols) {
    const double ns = (double)cnt;
    const double mr = re / ns, mi = im / ns;
    a[n] = cmake<float>(col_shift_part(mr, re2 / ns - mr * mr), col_shift_part(mi, im2 / ns - mi * mi));
  }
}

// Warp-specialised workgroup of 512 threads: waves 0-3 run the MFMAs of
// step k on one stage of the LDS tiles while waves 4-7 (the producers) split
// step k + 1 into the other stage; one barrier per step.  The producers keep
// the global loads of the next PQ_INFLIGHT steps in flight in a ring of
// register sets (a load takes about a microsecond under load, several
// steps).  A consumer and a producer wave share each SIMD, so the
// producers' VALU work runs in the shadow of the consumers' MFMAs.
// NCW consumer waves (4: one per SIMD; 8: two per SIMD with half the
// accumulators each) and 4 producer waves.
template <int WR, int WC, int GR, int TERMS, int NCW>
__global__ __launch_bounds__((NCW + 4) * 64, 1) void k_project_pair(ProjGeom g, const cx<float> *__restrict__ c,
                                                                    const uint32_t *__restrict__ wtab, double *__restrict__ part,
                                                                    const cx<float> *__restrict__ shift,
                                                                    int nrg, int ncg) {
  constexpr int NP = TERMS == 3 ? 2 : 3;
  constexpr int GC = NCW / GR;            // consumer waves along the columns
  constexpr int RB = WR * GR;             // row blocks per workgroup
  constexpr int NR = WC * GC * 16;        // real columns per workgroup
  constexpr int NCX = NR / 2;             // complex columns per workgroup
  // phase items (pair, 8 sites) per step, in groups of 16 pairs x 4 octets
  constexpr int AIT = (RB * 8 + 15) / 16 * 64;
  constexpr int API = (AIT + PQ_THREADS - 1) / PQ_THREADS;
  constexpr int BIT = NCX * (PQ_KT / 8);  // column items (complex column, 8 sites) per step
  constexpr int BPI = (BIT + PQ_THREADS - 1) / PQ_THREADS;
  constexpr int ASZ = RB * 16 * PQ_KT, BSZ = NR * PQ_KT;  // bf16 per piece of a stage
  static_assert(NCX % 16 == 0, "column tile of whole 16-column groups");
  // two stages of tiles (step k in stage k % 2), pieces p at p * ASZ / p * BSZ.
  // B rows: real part of the tile's complex column n at row n, imaginary
  // part at row NCX + n
  __shared__ __align__(16) uint16_t sA[2][NP * ASZ];
  __shared__ __align__(16) uint16_t sB[2][NP * BSZ];
  __shared__ u4 junk[64];
  __shared__ u4 sW[PQ_MMAX];  // entry k: cos_p | sin_p << 16 of cos/sin(2 pi k / M), p < 3

  // Workgroups are dispatched round-robin over the 8 XCDs (blockIdx.x mod
  // 8).  The row groups of one column group go to the same XCD at nearly the
  // same time, so that their common column tile is read from HBM once and
  // then from that XCD's L2.
  const int xcd = blockIdx.x & 7, slot = blockIdx.x >> 3;
  const int rg = slot % nrg, cg = (slot / nrg) * 8 + xcd;
  if (cg >= ncg) return;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const bool producer = wave >= NCW;
  const int64_t cx0 = (int64_t)cg * NCX;  // first complex column
  const int split = blockIdx.y;
  const int row0 = rg * RB * 16;
  const int R = 2 * g.npairs;               // partial rows (pairs x cos/sin)
  const int Rg = R - 2 * g.zero_pair;        // rows of the GEMM
  const bool zsum = g.zero_pair && rg == 0;  // this workgroup sums the columns for n = 0
  const int64_t PP = pidx_pairs(g.npairs);
  // K range of the split, whole steps of 32 sites
  int64_t xb, xe;
  {
    const int64_t per = ((g.V3 + g.nsplit - 1) / g.nsplit + PQ_KT - 1) / PQ_KT * PQ_KT;
    xb = (int64_t)split * per;
    xe = xb + per < g.V3 ? xb + per : g.V3;
  }
  const int nk = xe > xb ? (int)((xe - xb + PQ_KT - 1) / PQ_KT) : 0;
  for (int i = threadIdx.x; i < g.M; i += (NCW + 4) * 64) sW[i] = u4{wtab[6 * i], wtab[6 * i + 1], wtab[6 * i + 2], 0};
  __syncthreads();

  if (producer) {
    const int ptid = threadIdx.x - NCW * 64;
    // phase item q: pair aj of the tile, sites 8 ao .. 8 ao + 7 (rows 2 aj
    // cos, 2 aj + 1 sin); reads clamped into the index table (pairs past it
    // feed outputs that are not stored), items past the tile store into junk
    int aj[API], ao[API];
    bool alive[API];
    const uint16_t *asrc[API];
    const int64_t astep = PP * PQ_KT;
#pragma unroll
    for (int q = 0; q < API; q++) {
      const int it = ptid + PQ_THREADS * q;
      aj[q] = (it & 15) + 16 * (it >> 6);
      ao[q] = (it >> 4) & 3;
      alive[q] = (API * PQ_THREADS == AIT || it < AIT) && aj[q] < RB * 8;
      const int64_t pr = min((int64_t)(row0 / 2 + aj[q]), PP - 1);
      asrc[q] = g.pidx + (xb / PQ_KT * PP + pr) * PQ_KT + 8 * ao[q];
    }
    // column item q: complex column bn of the tile, sites 8 bo .. 8 bo + 7;
    // reads clamped to the last column, items past the tile store into junk.
    // The column's shift is subtracted from every site value as it is loaded
    // (sites past the split stay zero).
    int boff[BPI], bo[BPI];
    bool bin[BPI];
    const cx<float> *bsrc[BPI];
    cx<float> bsh[BPI];
    const int64_t bstep = 2 * (int64_t)g.ncols * 16;  // two site tiles
#pragma unroll
    for (int q = 0; q < BPI; q++) {
      const int it = ptid + PQ_THREADS * q;
      const int bn = (it & 15) + 16 * (it >> 6);
      bo[q] = (it >> 4) & 3;
      boff[q] = (BPI * PQ_THREADS == BIT || it < BIT) ? lds_q(bn, bo[q]) : -1;
      bin[q] = cx0 + bn < g.ncols;
      const int64_t col = min(cx0 + bn, (int64_t)g.ncols - 1);
      bsrc[q] = c + ctile_idx(col, xb + 8 * bo[q], g.ncols);
      bsh[q] = shift[col];
    }
    constexpr int NQ = PQ_INFLIGHT;
    double zre[BPI], zim[BPI];  // column sums of the staged steps (n = 0)
#pragma unroll
    for (int q = 0; q < BPI; q++) zre[q] = zim[q] = 0;
    cx<float> bvr[NQ][BPI][8];
    u4 avr[NQ][API];  // 8 phase indices (uint16) per item
    // full step k into register set SL: unconditional loads, so that the
    // compiler counts the loads in flight exactly
    auto load = [&](auto slot, int k) {
      constexpr int SL = decltype(slot)::value;
#pragma unroll
      for (int q = 0; q < API; q++) avr[SL][q] = *reinterpret_cast<const u4 *>(asrc[q] + k * astep);
#pragma unroll
      for (int q = 0; q < BPI; q++) {
        const f4 *src = reinterpret_cast<const f4 *>(bsrc[q] + k * bstep);
#pragma unroll
        for (int s = 0; s < 4; s++) {
          const f4 v = src[s];
          bvr[SL][q][2 * s] = cmake<float>(v[0] - bsh[q].re, v[1] - bsh[q].im);
          bvr[SL][q][2 * s + 1] = cmake<float>(v[2] - bsh[q].re, v[3] - bsh[q].im);
        }
      }
    };
    // the partial last step k (sites past the split zero) into register set 0
    auto load_tail = [&](int k) {
      const int64_t xt = xb + (int64_t)k * PQ_KT;
#pragma unroll
      for (int q = 0; q < API; q++) avr[0][q] = *reinterpret_cast<const u4 *>(asrc[q] + k * astep);
#pragma unroll
      for (int q = 0; q < BPI; q++) {
        const int64_t x0 = xt + 8 * bo[q];
#pragma unroll
        for (int s = 0; s < 8; s++) {
          cx<float> v = cmake<float>(0, 0);
          if (bin[q] && x0 + s < xe) {
            v = bsrc[q][k * bstep + s];
            v.re -= bsh[q].re;
            v.im -= bsh[q].im;
          }
          bvr[0][q][s] = v;
        }
      }
    };
    // register set SL -> tiles of one stage (column values split into bf16 pieces)
    auto stage = [&](auto slot, uint16_t *tA, uint16_t *tB) {
      constexpr int SL = decltype(slot)::value;
#pragma unroll
      for (int q = 0; q < BPI; q++) {
        if (zsum && boff[q] >= 0 && bin[q]) {
          float sr = 0, si = 0;
#pragma unroll
          for (int s = 0; s < 8; s++) {
            sr += bvr[SL][q][s].re;
            si += bvr[SL][q][s].im;
          }
          zre[q] += (double)sr;
          zim[q] += (double)si;
        }
        u4 wre[3], wim[3];
#pragma unroll
        for (int s = 0; s < 4; s++) {
          uint32_t a[3], b[3];
          split3(bvr[SL][q][2 * s].re, bvr[SL][q][2 * s + 1].re, a);
          split3(bvr[SL][q][2 * s].im, bvr[SL][q][2 * s + 1].im, b);
#pragma unroll
          for (int p = 0; p < 3; p++) {
            wre[p][s] = a[p];
            wim[p][s] = b[p];
          }
        }
#pragma unroll
        for (int p = 0; p < NP; p++) {
          *(boff[q] >= 0 ? reinterpret_cast<u4 *>(tB + p * BSZ + boff[q]) : &junk[lane]) = wre[p];
          *(boff[q] >= 0 ? reinterpret_cast<u4 *>(tB + p * BSZ + boff[q] + NCX * PQ_KT) : &junk[lane]) = wim[p];
        }
      }
      // phase items: table entries of 8 sites -> cos row and sin row, pieces
      // p, two sites per word (v_perm_b32 picks the 16-bit halves)
#pragma unroll
      for (int q = 0; q < API; q++) {
        u4 e[8];
#pragma unroll
        for (int t = 0; t < 4; t++) {
          e[2 * t] = sW[avr[SL][q][t] & 0xFFFFu];
          e[2 * t + 1] = sW[avr[SL][q][t] >> 16];
        }
        const int oc = lds_q(2 * aj[q], ao[q]), os = lds_q(2 * aj[q] + 1, ao[q]);
#pragma unroll
        for (int p = 0; p < NP; p++) {
          u4 wc, ws;
#pragma unroll
          for (int t = 0; t < 4; t++) {
            wc[t] = __builtin_amdgcn_perm(e[2 * t + 1][p], e[2 * t][p], 0x05040100u);
            ws[t] = __builtin_amdgcn_perm(e[2 * t + 1][p], e[2 * t][p], 0x07060302u);
          }
          *(alive[q] ? reinterpret_cast<u4 *>(tA + p * ASZ + oc) : &junk[lane]) = wc;
          *(alive[q] ? reinterpret_cast<u4 *>(tA + p * ASZ + os) : &junk[lane]) = ws;
        }
      }
    };
    using S0 = std::integral_constant<int, 0>;
    using S1 = std::integral_constant<int, 1>;
    using S2 = std::integral_constant<int, 2>;
    static_assert(NQ == 3, "three register sets");
    // step k sits in register set k % 3.  Consumers compute step k while
    // the producers stage step k + 1; barrier count: 1 + nk on both sides.
    const int nf = (int)((xe > xb ? xe - xb : 0) / PQ_KT);  // full steps
    const int last = nf > 0 ? nf - 1 : 0;
    if (nf > 0) {
      load(S0{}, 0);
      load(S1{}, min(1, last));
      load(S2{}, min(2, last));
      stage(S0{}, sA[0], sB[0]);
      load(S0{}, min(3, last));
    } else if (nk > 0) {
      load_tail(0);
      stage(S0{}, sA[0], sB[0]);
    }
    __syncthreads();
    // steady state: stage full steps k + 1 <= nf - 1 and refill their sets
    int k = 0;
    for (; k + 3 <= nf - 1; k += 3) {
      stage(S1{}, sA[(k + 1) & 1], sB[(k + 1) & 1]);
      load(S1{}, min(k + 4, last));
      __syncthreads();
      stage(S2{}, sA[(k + 2) & 1], sB[(k + 2) & 1]);
      load(S2{}, min(k + 5, last));
      __syncthreads();
      stage(S0{}, sA[(k + 3) & 1], sB[(k + 3) & 1]);
      load(S0{}, min(k + 6, last));
      __syncthreads();
    }
    // remaining full steps (their loads are in flight already)
    if (k + 1 <= nf - 1) {
      stage(S1{}, sA[(k + 1) & 1], sB[(k + 1) & 1]);
      __syncthreads();
      ++k;
      if (k + 1 <= nf - 1) {
        stage(S2{}, sA[(k + 1) & 1], sB[(k + 1) & 1]);
        __syncthreads();
        ++k;
      }
    }
    // the partial step after the full ones, then the consumers' last step
    if (nf > 0 && nk > nf) {
      load_tail(nf);
      stage(S0{}, sA[nf & 1], sB[nf & 1]);
      __syncthreads();
    }
    if (nk > 0) __syncthreads();
    if (zsum) {
      // the four octets of a column sit in lanes l, l ^ 16, l ^ 32, l ^ 48
      const int64_t nreal = 2 * (int64_t)g.ncols;
#pragma unroll
      for (int q = 0; q < BPI; q++) {
        double zr = zre[q], zi = zim[q];
        zr += __shfl_xor(zr, 16, 64);
        zi += __shfl_xor(zi, 16, 64);
        zr += __shfl_xor(zr, 32, 64);
        zi += __shfl_xor(zi, 32, 64);
        const int it = ptid + PQ_THREADS * q;
        const int64_t n = cx0 + (it & 15) + 16 * (it >> 6);
        if (bo[q] == 0 && boff[q] >= 0 && bin[q]) {
          double *pr = part + ((int64_t)split * nreal + 2 * n) * R + Rg;
          pr[0] = zr;      // Re column, cos row
          pr[1] = 0;       // Re column, sin row
          pr[R] = zi;      // Im column, cos row
          pr[R + 1] = 0;   // Im column, sin row
        }
      }
    }
    return;
  }

  // ---- consumers
  const int wr = wave / GC, wc = wave % GC;  // this wave's row and column group
  f4 acc[WR][WC];
  double acc64[WR][WC][4];
#pragma unroll
  for (int i = 0; i < WR; i++)
#pragma unroll
    for (int j = 0; j < WC; j++) {
      acc[i][j] = f4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; v++) acc64[i][j][v] = 0;
    }
  __syncthreads();
  int nacc = 0;
  for (int k = 0; k < nk; k++) {
    const uint16_t *tA = sA[k & 1], *tB = sB[k & 1];
    // MFMAs one A piece at a time (WR*WC independent accumulators between
    // two products into the same one)
    const int fq = lane >> 4;
    bf8 bf[WC][NP];
#pragma unroll
    for (int j = 0; j < WC; j++)
#pragma unroll
      for (int p = 0; p < NP; p++)
        bf[j][p] = *reinterpret_cast<const bf8 *>(tB + p * BSZ + lds_q((wc * WC + j) * 16 + (lane & 15), fq));
#pragma unroll
    for (int pa = NP - 1; pa >= 0; pa--) {
      bf8 af[WR];
#pragma unroll
      for (int i = 0; i < WR; i++)
        af[i] = *reinterpret_cast<const bf8 *>(tA + pa * ASZ + lds_q((wr * WR + i) * 16 + (lane & 15), fq));
      // B pieces that pair with A piece pa: pa + pb <= 2 (split 6), <= 1 (split 3)
#pragma unroll
      for (int pb = NP - 1 - pa; pb >= 0; pb--)
#pragma unroll
        for (int i = 0; i < WR; i++)
#pragma unroll
          for (int j = 0; j < WC; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bf[j][pb], acc[i][j], 0, 0, 0);
    }
    if (++nacc == PQ_FLUSH) {
      nacc = 0;
#pragma unroll
      for (int i = 0; i < WR; i++)
#pragma unroll
        for (int j = 0; j < WC; j++) {
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
  for (int i = 0; i < WR; i++)
#pragma unroll
    for (int j = 0; j < WC; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int r = row0 + (wr * WR + i) * 16 + 4 * (lane >> 4) + v;
        const int jt = (wc * WC + j) * 16 + (lane & 15);  // tile row of the column
        const int64_t n = 2 * (cx0 + (jt < NCX ? jt : jt - NCX)) + (jt < NCX ? 0 : 1);
        if (r < Rg && n < nreal) part[((int64_t)split * nreal + n) * R + r] = acc64[i][j][v] + (double)acc[i][j][v];
      }
}

// Tile codes 1000 (NCW / 4) + 100 WR + 10 WC + GR (1xxx: 4 consumer waves,
// 2xxx: 8): each consumer wave owns WR x WC blocks of 16 x 16 outputs, the
// consumer waves form GR x (NCW / GR) groups over rows x columns; a
// workgroup covers WR*GR row blocks and 16*WC*(NCW/GR) real columns.
#define PLG_PQ_CODES(X)                                                                                      \
  X(6, 1, 1, 4) X(5, 2, 1, 4) X(4, 2, 1, 4) X(3, 3, 1, 4) X(3, 3, 2, 4) X(4, 2, 2, 4) X(2, 4, 2, 4)           \
  X(3, 2, 2, 4) X(4, 3, 2, 4) X(3, 4, 2, 4) X(2, 2, 2, 4) X(2, 4, 4, 4) X(2, 6, 2, 4)                         \
  X(3, 2, 2, 8) X(4, 2, 2, 8) X(2, 2, 4, 8) X(3, 2, 4, 8) X(2, 3, 4, 8) X(8, 1, 2, 8) X(4, 2, 4, 8)      \
  X(4, 1, 4, 8) X(2, 2, 8, 8)

struct PqTile {
  int wr, wc, gr, ncw;
  int rb() const { return wr * gr; }
  int ncx() const { return 8 * wc * (ncw / gr); }
};
static PqTile pq_tile(int code) { return {(code / 100) % 10, (code / 10) % 10, code % 10, 4 * (code / 1000)}; }

int launch_project_bf16(const ProjGeom &g, const cx<float> *c, double *part, int terms, const uint32_t *wtab,
                        int code, hipStream_t st) {
  if (g.npairs <= 0 || !g.pidx) return PLG_ERR_INVALID;
  if (g.M > PQ_MMAX) return PLG_ERR_UNSUPPORTED;
  if (code < 1000 || code > 2999) return PLG_ERR_INVALID;
  const PqTile t = pq_tile(code);
  const int nblk = max(1, (2 * (g.npairs - g.zero_pair) + 15) / 16);
  const int nrg = (nblk + t.rb() - 1) / t.rb();
  const int64_t ncg = (g.ncols + t.ncx() - 1) / t.ncx();
  if ((ncg + 7) / 8 * 8 * nrg > INT32_MAX) return PLG_ERR_UNSUPPORTED;
  dim3 block((t.ncw + 4) * 64);
  dim3 grid((unsigned)((ncg + 7) / 8 * 8 * nrg), (unsigned)g.nsplit);
  if (!g.shifted) return PLG_ERR_INVALID;
  cx<float> *shift = proj_shift(g, part);
  k_col_shift<<<dim3((unsigned)((g.ncols + CS_COLS - 1) / CS_COLS)), dim3(256), 0, st>>>(c, g.ncols, g.V3, shift);
#define PLG_PQ(WR, WC, GR, NCW)                                                                            \
  if (t.wr == WR && t.wc == WC && t.gr == GR && t.ncw == NCW) {                                          \
    if (terms == 6) k_project_pair<WR, WC, GR, 6, NCW><<<grid, block, 0, st>>>(g, c, wtab, part, shift, nrg, (int)ncg); \
    else k_project_pair<WR, WC, GR, 3, NCW><<<grid, block, 0, st>>>(g, c, wtab, part, shift, nrg, (int)ncg);            \
    return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;                                       \
  }
  PLG_PQ_CODES(PLG_PQ)
#undef PLG_PQ
  return PLG_ERR_INVALID;
}

#define PLG_PQ_CODE(WR, WC, GR, NCW) 1000 * (NCW / 4) + 100 * WR + 10 * WC + GR,
static const int k_pq_codes[] = {PLG_PQ_CODES(PLG_PQ_CODE)};
#undef PLG_PQ_CODE

// tile codes the autotuner tries for a pair count: row tiles that waste at
// most one row block per row group (all when none qualifies)
std::vector<int> proj_candidates(int npairs) {
  const int nblk = (2 * npairs + 15) / 16;
  std::vector<int> v;
  for (int code : k_pq_codes) {
    const int rb = pq_tile(code).rb();
    const int groups = (nblk + rb - 1) / rb;
    if (groups * rb - nblk <= std::max(1, groups - 1)) v.push_back(code);
  }
  if (v.empty()) v.assign(std::begin(k_pq_codes), std::end(k_pq_codes));
  return v;
}

// workgroups of one K split for a tile code
int64_t proj_workgroups(int code, int64_t ncols, int npairs) {
  const PqTile t = pq_tile(code);
  const int nblk = (2 * npairs + 15) / 16;
  return ((ncols + t.ncx() - 1) / t.ncx()) * ((nblk + t.rb() - 1) / t.rb());
}

// default tile (the fastest on the measured workloads): 8 consumer waves of
// 4 x 2 blocks when 8 row blocks divide the rows, else 4 consumer waves of
// 3 x 4 blocks
int pick_proj(int npairs) {
  const int nblk = (2 * npairs + 15) / 16;
  if (nblk <= 2) return 1222;
  if (nblk <= 4) return 1242;
  return nblk % 8 == 0 ? 2422 : 1342;
}

// Pairs (n, -n) of a momentum list, on the scaled momenta ms (nmom x 3,
// n_k (M / L_k) mod M).  pair_ms: the representative of each pair;
// mom_pair[m] = 2 j + 1 when m is the -n member of pair j, else 2 j.  A
// momentum equal to its own negative (mod M: zero, or components L/2) has a
// zero sine row and stays alone, as does one whose partner is absent or
// already taken.
int make_pairs(const std::vector<int> &ms, int M, std::vector<int> &pair_ms, std::vector<int> &mom_pair,
               int *zero_last) {
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
  const int np = (int)pair_ms.size() / 3;
  // move a zero-momentum pair to the end
  *zero_last = 0;
  for (int j = 0; j < np; j++)
    if (pair_ms[3 * j] == 0 && pair_ms[3 * j + 1] == 0 && pair_ms[3 * j + 2] == 0) {
      const int l = np - 1;
      for (int i = 0; i < 3; i++) std::swap(pair_ms[3 * j + i], pair_ms[3 * l + i]);
      for (int &v : mom_pair) {
        if (v >> 1 == j) v = 2 * l + (v & 1);
        else if (v >> 1 == l) v = 2 * j + (v & 1);
      }
      *zero_last = 1;
      break;
    }
  return np;
}

}  // namespace plg
