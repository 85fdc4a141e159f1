// plg_project_bf16.hip -- mixed-precision momentum projection on MFMA.
//
// Same GEMM as the FP64 projection (interleaved re/im real form of
// sum_x exp(-i 2pi n.x/L) c[n][x]), with split-bf16 operands on
// v_mfma_f32_16x16x32_bf16:
//   a = a0 + a1 + a2, b = b0 + b1 + b2 (bf16 pieces, ~2^-24 relative)
//   3 products  a0b0 + a0b1 + a1b0               (split = 3)
//   6 products  + a0b2 + a1b1 + a2b0             (split = 6)
// smallest products first into an FP32 MFMA accumulator that is added into
// FP64 registers every FLUSH k-steps; FP64 partials per K split.
//
// Work decomposition: a workgroup owns 32 columns (two 16-column blocks) and
// 4*MBW row blocks of 16 real rows (2 momenta x re/im each) for one K range.
// Wave w owns row blocks w, w+4, ...  Each site-correlator element is read
// and split into pieces once per row group (once in total when one group
// covers all momenta).  The A operand is never stored: every lane builds its
// fragments from the packed phase table, whose 32-bit words already hold the
// (re, im) bf16 pair of a piece:
//   W[k][p]   = cos_p | sin_p << 16       (row re:  cos,  sin)
//   W[k][3+p] = -sin_p | cos_p << 16      (row im: -sin,  cos)
// for theta = 2 pi k / M, k = n.x (M/L) mod M.
#include "plg_internal.h"

namespace plg {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef uint32_t u4 __attribute__((ext_vector_type(4)));

constexpr int PB_THREADS = 256;
constexpr int PB_NC = 32;            // columns per workgroup
constexpr int PB_KT = 16;            // sites per k-step (32 real K)
constexpr int PB_LDB = 2 * PB_KT + 8;  // padded bf16 row of the B stage
constexpr int PB_MMAX = 512;         // phase-table entries held in LDS
constexpr int PB_FLUSH = 8;


// k mod M for 0 <= k < 2^24 with a float reciprocal (exact quotient +-1, fixed up)
__device__ __forceinline__ int mod_m(int k, int M, float invM) {
  int q = (int)((float)k * invM);
  int r = k - q * M;
  r += r < 0 ? M : 0;
  r -= r >= M ? M : 0;
  return r;
}

// (re, im) -> three packed bf16 pairs p0 + p1 + p2, each rounded to nearest
// even by v_cvt_pk_bf16_f32 (one instruction per pair)
typedef float f2v __attribute__((ext_vector_type(2)));
typedef __bf16 b2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void split3(cx<float> v, uint32_t *w) {
  const f2v x = {v.re, v.im};
  const b2v h0 = __builtin_convertvector(x, b2v);
  const f2v r1 = x - __builtin_convertvector(h0, f2v);
  const b2v h1 = __builtin_convertvector(r1, b2v);
  const f2v r2 = r1 - __builtin_convertvector(h1, f2v);
  const b2v h2 = __builtin_convertvector(r2, b2v);
  w[0] = __builtin_bit_cast(uint32_t, h0);
  w[1] = __builtin_bit_cast(uint32_t, h1);
  w[2] = __builtin_bit_cast(uint32_t, h2);
}

template <int MBW, int TERMS>
__global__ __launch_bounds__(PB_THREADS) void k_project_bf16(ProjGeom g, const cx<float> *__restrict__ c,
                                                             const uint32_t *__restrict__ wtab,
                                                             double *__restrict__ part) {
  constexpr int NP = TERMS == 3 ? 2 : 3;
  __shared__ uint32_t sW[PB_MMAX * 6];
  __shared__ __align__(16) uint16_t sB[2][NP][PB_NC][PB_LDB];
  __shared__ int sXYZ[2][PB_KT];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t n0 = (int64_t)blockIdx.x * PB_NC;
  const int split = blockIdx.y;
  const int mgroup = blockIdx.z;
  const int Mreal = 2 * g.nmom;
  const int nblk = (Mreal + 15) >> 4;
  const int M = g.M;
  const float invM = 1.0f / (float)M;
  int64_t xb, xe;
  {
    const int64_t per = (g.V3 + g.nsplit - 1) / g.nsplit;
    xb = (int64_t)split * per;
    xe = xb + per < g.V3 ? xb + per : g.V3;
  }
  for (int i = tid; i < M * 6; i += PB_THREADS) sW[i] = wtab[i];

  // per row block: packed momentum (n_x, n_y, n_z scaled, 10 bits each),
  // bit 30 = imaginary row, bit 31 = row exists
  int rowinfo[MBW];
#pragma unroll
  for (int i = 0; i < MBW; i++) {
    const int mb = mgroup * 4 * MBW + wave + 4 * i;
    const int mr = mb * 16 + (lane & 15);
    int info = 0;
    if (mb < nblk && mr < Mreal) {
      const int m = mr >> 1;
      info = g.mom_scaled[3 * m] | (g.mom_scaled[3 * m + 1] << 10) | (g.mom_scaled[3 * m + 2] << 20) |
             ((mr & 1) << 30) | (int)(1u << 31);
    }
    rowinfo[i] = info;
  }

  f4 acc[MBW][2];
  double acc64[MBW][2][4];
#pragma unroll
  for (int i = 0; i < MBW; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      acc[i][j] = f4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; v++) acc64[i][j][v] = 0;
    }

  // staging of one k-step, in two halves so that the global loads of step
  // t+1 are in flight during the MFMAs of step t:
  //   load():  thread -> column tid/8, sites 2*(tid%8) + {0,1} into registers
  //   store(): split into bf16 pieces, write the LDS stage, site coordinates
  const int scol = tid >> 3, s0 = 2 * (tid & 7);
  float v[4];
  auto load = [&](int64_t xt) {
    const int64_t n = n0 + scol;
#pragma unroll
    for (int e = 0; e < 4; e++) v[e] = 0;
    if (n < g.ncols) {
#pragma unroll
      for (int q = 0; q < 2; q++) {
        const int64_t x = xt + s0 + q;
        if (x < xe) {
          const cx<float> z = c[cidx(g, n, x)];
          v[2 * q] = z.re;
          v[2 * q + 1] = z.im;
        }
      }
    }
  };
  auto store = [&](int buf, int64_t xt) {
    const int col = scol;
    uint32_t w0[3], w1[3];
    split3(cmake<float>(v[0], v[1]), w0);
    split3(cmake<float>(v[2], v[3]), w1);
#pragma unroll
    for (int p = 0; p < NP; p++) {
      uint2 w;
      w.x = w0[p];
      w.y = w1[p];
      *reinterpret_cast<uint2 *>(&sB[buf][p][col][2 * s0]) = w;
    }
    if (tid < PB_KT) {
      const int64_t x = xt + tid;
      int packed = -1;  // site outside the range
      if (x < xe) {
        const int cx_ = (int)(x % g.L[0]);
        const int64_t q = x / g.L[0];
        packed = cx_ | ((int)(q % g.L[1]) << 10) | ((int)(q / g.L[1]) << 20);
      }
      sXYZ[buf][tid] = packed;
    }
  };

  int buf = 0, step = 0;
  load(xb);
  store(0, xb);
  __syncthreads();
  for (int64_t xt = xb; xt < xe; xt += PB_KT) {
    const bool more = xt + PB_KT < xe;
    if (more) load(xt + PB_KT);
    // B fragments: lane l holds K = 8*(l/16) .. +7 of column l%16
    const int kk = 8 * (lane >> 4);
    bf8 bf[NP][2];
#pragma unroll
    for (int p = 0; p < NP; p++)
#pragma unroll
      for (int j = 0; j < 2; j++)
        bf[p][j] = *reinterpret_cast<const bf8 *>(&sB[buf][p][j * 16 + (lane & 15)][kk]);
    int xyz[4];
#pragma unroll
    for (int q = 0; q < 4; q++) xyz[q] = sXYZ[buf][4 * (lane >> 4) + q];
#pragma unroll
    for (int i = 0; i < MBW; i++) {
      // skip whole padding blocks only: the test is uniform across the wave
      if (mgroup * 4 * MBW + wave + 4 * i >= nblk) continue;
      const int info = rowinfo[i];
      const int mx = info & 1023, my = (info >> 10) & 1023, mz = (info >> 20) & 1023;
      const int hi = (info >> 30) & 1;
      u4 w[NP];
#pragma unroll
      for (int q = 0; q < 4; q++) {
        const int s = xyz[q];
        uint32_t wp[3] = {0, 0, 0};
        if (info < 0 && s >= 0) {
          const int k = mod_m(mx * (s & 1023) + my * ((s >> 10) & 1023) + mz * (s >> 20), M, invM);
          const uint32_t *t = &sW[k * 6 + 3 * hi];
#pragma unroll
          for (int p = 0; p < NP; p++) wp[p] = t[p];
        }
#pragma unroll
        for (int p = 0; p < NP; p++) w[p][q] = wp[p];
      }
      bf8 af[NP];
#pragma unroll
      for (int p = 0; p < NP; p++) af[p] = __builtin_bit_cast(bf8, w[p]);
#pragma unroll
      for (int j = 0; j < 2; j++) {
        f4 a = acc[i][j];
        if (TERMS == 6) {
          a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[2], bf[0][j], a, 0, 0, 0);
          a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1], bf[1][j], a, 0, 0, 0);
          a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bf[2][j], a, 0, 0, 0);
        }
        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1], bf[0][j], a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bf[1][j], a, 0, 0, 0);
        a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bf[0][j], a, 0, 0, 0);
        acc[i][j] = a;
      }
    }
    if (more) store(buf ^ 1, xt + PB_KT);
    if (++step == PB_FLUSH) {
      step = 0;
#pragma unroll
      for (int i = 0; i < MBW; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
#pragma unroll
          for (int v = 0; v < 4; v++) acc64[i][j][v] += (double)acc[i][j][v];
          acc[i][j] = f4{0, 0, 0, 0};
        }
    }
    __syncthreads();
    buf ^= 1;
  }
  // D of 16x16x32 (f32): lane l, value v -> row 4*(l/16) + v, column l%16
#pragma unroll
  for (int i = 0; i < MBW; i++) {
    const int mb = mgroup * 4 * MBW + wave + 4 * i;
    if (mb >= nblk) continue;
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int mr = mb * 16 + 4 * (lane >> 4) + v;
        const int64_t n = n0 + j * 16 + (lane & 15);
        if (mr < Mreal && n < g.ncols)
          part[((int64_t)split * g.ncols + n) * Mreal + mr] = acc64[i][j][v] + (double)acc[i][j][v];
      }
  }
}

// ------------------------------------------------------------------ shared-A variant
// A workgroup owns RB row blocks (one row group) and 4*NBW column blocks; the
// phase tile of each k-step (RB*16 rows x 16 sites, packed operand words per
// piece) is built once in LDS by all 256 threads and read by the four waves,
// which split the columns.  Two barriers per k-step; the site-correlator
// loads of step t+1 are issued before the MFMAs of step t.
constexpr int PV_AST = 20;  // words per A-tile row (16 sites, padded for alignment)

template <int RB, int NBW, int TERMS>
__global__ __launch_bounds__(PB_THREADS, 2) void k_project_bf16_sa(ProjGeom g, const cx<float> *__restrict__ c,
                                                                   const uint32_t *__restrict__ wtab,
                                                                   double *__restrict__ part) {
  constexpr int NP = TERMS == 3 ? 2 : 3;
  constexpr int NC = 4 * NBW * 16;            // columns per workgroup
  constexpr int RM = RB * 8;                  // momenta per row group
  constexpr int MPT = (RM + 15) / 16;         // momentum slots per thread in the tile build
  constexpr int CPT = NC * PB_KT / PB_THREADS;  // complex B values staged per thread
  // phase table, 8 words per k: (re-row, im-row) word pairs of pieces 0..2
  __shared__ __align__(16) uint32_t sWt[PB_MMAX * 8];
  __shared__ __align__(16) uint32_t sA[NP][MPT * 32][PV_AST];
  __shared__ __align__(16) uint16_t sB[NP][NC][PB_LDB];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t n0 = (int64_t)blockIdx.x * NC;
  const int split = blockIdx.y;
  const int row0 = blockIdx.z * RB * 16;
  const int Mreal = 2 * g.nmom;
  const int M = g.M;
  const float invM = 1.0f / (float)M;
  int64_t xb, xe;
  {
    const int64_t per = (g.V3 + g.nsplit - 1) / g.nsplit;
    xb = (int64_t)split * per;
    xe = xb + per < g.V3 ? xb + per : g.V3;
  }
  for (int i = tid; i < M; i += PB_THREADS) {
#pragma unroll
    for (int p = 0; p < 3; p++) {
      sWt[8 * i + 2 * p] = wtab[6 * i + p];
      sWt[8 * i + 2 * p + 1] = wtab[6 * i + 3 + p];
    }
    sWt[8 * i + 6] = 0;
    sWt[8 * i + 7] = 0;
  }

  // tile build assignment: site tid%16, momenta (tid/16) + 16 j of this group
  const int ts = tid & 15;
  int mpk[MPT];
#pragma unroll
  for (int j = 0; j < MPT; j++) {
    const int mi = (tid >> 4) + 16 * j;
    const int m = row0 / 2 + mi;
    mpk[j] = (mi < RM && m < g.nmom)
                 ? (g.mom_scaled[3 * m] | (g.mom_scaled[3 * m + 1] << 10) | (g.mom_scaled[3 * m + 2] << 20))
                 : -1;
  }
  // staging assignment: column tid / (16/CPT) ... CPT consecutive sites of one column
  constexpr int TPC = PB_KT / CPT;            // threads per column
  const int bcol = tid / TPC, bs0 = (tid % TPC) * CPT;

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

  cx<float> bv[CPT];
  auto load = [&](int64_t xt) {
    const int64_t n = n0 + bcol;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const int64_t x = xt + bs0 + q;
      bv[q] = (n < g.ncols && x < xe) ? c[cidx(g, n, x)] : cmake<float>(0, 0);
    }
  };

  int step = 0;
  load(xb);
  // lattice coordinates of this thread's site (xb + ts), advanced by PB_KT per step
  int cx_, cy_, cz_;
  {
    const int x = (int)(xb + ts);
    cx_ = x % g.L[0];
    const int q = x / g.L[0];
    cy_ = q % g.L[1];
    cz_ = q / g.L[1];
  }
  __syncthreads();  // sWt ready
  for (int64_t xt = xb; xt < xe; xt += PB_KT) {
    // ---- stage B pieces of this step
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      uint32_t w[3];
      split3(bv[q], w);
#pragma unroll
      for (int p = 0; p < NP; p++) *reinterpret_cast<uint32_t *>(&sB[p][bcol][2 * (bs0 + q)]) = w[p];
    }
    // ---- build the phase tile of this step (rows beyond the group are
    // padding rows of sA that no MFMA reads)
    {
      const bool sv = xt + ts < xe;
#pragma unroll
      for (int j = 0; j < MPT; j++) {
        const int mi = (tid >> 4) + 16 * j;
        const int pk = mpk[j];
        const bool ok = pk >= 0 && sv;
        const int k = ok ? mod_m((pk & 1023) * cx_ + ((pk >> 10) & 1023) * cy_ + (pk >> 20) * cz_, M, invM) : 0;
        const uint2 *e = reinterpret_cast<const uint2 *>(&sWt[8 * k]);
#pragma unroll
        for (int p = 0; p < NP; p++) {
          const uint2 w = e[p];
          sA[p][2 * mi][ts] = ok ? w.x : 0u;
          sA[p][2 * mi + 1][ts] = ok ? w.y : 0u;
        }
      }
      cx_ += PB_KT;
      while (cx_ >= g.L[0]) {
        cx_ -= g.L[0];
        if (++cy_ == g.L[1]) {
          cy_ = 0;
          ++cz_;
        }
      }
    }
    const bool more = xt + PB_KT < xe;
    if (more) load(xt + PB_KT);
    __syncthreads();
    // ---- MFMAs
    {
      const int kk = 8 * (lane >> 4);
      bf8 bf[NP][NBW];
#pragma unroll
      for (int p = 0; p < NP; p++)
#pragma unroll
        for (int j = 0; j < NBW; j++)
          bf[p][j] = *reinterpret_cast<const bf8 *>(&sB[p][(wave * NBW + j) * 16 + (lane & 15)][kk]);
#pragma unroll
      for (int i = 0; i < RB; i++) {
        if (row0 + i * 16 >= Mreal) continue;  // uniform: a padding block
        bf8 af[NP];
#pragma unroll
        for (int p = 0; p < NP; p++)
          af[p] = __builtin_bit_cast(bf8, *reinterpret_cast<const u4 *>(&sA[p][i * 16 + (lane & 15)][4 * (lane >> 4)]));
#pragma unroll
        for (int j = 0; j < NBW; j++) {
          f4 a = acc[i][j];
          if (TERMS == 6) {
            a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[2], bf[0][j], a, 0, 0, 0);
            a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1], bf[1][j], a, 0, 0, 0);
            a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bf[2][j], a, 0, 0, 0);
          }
          a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[1], bf[0][j], a, 0, 0, 0);
          a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bf[1][j], a, 0, 0, 0);
          a = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[0], bf[0][j], a, 0, 0, 0);
          acc[i][j] = a;
        }
      }
    }
    if (++step == PB_FLUSH) {
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
#pragma unroll
  for (int i = 0; i < RB; i++)
#pragma unroll
    for (int j = 0; j < NBW; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int mr = row0 + i * 16 + 4 * (lane >> 4) + v;
        const int64_t n = n0 + (wave * NBW + j) * 16 + (lane & 15);
        if (mr < Mreal && n < g.ncols)
          part[((int64_t)split * g.ncols + n) * Mreal + mr] = acc64[i][j][v] + (double)acc[i][j][v];
      }
}


// ------------------------------------------------------------------ 32x32 variant
// v_mfma_f32_32x32x16_bf16 sustains close to the BF16 peak with only a couple
// of independent accumulators, where 16x16x32 needs many more than the
// FP64-shadowed accumulators leave registers for.  256 threads (4 waves, one
// per SIMD, up to 512 registers each); a workgroup owns RT row tiles of 32
// real rows and 128 columns, each wave one 32-column tile and all RT row
// tiles.  Operand layouts (lane l, t < 8):
//   A: A[l % 32][8 (l / 32) + t]     B: B[8 (l / 32) + t][l % 32]
//   D (16 values v): row 8 (v / 4) + 4 (l / 32) + v % 4, column l % 32
// A k-step covers 16 sites = two MFMA K-slices of 8 sites.  Phase and
// site-correlator tiles are double-buffered; one barrier per step.
constexpr int M32_THREADS = 256;
constexpr int M32_NC = 128;

template <int RT, int TERMS>
__global__ __launch_bounds__(M32_THREADS) __attribute__((amdgpu_waves_per_eu(1, 1)))
void k_project_bf16_m32(ProjGeom g, const cx<float> *__restrict__ c, const uint32_t *__restrict__ wtab,
                        double *__restrict__ part) {
  typedef float f16v __attribute__((ext_vector_type(16)));
  constexpr int NP = TERMS == 3 ? 2 : 3;
  constexpr int NC = M32_NC;
  constexpr int RM = RT * 16;                      // momenta per row group
  constexpr int MPT = RM * PB_KT / M32_THREADS;    // phase entries per thread per step (= RT)
  constexpr int CPT = NC * PB_KT / M32_THREADS;    // complex B values staged per thread (8)
  constexpr int TPC = PB_KT / CPT;
  __shared__ __align__(16) uint32_t sA[2][NP][RT * 32][PV_AST];
  __shared__ __align__(16) uint16_t sB[2][NP][NC][PB_LDB];
  extern __shared__ __align__(16) uint32_t sWt[];  // M * 8 words

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int64_t n0 = (int64_t)blockIdx.x * NC;
  const int split = blockIdx.y;
  const int row0 = blockIdx.z * RT * 32;
  const int Mreal = 2 * g.nmom;
  const int M = g.M;
  const float invM = 1.0f / (float)M;
  int64_t xb, xe;
  {
    const int64_t per = (g.V3 + g.nsplit - 1) / g.nsplit;
    xb = (int64_t)split * per;
    xe = xb + per < g.V3 ? xb + per : g.V3;
  }
  for (int i = tid; i < M; i += M32_THREADS) {
#pragma unroll
    for (int p = 0; p < 3; p++) {
      sWt[8 * i + 2 * p] = wtab[6 * i + p];
      sWt[8 * i + 2 * p + 1] = wtab[6 * i + 3 + p];
    }
    sWt[8 * i + 6] = 0;
    sWt[8 * i + 7] = 0;
  }
  // tile build: site tid % 16, momenta (tid / 16) + 16 j of this group
  const int ts = tid & 15;
  int mpk[MPT];
#pragma unroll
  for (int j = 0; j < MPT; j++) {
    const int mi = (tid >> 4) + 16 * j;
    const int m = row0 / 2 + mi;
    mpk[j] = (m < g.nmom)
                 ? (g.mom_scaled[3 * m] | (g.mom_scaled[3 * m + 1] << 10) | (g.mom_scaled[3 * m + 2] << 20))
                 : -1;
  }
  const int bcol = tid / TPC, bs0 = (tid % TPC) * CPT;

  f16v acc[RT];
  double acc64[RT][16];
#pragma unroll
  for (int i = 0; i < RT; i++)
#pragma unroll
    for (int v = 0; v < 16; v++) {
      acc[i][v] = 0;
      acc64[i][v] = 0;
    }

  cx<float> bv[CPT];
  auto load = [&](int64_t xt) {
    const int64_t n = n0 + bcol;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const int64_t x = xt + bs0 + q;
      bv[q] = (n < g.ncols && x < xe) ? c[cidx(g, n, x)] : cmake<float>(0, 0);
    }
  };
  int cx_, cy_, cz_;
  {
    const int x = (int)(xb + ts);
    cx_ = x % g.L[0];
    const int q = x / g.L[0];
    cy_ = q % g.L[1];
    cz_ = q / g.L[1];
  }
  auto build = [&](int b, int64_t xt) {
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      uint32_t w[3];
      split3(bv[q], w);
#pragma unroll
      for (int p = 0; p < NP; p++) *reinterpret_cast<uint32_t *>(&sB[b][p][bcol][2 * (bs0 + q)]) = w[p];
    }
    const bool sv = xt + ts < xe;
#pragma unroll
    for (int j = 0; j < MPT; j++) {
      const int mi = (tid >> 4) + 16 * j;
      const int pk = mpk[j];
      const bool ok = pk >= 0 && sv;
      const int k = ok ? mod_m((pk & 1023) * cx_ + ((pk >> 10) & 1023) * cy_ + (pk >> 20) * cz_, M, invM) : 0;
      const uint2 *e = reinterpret_cast<const uint2 *>(&sWt[8 * k]);
#pragma unroll
      for (int p = 0; p < NP; p++) {
        const uint2 w = e[p];
        sA[b][p][2 * mi][ts] = ok ? w.x : 0u;
        sA[b][p][2 * mi + 1][ts] = ok ? w.y : 0u;
      }
    }
    cx_ += PB_KT;
    while (cx_ >= g.L[0]) {
      cx_ -= g.L[0];
      if (++cy_ == g.L[1]) {
        cy_ = 0;
        ++cz_;
      }
    }
  };

  int step = 0, buf = 0;
  __syncthreads();  // phase table
  if (xb < xe) {
    load(xb);
    build(0, xb);
    if (xb + PB_KT < xe) load(xb + PB_KT);
  }
  __syncthreads();
  const int h = lane >> 5, r32 = lane & 31;
  for (int64_t xt = xb; xt < xe; xt += PB_KT) {
    // ---- MFMAs of tile t: two K slices of 8 sites
#pragma unroll
    for (int s = 0; s < 2; s++) {
      bf8 bf[NP];
#pragma unroll
      for (int p = 0; p < NP; p++)
        bf[p] = *reinterpret_cast<const bf8 *>(&sB[buf][p][wave * 32 + r32][16 * s + 8 * h]);
      bf8 af[RT][NP];
#pragma unroll
      for (int i = 0; i < RT; i++)
#pragma unroll
        for (int p = 0; p < NP; p++)
          af[i][p] = __builtin_bit_cast(bf8, *reinterpret_cast<const u4 *>(&sA[buf][p][i * 32 + r32][8 * s + 4 * h]));
      auto term = [&](int pa, int pb) {
#pragma unroll
        for (int i = 0; i < RT; i++)
          acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[i][pa], bf[pb], acc[i], 0, 0, 0);
      };
      if constexpr (TERMS == 6) {
        term(2, 0);
        term(1, 1);
        term(0, 2);
      }
      term(1, 0);
      term(0, 1);
      term(0, 0);
    }
    // ---- tile t+1 into the other buffers while the MFMAs run
    if (xt + PB_KT < xe) {
      build(buf ^ 1, xt + PB_KT);
      if (xt + 2 * PB_KT < xe) load(xt + 2 * PB_KT);
    }
    if (++step == PB_FLUSH) {
      step = 0;
#pragma unroll
      for (int i = 0; i < RT; i++)
#pragma unroll
        for (int v = 0; v < 16; v++) {
          acc64[i][v] += (double)acc[i][v];
          acc[i][v] = 0;
        }
    }
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int i = 0; i < RT; i++)
#pragma unroll
    for (int v = 0; v < 16; v++) {
      const int mr = row0 + i * 32 + 8 * (v >> 2) + 4 * h + (v & 3);
      const int64_t n = n0 + wave * 32 + r32;
      if (mr < Mreal && n < g.ncols)
        part[((int64_t)split * g.ncols + n) * Mreal + mr] = acc64[i][v] + (double)acc[i][v];
    }
}

// LDS bytes of a 32x32 workgroup (static tiles + phase table)
static size_t m32_lds(int rt, int np, int M) {
  return (size_t)2 * np * rt * 32 * PV_AST * 4 + (size_t)2 * np * M32_NC * PB_LDB * 2 + (size_t)M * 32;
}

// ------------------------------------------------------------------ pipelined shared-A variant
// 512 threads (8 waves) share the phase tile of RB row blocks and a B tile of
// 128 columns.  Waves form a 2 x 4 grid: each owns RB/2 row blocks and two
// column blocks, so every A fragment read from LDS feeds 2 column blocks and
// every B fragment RB/2 row blocks.  Phase and site-correlator tiles are
// double-buffered: in step t a wave issues the MFMAs of tile t and then builds
// its share of tile t+1 while they run; one barrier per step.  The phase table
// (M entries) is dynamic LDS.
constexpr int PP_THREADS = 512;

template <int RB, int TERMS>
__global__ __launch_bounds__(PP_THREADS) void k_project_bf16_pp(ProjGeom g, const cx<float> *__restrict__ c,
                                                                const uint32_t *__restrict__ wtab,
                                                                double *__restrict__ part) {
  static_assert(RB % 2 == 0, "two row halves");
  constexpr int NP = TERMS == 3 ? 2 : 3;
  constexpr int RH = RB / 2;                      // row blocks per wave
  constexpr int NC = 128;                         // columns per workgroup
  constexpr int RM = RB * 8;                      // momenta per row group
  constexpr int MPT = (RM + 31) / 32;             // momentum slots per thread in the tile build
  constexpr int CPT = NC * PB_KT / PP_THREADS;    // complex B values staged per thread
  constexpr int TPC = PB_KT / CPT;
  __shared__ __align__(16) uint32_t sA[2][NP][RB * 16][PV_AST];
  __shared__ __align__(16) uint16_t sB[2][NP][NC][PB_LDB];
  extern __shared__ __align__(16) uint32_t sWt[];  // M * 8 words

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 2, wc = wave & 3;        // row half, column pair of this wave
  const int64_t n0 = (int64_t)blockIdx.x * NC;
  const int split = blockIdx.y;
  const int row0 = blockIdx.z * RB * 16;
  const int Mreal = 2 * g.nmom;
  const int M = g.M;
  const float invM = 1.0f / (float)M;
  int64_t xb, xe;
  {
    const int64_t per = (g.V3 + g.nsplit - 1) / g.nsplit;
    xb = (int64_t)split * per;
    xe = xb + per < g.V3 ? xb + per : g.V3;
  }
  for (int i = tid; i < M; i += PP_THREADS) {
#pragma unroll
    for (int p = 0; p < 3; p++) {
      sWt[8 * i + 2 * p] = wtab[6 * i + p];
      sWt[8 * i + 2 * p + 1] = wtab[6 * i + 3 + p];
    }
    sWt[8 * i + 6] = 0;
    sWt[8 * i + 7] = 0;
  }
  // tile build: site tid % 16, momenta (tid / 16) + 32 j of this group
  const int ts = tid & 15;
  int mpk[MPT];
#pragma unroll
  for (int j = 0; j < MPT; j++) {
    const int mi = (tid >> 4) + 32 * j;
    const int m = row0 / 2 + mi;
    mpk[j] = (mi < RM && m < g.nmom)
                 ? (g.mom_scaled[3 * m] | (g.mom_scaled[3 * m + 1] << 10) | (g.mom_scaled[3 * m + 2] << 20))
                 : -1;
  }
  const int bcol = tid / TPC, bs0 = (tid % TPC) * CPT;

  f4 acc[RH][2];
  double acc64[RH][2][4];
#pragma unroll
  for (int i = 0; i < RH; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
      acc[i][j] = f4{0, 0, 0, 0};
#pragma unroll
      for (int v = 0; v < 4; v++) acc64[i][j][v] = 0;
    }

  cx<float> bv[CPT];
  auto load = [&](int64_t xt) {
    const int64_t n = n0 + bcol;
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      const int64_t x = xt + bs0 + q;
      bv[q] = (n < g.ncols && x < xe) ? c[cidx(g, n, x)] : cmake<float>(0, 0);
    }
  };
  int cx_, cy_, cz_;
  {
    const int x = (int)(xb + ts);
    cx_ = x % g.L[0];
    const int q = x / g.L[0];
    cy_ = q % g.L[1];
    cz_ = q / g.L[1];
  }
  // tile of the step starting at xt (coordinates: this thread's site of it)
  auto build = [&](int b, int64_t xt) {
#pragma unroll
    for (int q = 0; q < CPT; q++) {
      uint32_t w[3];
      split3(bv[q], w);
#pragma unroll
      for (int p = 0; p < NP; p++) *reinterpret_cast<uint32_t *>(&sB[b][p][bcol][2 * (bs0 + q)]) = w[p];
    }
    const bool sv = xt + ts < xe;
#pragma unroll
    for (int j = 0; j < MPT; j++) {
      const int mi = (tid >> 4) + 32 * j;
      if (32 * (j + 1) <= RM || mi < RM) {  // only the last slot can fall outside the group
        const int pk = mpk[j];
        const bool ok = pk >= 0 && sv;
        const int k = ok ? mod_m((pk & 1023) * cx_ + ((pk >> 10) & 1023) * cy_ + (pk >> 20) * cz_, M, invM) : 0;
        const uint2 *e = reinterpret_cast<const uint2 *>(&sWt[8 * k]);
#pragma unroll
        for (int p = 0; p < NP; p++) {
          const uint2 w = e[p];
          sA[b][p][2 * mi][ts] = ok ? w.x : 0u;
          sA[b][p][2 * mi + 1][ts] = ok ? w.y : 0u;
        }
      }
    }
    cx_ += PB_KT;
    while (cx_ >= g.L[0]) {
      cx_ -= g.L[0];
      if (++cy_ == g.L[1]) {
        cy_ = 0;
        ++cz_;
      }
    }
  };

  int step = 0, buf = 0;
  __syncthreads();  // phase table
  if (xb < xe) {
    load(xb);
    build(0, xb);
    if (xb + PB_KT < xe) load(xb + PB_KT);
  }
  __syncthreads();
  const int kk = 8 * (lane >> 4);
  for (int64_t xt = xb; xt < xe; xt += PB_KT) {
    // ---- MFMAs of tile t: rows wr*RH + i, columns (2 wc + j) * 16
    {
      bf8 bf[2][NP];
#pragma unroll
      for (int j = 0; j < 2; j++)
#pragma unroll
        for (int p = 0; p < NP; p++)
          bf[j][p] = *reinterpret_cast<const bf8 *>(&sB[buf][p][(2 * wc + j) * 16 + (lane & 15)][kk]);
      bf8 af[RH][NP];
#pragma unroll
      for (int i = 0; i < RH; i++)
#pragma unroll
        for (int p = 0; p < NP; p++)
          af[i][p] = __builtin_bit_cast(
              bf8, *reinterpret_cast<const u4 *>(&sA[buf][p][(wr * RH + i) * 16 + (lane & 15)][4 * (lane >> 4)]));
      // smallest products first, term-major over the tiles
      auto term = [&](int pa, int pb) {
#pragma unroll
        for (int i = 0; i < RH; i++)
#pragma unroll
          for (int j = 0; j < 2; j++)
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i][pa], bf[j][pb], acc[i][j], 0, 0, 0);
      };
      if constexpr (TERMS == 6) {
        term(2, 0);
        term(1, 1);
        term(0, 2);
      }
      term(1, 0);
      term(0, 1);
      term(0, 0);
    }
    // ---- tile t+1 into the other buffers while the MFMAs run
    if (xt + PB_KT < xe) {
      build(buf ^ 1, xt + PB_KT);
      if (xt + 2 * PB_KT < xe) load(xt + 2 * PB_KT);
    }
    if (++step == PB_FLUSH) {
      step = 0;
#pragma unroll
      for (int i = 0; i < RH; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
#pragma unroll
          for (int v = 0; v < 4; v++) acc64[i][j][v] += (double)acc[i][j][v];
          acc[i][j] = f4{0, 0, 0, 0};
        }
    }
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int i = 0; i < RH; i++)
#pragma unroll
    for (int j = 0; j < 2; j++)
#pragma unroll
      for (int v = 0; v < 4; v++) {
        const int mr = row0 + (wr * RH + i) * 16 + 4 * (lane >> 4) + v;
        const int64_t n = n0 + (2 * wc + j) * 16 + (lane & 15);
        if (mr < Mreal && n < g.ncols)
          part[((int64_t)split * g.ncols + n) * Mreal + mr] = acc64[i][j][v] + (double)acc[i][j][v];
      }
}

// row blocks per wave for a momentum count: fewest padded blocks, then fewest groups
int pick_mbw(int nmom) {
  const int nblk = (2 * nmom + 15) / 16;
  int best = 1, bestw = 1 << 30;
  for (int mbw : {1, 2, 3, 5, 9}) {
    const int per = 4 * mbw;
    const int waste = ((nblk + per - 1) / per) * per - nblk;
    if (waste <= bestw) {  // ties go to the larger tile (fewer row groups)
      best = mbw;
      bestw = waste;
    }
  }
  return best;
}

int launch_project_bf16(const ProjGeom &g, const cx<float> *c, cx<double> *part, int terms, const uint32_t *wtab,
                        int mbw, hipStream_t st) {
  if (g.M > PB_MMAX) return PLG_ERR_UNSUPPORTED;
  const int nblk = (2 * g.nmom + 15) / 16;
  const int groups = (nblk + 4 * mbw - 1) / (4 * mbw);
  dim3 block(PB_THREADS), grid((unsigned)((g.ncols + PB_NC - 1) / PB_NC), (unsigned)g.nsplit, (unsigned)groups);
  double *pd = reinterpret_cast<double *>(part);
#define PLG_PB(MB)                                                                 \
  case MB:                                                                         \
    if (terms == 6) k_project_bf16<MB, 6><<<grid, block, 0, st>>>(g, c, wtab, pd); \
    else k_project_bf16<MB, 3><<<grid, block, 0, st>>>(g, c, wtab, pd);            \
    break;
  if (mbw >= 500) {
    // 32x32 variant: code = 500 + RT
    const int rt = mbw - 500;
    if (m32_lds(rt, terms == 3 ? 2 : 3, g.M) > 160 * 1024) return PLG_ERR_UNSUPPORTED;
    const int nt32 = (2 * g.nmom + 31) / 32;
    dim3 g3((unsigned)((g.ncols + M32_NC - 1) / M32_NC), (unsigned)g.nsplit, (unsigned)((nt32 + rt - 1) / rt));
    const size_t dyn = (size_t)g.M * 8 * sizeof(uint32_t);
#define PLG_M32(RT)                                                                                 \
  if (rt == RT) {                                                                                   \
    if (terms == 6) k_project_bf16_m32<RT, 6><<<g3, block, dyn, st>>>(g, c, wtab, pd);               \
    else k_project_bf16_m32<RT, 3><<<g3, block, dyn, st>>>(g, c, wtab, pd);                          \
    return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;                                  \
  }
    PLG_M32(6)
    PLG_M32(5)
    PLG_M32(4)
    PLG_M32(3)
    PLG_M32(2)
    PLG_M32(1)
#undef PLG_M32
    return PLG_ERR_INVALID;
  }
  if (mbw >= 400) {
    // pipelined shared-A variant: code = 400 + RB
    const int rb = mbw - 400;
    dim3 g3((unsigned)((g.ncols + 127) / 128), (unsigned)g.nsplit, (unsigned)((nblk + rb - 1) / rb));
    dim3 b512(PP_THREADS);
    const size_t dyn = (size_t)g.M * 8 * sizeof(uint32_t);
#define PLG_PP(RB)                                                                                  \
  if (rb == RB) {                                                                                   \
    if (terms == 6) k_project_bf16_pp<RB, 6><<<g3, b512, dyn, st>>>(g, c, wtab, pd);                 \
    else k_project_bf16_pp<RB, 3><<<g3, b512, dyn, st>>>(g, c, wtab, pd);                            \
    return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;                                  \
  }
    PLG_PP(10)
    PLG_PP(8)
    PLG_PP(6)
    PLG_PP(4)
    PLG_PP(2)
#undef PLG_PP
    return PLG_ERR_INVALID;
  }
  if (mbw >= 100) {
    // shared-A variant: code = 100 + 10*RB + NBW
    const int rb = (mbw - 100) / 10, nbw = mbw % 10;
    const int nc = 64 * nbw;
    dim3 g3((unsigned)((g.ncols + nc - 1) / nc), (unsigned)g.nsplit, (unsigned)((nblk + rb - 1) / rb));
#define PLG_SA(RB, NBW)                                                                             \
  if (rb == RB && nbw == NBW) {                                                                     \
    if (terms == 6) k_project_bf16_sa<RB, NBW, 6><<<g3, block, 0, st>>>(g, c, wtab, pd);             \
    else k_project_bf16_sa<RB, NBW, 3><<<g3, block, 0, st>>>(g, c, wtab, pd);                        \
    return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;                                  \
  }
    PLG_SA(11, 1)
    PLG_SA(9, 1)
    PLG_SA(5, 2)
    PLG_SA(3, 2)
    PLG_SA(2, 2)
    PLG_SA(1, 2)
#undef PLG_SA
    return PLG_ERR_INVALID;
  }
  switch (mbw) {
    PLG_PB(1)
    PLG_PB(2)
    PLG_PB(3)
    PLG_PB(5)
    PLG_PB(9)
  default:
    return PLG_ERR_INVALID;
  }
#undef PLG_PB
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

// tile candidates of the autotuner for a momentum count (projection codes as
// accepted by launch_project_bf16)
std::vector<int> proj_candidates(int nmom) {
  const int nblk = (2 * nmom + 15) / 16;
  std::vector<int> v;
  for (int code : {211, 191, 152, 132, 122, 112}) {
    const int rb = (code - 100) / 10;
    if (rb <= 2 * nblk) v.push_back(code);
  }
  for (int code : {410, 408, 406, 404, 402}) {
    const int rb = code - 400;
    if (rb <= 2 * nblk) v.push_back(code);
  }
  const int nt32 = (2 * nmom + 31) / 32;
  for (int code : {506, 505, 504, 503, 502, 501})
    if (code - 500 <= nt32) v.push_back(code);
  v.push_back(pick_mbw(nmom));
  return v;
}

// workgroups of one K split for a projection code
int64_t proj_workgroups(int code, int64_t ncols, int nmom) {
  const int nblk = (2 * nmom + 15) / 16;
  if (code >= 500) {
    const int rt = code - 500, nt32 = (2 * nmom + 31) / 32;
    return ((ncols + M32_NC - 1) / M32_NC) * ((nt32 + rt - 1) / rt);
  }
  if (code >= 400) {
    const int rb = code - 400;
    return ((ncols + 127) / 128) * ((nblk + rb - 1) / rb);
  }
  if (code >= 100) {
    const int rb = (code - 100) / 10, nbw = code % 10;
    return ((ncols + 64 * nbw - 1) / (64 * nbw)) * ((nblk + rb - 1) / rb);
  }
  return ((ncols + PB_NC - 1) / PB_NC) * ((nblk + 4 * code - 1) / (4 * code));
}

// default: the shared-A tile whose row groups waste the fewest padded blocks
int pick_proj(int nmom) {
  const int nblk = (2 * nmom + 15) / 16;
  int best = 112, bestw = 1 << 30;
  for (int code : {112, 122, 132, 152, 191, 211}) {
    const int rb = (code - 100) / 10;
    const int waste = ((nblk + rb - 1) / rb) * rb - nblk;
    if (waste <= bestw) {
      best = code;
      bestw = waste;
    }
  }
  return best;
}

}  // namespace plg
