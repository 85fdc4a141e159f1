// plg_tetra.hip -- four-propagator table contractions (gfx950): the term
// sums of lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:6-62,
//   O_k = sum_t v_t sum_{a0..a3} P1[i0][i1][a[c0]][a[c1]] P2[i2][i3][a[c2]][a[c3]]
//                 conj(P3[i5+2][i4+2][a[c5]][a[c4]]) conj(P4[i7+2][i6+2][a[c7]][a[c6]])
// (spin indices + 2 mod 4: the gamma5 of the backward propagators), for
// tables in the reference's format (8 spin indices, 8 colour slots, a float2
// per term).
//
// The reference sums the 81 colour assignments of every term.  Every term of
// its heavy-light tables has one of two colour structures, which close into
// colour traces of the 3x3 blocks B1 = P1[i0][i1], ..., B4 = P4[i7+2][i6+2]:
//   direct  (c0 = c5, c1 = c4, c2 = c7, c3 = c6):  <B1, B3> <B2, B4>,   <X, Y> = sum X conj(Y)
//   crossed (c0 = c5, c1 = c6, c2 = c7, c3 = c4):  Tr[(B3^dag B1)(B4^dag B2)]
// Terms that pair P1 with P4 and P2 with P3 instead (the bcud tables) close
// into the same two structures on the views P1, P2, P4, P3; the host packs them
// as a second set, contracted by a second launch that adds to the first.
// The host classifies every term; a table with another structure is refused.
//
// Both structures factor through the spin pairs p = (B1, B3) and
// q = (B2, B4), 256 of each per site: per site the workgroup builds the
// tables <B1,B3>, B3^dag B1 (over p) and <B2,B4>, B4^dag B2 (over q), 36
// complex products per entry, and then a direct term costs one complex
// product of two table entries and a crossed term the trace of a product of
// two (9 products), instead of 19 and 63 from the blocks.  The terms of the
// reference's tables use every p and q, about 50 terms per entry.
//
// A workgroup of 4 waves stages 64-byte lines of the four propagators (8
// FP32 or 4 FP64 consecutive sites) in LDS; per site, each thread builds one
// p entry and one q entry of the tables, then each wave takes every fourth
// operator, its lanes splitting the operator's terms, and a wave reduction.
// In FP32 a spin pair's entries are ten (the 3x3 product, then the scalar
// product), read as five 16-byte pairs: 83.0 -> 76.3 ms on the benchmark's
// table against nine entries read one by one.
#include "plg_internal.h"

#include <algorithm>
#include <cstring>
#include <map>

namespace plg {

constexpr int TETRA_TB = 12;      // terms per lane loaded together
constexpr int TETRA_LDS_COEF = 256;  // coefficients kept in LDS

template <typename T> struct TetraGeom {
  static constexpr int G = 64 / (int)sizeof(cx<T>);  // sites per staged line
  static constexpr int SPAD = 577;                   // LDS stride of a site (576 + 1)
};

template <typename T>
__global__ __launch_bounds__(256) void k_tetra(TetraSched S, const PropDev<T> *__restrict__ props,
                                               const RowDev4 *__restrict__ rows, cx<T> *__restrict__ out,
                                               int64_t V3, int T_lat, int nch_total, int groups_per_block) {
  constexpr int G = TetraGeom<T>::G, SPAD = TetraGeom<T>::SPAD;
  __shared__ cx<T> sp[G * SPAD];                 // [site][f][m01][a b]
  // per spin pair: the 9 entries of (B3^dag B1)[z][y] (over p) or
  // (B4^dag B2)[y][z] (over q) at z*3+y; in FP32 then <B1,B3> or <B2,B4>,
  // ten entries, so that entry pairs are 16-byte reads (in FP64 an entry is
  // one, and nine entries keep the spin pairs on distinct banks)
  constexpr bool F32 = sizeof(T) == 4;
  constexpr int SPP = F32 ? 10 : 9;
  __shared__ __align__(16) cx<T> m13[256 * SPP], n24[256 * SPP];
  __shared__ cx<T> d13[F32 ? 1 : 256], d24[F32 ? 1 : 256];  // FP64: <B1,B3> over p, <B2,B4> over q
  __shared__ cx<T> scoef[TETRA_LDS_COEF];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const RowDev4 row = rows[blockIdx.y];
  const PropDev<T> P0 = props[row.p[0]], P1 = props[row.p[1]], P2 = props[row.p[2]], P3 = props[row.p[3]];
  auto plane_base = [&](const PropDev<T> &Q) {
    int plane = row.t - Q.t_first;
    plane += plane < 0 ? T_lat : 0;
    return Q.data + (int64_t)plane * V3;
  };
  const cx<T> *b0 = plane_base(P0), *b1 = plane_base(P1), *b2 = plane_base(P2), *b3 = plane_base(P3);
  const T sg = prop_sign(P0, row.t) * prop_sign(P1, row.t) * prop_sign(P2, row.t) * prop_sign(P3, row.t);
  const int rot_any = P0.rot | P1.rot | P2.rot | P3.rot;
  const int64_t ncols = (int64_t)gridDim.y * nch_total, col0 = (int64_t)blockIdx.y * nch_total;
  // the coefficients in LDS when they fit (the reference's tables have six)
  const bool coef_lds = S.ncoef <= TETRA_LDS_COEF;
  if (coef_lds)
    for (int i = tid; i < S.ncoef; i += 256) scoef[i] = cmake<T>((T)S.coef[2 * i], (T)S.coef[2 * i + 1]);

  for (int gi = 0; gi < groups_per_block; gi++) {
    const int64_t x0 = ((int64_t)blockIdx.x * groups_per_block + gi) * G;
    if (x0 >= V3) break;  // uniform over the workgroup
    __syncthreads();      // the previous group's sites are done
    for (int idx = tid; idx < 576 * G; idx += 256) {
      const int comp = idx / G, s = idx - comp * G;
      const int f = comp / 144, cl = comp - f * 144;
      const cx<T> *b = f == 0 ? b0 : (f == 1 ? b1 : (f == 2 ? b2 : b3));
      const int64_t cs = f == 0 ? P0.cs : (f == 1 ? P1.cs : (f == 2 ? P2.cs : P3.cs));
      const int64_t x = x0 + s;
      sp[s * SPAD + comp] = x < V3 ? gload(b + (int64_t)cl * cs + x) : cmake<T>(0, 0);
    }
    __syncthreads();
    if (rot_any) {  // rotation to the physical basis, per (site, propagator, colour pair)
      for (int k = tid; k < G * 36; k += 256) {
        const int s = k / 36, f = (k % 36) / 9, cc = k % 9;
        const int rot = f == 0 ? P0.rot : (f == 1 ? P1.rot : (f == 2 ? P2.rot : P3.rot));
        if (rot) {
          cx<T> *b = sp + s * SPAD + f * 144 + cc;
          cx<T> v[16];
#pragma unroll
          for (int m = 0; m < 16; m++) v[m] = b[m * 9];
          rot_spins(v, rot);
#pragma unroll
          for (int m = 0; m < 16; m++) b[m * 9] = v[m];
        }
      }
      __syncthreads();
    }
    for (int s = 0; s < G; s++) {
      const int64_t x = x0 + s;
      if (x >= V3) break;  // uniform
      const cx<T> *spx = sp + s * SPAD;
      // tables of the site: thread tid = 16 hi + lo builds entry tid of both
      {
        const int hi = tid >> 4, lo = tid & 15;
        cx<T> a[9], b[9];
#pragma unroll
        for (int j = 0; j < 9; j++) {
          a[j] = spx[hi * 9 + j];          // B1 = P1 block s1 = hi
          b[j] = spx[288 + lo * 9 + j];    // B3 = P3 block s3 = lo
        }
        cx<T> d = cmake<T>(0, 0);
#pragma unroll
        for (int j = 0; j < 9; j++) cfma_conj(d, a[j], b[j]);
        if constexpr (F32) m13[tid * SPP + 9] = d;
        else d13[tid] = d;
#pragma unroll
        for (int z = 0; z < 3; z++)
#pragma unroll
          for (int y = 0; y < 3; y++) {
            cx<T> m = cmake<T>(0, 0);
#pragma unroll
            for (int w = 0; w < 3; w++) cfma_conj(m, a[w * 3 + y], b[w * 3 + z]);
            m13[tid * SPP + z * 3 + y] = m;
          }
#pragma unroll
        for (int j = 0; j < 9; j++) {
          a[j] = spx[144 + hi * 9 + j];    // B2 = P2 block s2 = hi
          b[j] = spx[432 + lo * 9 + j];    // B4 = P4 block s4 = lo
        }
        d = cmake<T>(0, 0);
#pragma unroll
        for (int j = 0; j < 9; j++) cfma_conj(d, a[j], b[j]);
        if constexpr (F32) n24[tid * SPP + 9] = d;
        else d24[tid] = d;
#pragma unroll
        for (int z = 0; z < 3; z++)
#pragma unroll
          for (int y = 0; y < 3; y++) {
            cx<T> m = cmake<T>(0, 0);
#pragma unroll
            for (int w = 0; w < 3; w++) cfma_conj(m, a[w * 3 + z], b[w * 3 + y]);
            n24[tid * SPP + z * 3 + y] = m;
          }
      }
      __syncthreads();
      // one term of the tables: direct <B1,B3><B2,B4> or crossed trace
      auto term = [&](uint32_t u, bool crossed) {
        const int p = u & 255, q = (u >> 8) & 255, ci = (int)(u >> 16);
        cx<T> v;
        if constexpr (F32) {
          // 16-byte reads: entry pairs (0,1) .. (6,7), then (8, the <,> value)
          typedef float f4 __attribute__((ext_vector_type(4)));
          const f4 *a = reinterpret_cast<const f4 *>(m13 + p * SPP);
          const f4 *b = reinterpret_cast<const f4 *>(n24 + q * SPP);
          if (!crossed) {
            const f4 x = a[4], y = b[4];
            v = cmake<T>(x[2], x[3]) * cmake<T>(y[2], y[3]);
          } else {
            v = cmake<T>(0, 0);
#pragma unroll
            for (int k = 0; k < 4; k++) {
              const f4 x = a[k], y = b[k];
              cfma(v, cmake<T>(x[0], x[1]), cmake<T>(y[0], y[1]));
              cfma(v, cmake<T>(x[2], x[3]), cmake<T>(y[2], y[3]));
            }
            const f4 x = a[4], y = b[4];
            cfma(v, cmake<T>(x[0], x[1]), cmake<T>(y[0], y[1]));
          }
        } else {
          if (!crossed) {
            v = d13[p] * d24[q];
          } else {
            v = cmake<T>(0, 0);
#pragma unroll
            for (int j = 0; j < 9; j++) cfma(v, m13[p * SPP + j], n24[q * SPP + j]);
          }
        }
        const cx<T> c = coef_lds ? scoef[ci] : cmake<T>((T)S.coef[2 * ci], (T)S.coef[2 * ci + 1]);
        return c * v;
      };
      auto emit = [&](int k, cx<T> acc) {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          acc.re += __shfl_xor(acc.re, off);
          acc.im += __shfl_xor(acc.im, off);
        }
        if (lane == 0) {
          cx<T> *o = out + ctile_idx(col0 + k, x, ncols);
          cx<T> v = cmake<T>(sg * acc.re, sg * acc.im);
          if (S.accumulate) v = v + gload(o);  // the set written by the previous launch
          gstore(o, v);
        }
      };
      // wave w takes the operators k = w, w + 4, ...; its lanes split an
      // operator's terms, TETRA_TB terms per lane loaded together, the next
      // batch (of this operator or the wave's next one) in flight while one
      // is summed
      int k = wave, base = 0;
      uint32_t u[TETRA_TB];
      auto fetch = [&](int kk, int b, uint32_t *dst) {
        const int n = kk < S.nops ? S.count[kk] : 0;
        const TetraTerm *E = S.terms + (kk < S.nops ? S.first[kk] : 0);
#pragma unroll
        for (int i = 0; i < TETRA_TB; i++) {
          const int t = b + lane + 64 * i;
          dst[i] = t < n ? E[t] : 0u;
        }
      };
      fetch(k, 0, u);
      cx<T> acc = cmake<T>(0, 0);
      while (k < S.nops) {
        const int n = S.count[k], nd = S.ndirect[k];
        // the batch after this one
        int k2 = k, b2 = base + 64 * TETRA_TB;
        if (b2 >= n) {
          k2 = k + 4;
          b2 = 0;
        }
        uint32_t v[TETRA_TB];
        fetch(k2, b2, v);
#pragma unroll
        for (int i = 0; i < TETRA_TB; i++) {
          const int t = base + lane + 64 * i;
          if (t < n) acc = acc + term(u[i], t >= nd);
        }
        if (k2 != k) {
          emit(k, acc);
          acc = cmake<T>(0, 0);
        }
        k = k2;
        base = b2;
#pragma unroll
        for (int i = 0; i < TETRA_TB; i++) u[i] = v[i];
      }
      __syncthreads();  // the tables may be rebuilt for the next site
    }
  }
}

template <typename T>
int launch_tetra(const TetraSched &S, const PropDev<T> *props, const RowDev4 *rows, int nrows, cx<T> *out,
                 int64_t V3, int T_lat, int nch_total, int spw, hipStream_t st) {
  constexpr int G = TetraGeom<T>::G;
  const int gpb = std::max(1, spw / G);
  const int64_t ngroups = (V3 + G - 1) / G;
  dim3 block(256), grid((unsigned)((ngroups + gpb - 1) / gpb), (unsigned)nrows);
  k_tetra<T><<<grid, block, 0, st>>>(S, props, rows, out, V3, T_lat, nch_total, gpb);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

template int launch_tetra<double>(const TetraSched &, const PropDev<double> *, const RowDev4 *, int, cx<double> *,
                                  int64_t, int, int, int, hipStream_t);
template int launch_tetra<float>(const TetraSched &, const PropDev<float> *, const RowDev4 *, int, cx<float> *,
                                 int64_t, int, int, int, hipStream_t);

// Order of an operator's terms of one structure.  The kernel's lanes take
// consecutive terms, and the 16 lanes of a row read LDS together, each a table
// entry p * 9 + j (or q * 9 + j): two different spin pairs with the same
// residue mod 16 in one run of 16 fall on one bank, and the read takes as many
// passes as the largest number of distinct spin pairs on a bank (equal pairs
// broadcast).  The host therefore builds runs of 16 (the first one filling the
// 16-term group the previous structure's terms end in) with distinct residues
// of p and of q where the table allows (the order was fitted to 8-byte
// reads; for the FP32 16-byte pair reads an order of runs of 8 with distinct
// residues mod 8 measured the same): each slot takes a term from the
// (p mod 16, q mod 16) bucket whose residues are still free in the run,
// preferring the residues with most terms left; then pairwise swaps between
// runs that do not raise the two runs' summed passes.  On the reference's
// tables the passes per read drop from 1.6 (a Latin-square walk over the
// buckets) to 1.3.  The sum is the same up to rounding.
static int run_passes(const TetraTerm *r, int n) {
  uint16_t mp[16] = {0}, mq[16] = {0};  // distinct spin pairs per residue (bit s1 / s2)
  for (int i = 0; i < n; i++) {
    const int p = r[i] & 255, q = (r[i] >> 8) & 255;
    mp[p & 15] |= (uint16_t)(1u << (p >> 4));
    mq[q & 15] |= (uint16_t)(1u << (q >> 4));
  }
  int dp = 0, dq = 0;
  for (int k = 0; k < 16; k++) {
    dp = std::max(dp, __builtin_popcount(mp[k]));
    dq = std::max(dq, __builtin_popcount(mq[k]));
  }
  return dp + dq;
}

static void order_runs(std::vector<TetraTerm> &v, int first_run) {
  const int n = (int)v.size();
  if (n <= 1) return;
  std::vector<TetraTerm> bucket[256];
  int cp[16] = {0}, cq[16] = {0};
  for (TetraTerm u : v) {
    const int pr = (u & 255) & 15, qr = ((u >> 8) & 255) & 15;
    bucket[pr * 16 + qr].push_back(u);
    cp[pr]++;
    cq[qr]++;
  }
  std::vector<int> start;  // run boundaries in the output
  std::vector<TetraTerm> out;
  out.reserve(n);
  int sz = first_run > 0 ? first_run : 16;
  while ((int)out.size() < n) {
    start.push_back((int)out.size());
    unsigned usedp = 0, usedq = 0;
    for (int slot = 0; slot < sz && (int)out.size() < n; slot++) {
      int best = -1, bs = -1;
      for (int b = 0; b < 256; b++) {
        if (bucket[b].empty()) continue;
        const int pr = b >> 4, qr = b & 15;
        const int sc = ((usedp >> pr & 1) ? 0 : 1 << 22) + ((usedq >> qr & 1) ? 0 : 1 << 22) + cp[pr] + cq[qr];
        if (sc > bs) {
          bs = sc;
          best = b;
        }
      }
      const int pr = best >> 4, qr = best & 15;
      out.push_back(bucket[best].back());
      bucket[best].pop_back();
      cp[pr]--;
      cq[qr]--;
      usedp |= 1u << pr;
      usedq |= 1u << qr;
    }
    sz = 16;
  }
  start.push_back(n);
  const int nr = (int)start.size() - 1;
  if (nr > 1) {
    std::vector<int> cost(nr);
    for (int r = 0; r < nr; r++) cost[r] = run_passes(&out[start[r]], start[r + 1] - start[r]);
    uint64_t st = 0x9e3779b97f4a7c15ull ^ (uint64_t)n;  // fixed seed: a table always packs alike
    auto rnd = [&](uint64_t m) {
      st = st * 6364136223846793005ull + 1442695040888963407ull;
      return (int)((st >> 33) % m);
    };
    for (int it = 0; it < 16 * n; it++) {
      const int a = rnd(nr), b = rnd(nr);
      if (a == b) continue;
      const int i = start[a] + rnd(start[a + 1] - start[a]), j = start[b] + rnd(start[b + 1] - start[b]);
      std::swap(out[i], out[j]);
      const int ca = run_passes(&out[start[a]], start[a + 1] - start[a]);
      const int cb = run_passes(&out[start[b]], start[b + 1] - start[b]);
      if (ca + cb <= cost[a] + cost[b]) {
        cost[a] = ca;
        cost[b] = cb;
      } else {
        std::swap(out[i], out[j]);
      }
    }
  }
  v.swap(out);
}

// host: classify and pack the terms of a table (reference format)
int tetra_pack(const int16_t *idxs, const int16_t *cols, const float *vals, const int16_t *counts, int nops,
               TetraPacked &out) {
  if (nops <= 0 || nops > TETRA_MAXOPS) return PLG_ERR_UNSUPPORTED;
  out = TetraPacked{};
  for (int h = 0; h < 2; h++) {
    out.first[h].assign(nops, 0);
    out.count[h].assign(nops, 0);
    out.ndirect[h].assign(nops, 0);
  }
  std::map<std::pair<uint32_t, uint32_t>, uint32_t> cidx;  // coefficient bits -> index
  cidx.emplace(std::make_pair(0u, 0u), 0u);                 // entry 0: zero
  out.coef.push_back(0.0f);
  out.coef.push_back(0.0f);
  size_t t = 0;
  for (int k = 0; k < nops; k++) {
    if (counts[k] < 0) return PLG_ERR_INVALID;
    std::vector<TetraTerm> direct[2], crossed[2];
    for (int i = 0; i < counts[k]; i++, t++) {
      const int16_t *m = idxs + 8 * t, *c = cols + 8 * t;
      for (int j = 0; j < 8; j++)
        if (m[j] < 0 || m[j] > 3 || c[j] < 0 || c[j] > 3) return PLG_ERR_INVALID;
      bool perm = true;  // c0..c3 a permutation of the four colour sums
      for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) perm &= c[a] != c[b];
      if (!perm) return PLG_ERR_UNSUPPORTED;
      const uint32_t s1 = m[0] * 4 + m[1], s2 = m[2] * 4 + m[3];
      const uint32_t s3 = ((m[5] + 2) & 3) * 4 + ((m[4] + 2) & 3), s4 = ((m[7] + 2) & 3) * 4 + ((m[6] + 2) & 3);
      uint32_t re, im;
      std::memcpy(&re, vals + 2 * t, 4);
      std::memcpy(&im, vals + 2 * t + 1, 4);
      auto it = cidx.find({re, im});
      uint32_t ci;
      if (it != cidx.end()) {
        ci = it->second;
      } else {
        ci = (uint32_t)cidx.size();
        if (ci >= (uint32_t)TETRA_MAXCOEF) return PLG_ERR_UNSUPPORTED;
        cidx.emplace(std::make_pair(re, im), ci);
        out.coef.push_back(vals[2 * t]);
        out.coef.push_back(vals[2 * t + 1]);
      }
      // set 0 on (P1, P2, P3, P4); set 1 on (P1, P2, P4, P3): P4's spin pair
      // takes the third place and P3's the fourth
      const TetraTerm u0 = (s1 * 16 + s3) | ((s2 * 16 + s4) << 8) | (ci << 16);
      const TetraTerm u1 = (s1 * 16 + s4) | ((s2 * 16 + s3) << 8) | (ci << 16);
      if (c[0] == c[5] && c[1] == c[4] && c[2] == c[7] && c[3] == c[6]) direct[0].push_back(u0);
      else if (c[0] == c[5] && c[1] == c[6] && c[2] == c[7] && c[3] == c[4]) crossed[0].push_back(u0);
      else if (c[0] == c[7] && c[1] == c[6] && c[2] == c[5] && c[3] == c[4]) direct[1].push_back(u1);
      else if (c[0] == c[7] && c[1] == c[4] && c[2] == c[5] && c[3] == c[6]) crossed[1].push_back(u1);
      else return PLG_ERR_UNSUPPORTED;
    }
    for (int h = 0; h < 2; h++) {
      std::vector<TetraTerm> &terms = out.terms[h];
      out.first[h][k] = (int)terms.size();
      order_runs(direct[h], 16);
      // the crossed terms start in the 16-term group the direct ones end in
      order_runs(crossed[h], (int)(16 - direct[h].size() % 16));
      terms.insert(terms.end(), direct[h].begin(), direct[h].end());
      terms.insert(terms.end(), crossed[h].begin(), crossed[h].end());
      out.ndirect[h][k] = (int)direct[h].size();
      out.count[h][k] = (int)terms.size() - out.first[h][k];
    }
  }
  return PLG_OK;
}

}  // namespace plg
