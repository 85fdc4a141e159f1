// plg_tetra.hip -- four-propagator table contractions (gfx950): the term
// sums of lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:6-62,
//   O_k = sum_t v_t sum_{a0..a3} P1[i0][i1][a[c0]][a[c1]] P2[i2][i3][a[c2]][a[c3]]
//                 conj(P3[i5+2][i4+2][a[c5]][a[c4]]) conj(P4[i7+2][i6+2][a[c7]][a[c6]])
// (spin indices + 2 mod 4: the gamma5 of the backward propagators), for
// tables in the reference's format (8 spin indices, 8 colour slots, a float2
// per term).
//
// The reference sums the 81 colour assignments of every term.  Every term of
// the reference's tables has one of two colour structures, which close into
// colour traces of the 3x3 blocks B1 = P1[i0][i1], ..., B4 = P4[i7+2][i6+2]:
//   direct  (c0 = c5, c1 = c4, c2 = c7, c3 = c6):  <B1, B3> <B2, B4>,   <X, Y> = sum X conj(Y)
//   crossed (c0 = c5, c1 = c6, c2 = c7, c3 = c4):  Tr[(B3^dag B1)(B4^dag B2)]
// (19 and 63 complex products instead of 243).  The host classifies every
// term; a table with another structure is refused.
//
// A workgroup of 4 waves stages 128-byte lines of the four propagators (16
// FP32 or 8 FP64 consecutive sites) in LDS; each wave contracts a site at a
// time, its lanes splitting an operator's terms, then a wave reduction.
#include "plg_internal.h"

#include <algorithm>

namespace plg {

template <typename T> struct TetraGeom {
  static constexpr int G = 128 / (int)sizeof(cx<T>);  // sites per line
  static constexpr int SPAD = 577;                    // LDS stride of a site (576 + 1)
};

// block pointers of a term inside a staged site (each 9 complex, [a][b])
struct TetraBlocks {
  int b1, b2, b3, b4;
};
__device__ __forceinline__ TetraBlocks tetra_blocks(uint32_t sp) {
  // sp: 4 bits per propagator: P1 i0 i1, P2 i2 i3, P3 (i5+2, i4+2), P4 (i7+2, i6+2)
  return TetraBlocks{(int)(sp & 15) * 9, 144 + (int)((sp >> 4) & 15) * 9, 288 + (int)((sp >> 8) & 15) * 9,
                     432 + (int)((sp >> 12) & 15) * 9};
}

template <typename T>
__global__ __launch_bounds__(256) void k_tetra(TetraSched S, const PropDev<T> *__restrict__ props,
                                               const RowDev4 *__restrict__ rows, cx<T> *__restrict__ out,
                                               int64_t V3, int T_lat, int nch_total, int groups_per_block) {
  constexpr int G = TetraGeom<T>::G, SPAD = TetraGeom<T>::SPAD;
  __shared__ cx<T> sp[G * SPAD];  // [site][f][m01][a b]
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
    for (int s = wave; s < G; s += 4) {
      const cx<T> *spx = sp + s * SPAD;
      const int64_t x = x0 + s;
      for (int k = 0; k < S.nops; k++) {
        cx<T> acc = cmake<T>(0, 0);
        const int n = S.count[k];
        const TetraTerm *E = S.terms + S.first[k];
        for (int t = lane; t < n; t += 64) {
          const TetraTerm u = E[t];
          const TetraBlocks bl = tetra_blocks(u.spins);
          const cx<T> *B1 = spx + bl.b1, *B2 = spx + bl.b2, *B3 = spx + bl.b3, *B4 = spx + bl.b4;
          cx<T> v;
          if (u.kind == 0) {  // direct: <B1, B3> <B2, B4>
            cx<T> x13 = cmake<T>(0, 0), x24 = cmake<T>(0, 0);
#pragma unroll
            for (int j = 0; j < 9; j++) {
              cfma_conj(x13, B1[j], B3[j]);
              cfma_conj(x24, B2[j], B4[j]);
            }
            v = x13 * x24;
          } else {  // crossed: Tr[(B3^dag B1)(B4^dag B2)]
            cx<T> m[9], nn[9];
#pragma unroll
            for (int z = 0; z < 3; z++)
#pragma unroll
              for (int y = 0; y < 3; y++) {
                cx<T> a = cmake<T>(0, 0), b = cmake<T>(0, 0);
#pragma unroll
                for (int w = 0; w < 3; w++) {
                  cfma_conj(a, B1[w * 3 + y], B3[w * 3 + z]);  // (B3^dag B1)[z][y]
                  cfma_conj(b, B2[w * 3 + z], B4[w * 3 + y]);  // (B4^dag B2)[y][z]
                }
                m[z * 3 + y] = a;
                nn[y * 3 + z] = b;
              }
            v = cmake<T>(0, 0);
#pragma unroll
            for (int z = 0; z < 3; z++)
#pragma unroll
              for (int y = 0; y < 3; y++) cfma(v, m[z * 3 + y], nn[y * 3 + z]);
          }
          cfma(acc, cmake<T>((T)u.re, (T)u.im), v);
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
          acc.re += __shfl_xor(acc.re, off);
          acc.im += __shfl_xor(acc.im, off);
        }
        if (lane == 0 && x < V3) gstore(out + ctile_idx(col0 + k, x, ncols), cmake<T>(sg * acc.re, sg * acc.im));
      }
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

// host: classify and pack the terms of a table (reference format)
int tetra_pack(const int16_t *idxs, const int16_t *cols, const float *vals, const int16_t *counts, int nops,
               std::vector<TetraTerm> &terms, std::vector<int> &first, std::vector<int> &count) {
  if (nops <= 0 || nops > TETRA_MAXOPS) return PLG_ERR_UNSUPPORTED;
  terms.clear();
  first.assign(nops, 0);
  count.assign(nops, 0);
  size_t t = 0;
  for (int k = 0; k < nops; k++) {
    if (counts[k] < 0) return PLG_ERR_INVALID;
    first[k] = (int)terms.size();
    std::vector<TetraTerm> direct, crossed;
    for (int i = 0; i < counts[k]; i++, t++) {
      const int16_t *m = idxs + 8 * t, *c = cols + 8 * t;
      for (int j = 0; j < 8; j++)
        if (m[j] < 0 || m[j] > 3 || c[j] < 0 || c[j] > 3) return PLG_ERR_INVALID;
      bool perm = true;  // c0..c3 a permutation of the four colour sums
      for (int a = 0; a < 4; a++)
        for (int b = a + 1; b < 4; b++) perm &= c[a] != c[b];
      if (!perm) return PLG_ERR_UNSUPPORTED;
      TetraTerm u;
      u.spins = (uint32_t)((m[0] * 4 + m[1]) | ((m[2] * 4 + m[3]) << 4) |
                           ((((m[5] + 2) & 3) * 4 + ((m[4] + 2) & 3)) << 8) |
                           ((((m[7] + 2) & 3) * 4 + ((m[6] + 2) & 3)) << 12));
      u.re = vals[2 * t];
      u.im = vals[2 * t + 1];
      if (c[0] == c[5] && c[1] == c[4] && c[2] == c[7] && c[3] == c[6]) {
        u.kind = 0;
        direct.push_back(u);
      } else if (c[0] == c[5] && c[1] == c[6] && c[2] == c[7] && c[3] == c[4]) {
        u.kind = 1;
        crossed.push_back(u);
      } else {
        return PLG_ERR_UNSUPPORTED;
      }
    }
    // one colour structure after the other: lanes of a wave take the same branch
    terms.insert(terms.end(), direct.begin(), direct.end());
    terms.insert(terms.end(), crossed.begin(), crossed.end());
    count[k] = (int)terms.size() - first[k];
  }
  return PLG_OK;
}

}  // namespace plg
