// plg_udsc.hip -- the udsc baryon two-point functions (gfx950), the table-
// driven three-propagator contraction of lib/kernels/PLEGMA_baryons_udsc.cu:
//   O_k(x) = sum_t v_kt sum_{eps, eps'} sgn sgn' S1[m0][m1][a][a'] S2[m2][m3][b][b'] S3[m4][m5][c][c']
// with the term tables of plg_udsc_tables.h (generated from the interpolating
// fields by plegma_amd/gen_udsc.py).
//
// The reference gives every site one thread that walks all terms with its
// three propagators in registers (PLEGMA_baryons_udsc.cu:86-121).  Here one
// wave works on one site at a time with the site's 432 propagator entries in
// LDS, and the terms are factorised through the colour "diquark" of the
// first two propagators:
//   D_q[c][c'] = sum eps_abc eps_a'b'c' S1[m0m1][a][a'] S2[m2m3][b][b']   (q = m0m1, m2m3)
//   T_m        = sum_cc' D_q[c][c'] S3[m4m5][c][c']                     (distinct terms m)
//   O_k        = sum_m v_km T_m
// Quads are processed in chunks of 64, one per lane: the lane forms D_q in
// registers and the (at most four) distinct terms of its quad, which go to
// LDS slots; then each operator sums its entries over those slots (4 lanes per
// operator).  D costs 36 complex products per quad, a term 9, an entry 1.
// A workgroup stages whole 128-byte lines of the propagators (16 FP32 or 8
// FP64 consecutive sites), so every line is read from HBM once.
#include "plg_internal.h"
#include "plg_udsc_tables.h"

#include <algorithm>
#include <map>
#include <vector>

namespace plg {

// ------------------------------------------------------------------ schedule
// host: the chunked schedule of nops operators, operator k being the terms
// [first[k], first[k] + count[k]) of `terms`
int udsc_schedule_terms(const UdscTerm *terms, const int *first, const int *count, int nops, UdscSchedHost &h) {
  if (nops <= 0 || nops > UD_MAXOPS) return PLG_ERR_UNSUPPORTED;
  auto key6 = [](const UdscTerm &t) {
    int k = 0;
    for (int i = 0; i < 6; i++) k = k * 4 + t.m[i];
    return k;
  };
  auto quad_of = [](int k6) {  // m01 | m23 << 4
    const int m0 = (k6 >> 10) & 3, m1 = (k6 >> 8) & 3, m2 = (k6 >> 6) & 3, m3 = (k6 >> 4) & 3;
    return (m0 * 4 + m1) | ((m2 * 4 + m3) << 4);
  };
  for (int k = 0; k < nops; k++)
    for (int i = 0; i < count[k]; i++)
      for (int j = 0; j < 6; j++)
        if (terms[first[k] + i].m[j] < 0 || terms[first[k] + i].m[j] > 3) return PLG_ERR_INVALID;
  // distinct terms grouped by quad
  std::map<int, std::vector<int>> by_quad;  // quad -> distinct key6 terms
  for (int k = 0; k < nops; k++)
    for (int i = 0; i < count[k]; i++) {
      const int k6 = key6(terms[first[k] + i]);
      auto &v = by_quad[quad_of(k6)];
      if (std::find(v.begin(), v.end(), k6) == v.end()) v.push_back(k6);
    }
  std::vector<int> quads;
  for (auto &kv : by_quad) {
    if ((int)kv.second.size() > UD_TPQ) return PLG_ERR_UNSUPPORTED;
    quads.push_back(kv.first);
  }
  std::sort(quads.begin(), quads.end(), [](int a, int b) {
    return (a & 15) != (b & 15) ? (a & 15) < (b & 15) : (a >> 4) < (b >> 4);
  });
  const int nch = std::max(1, ((int)quads.size() + 63) / 64);
  if (nch > UD_MAXC) return PLG_ERR_UNSUPPORTED;
  std::map<int, std::pair<int, int>> slot;  // key6 -> (chunk, chunk-local slot)
  h = UdscSchedHost{};
  h.s.nops = nops;
  h.s.nchunks = nch;
  h.quads.assign((size_t)nch * 64, (short)-1);
  h.tp.assign((size_t)nch * 64 * UD_TPQ, (signed char)-1);
  for (size_t i = 0; i < quads.size(); i++) {
    const int c = (int)i / 64, l = (int)i % 64;
    h.quads[i] = (short)quads[i];
    const auto &v = by_quad[quads[i]];
    for (size_t e = 0; e < v.size(); e++) {
      h.tp[i * UD_TPQ + e] = (signed char)((v[e] >> 2 & 3) * 4 + (v[e] & 3));
      slot[v[e]] = {c, l * UD_TPQ + (int)e};
    }
  }
  // each operator's entries, per chunk: slot | coefficient index << 8, in
  // four sublists padded with zero-coefficient entries
  std::map<std::pair<float, float>, int> cix;
  h.coef = {0.f, 0.f};
  cix[{0.f, 0.f}] = 0;
  for (int c = 0; c < nch; c++)
    for (int k = 0; k < nops; k++) {
      std::vector<uint16_t> v;
      for (int i = 0; i < count[k]; i++) {
        const UdscTerm &t = terms[first[k] + i];
        const auto &sl = slot[key6(t)];
        if (sl.first != c) continue;
        auto it = cix.find({t.re, t.im});
        if (it == cix.end()) {
          if ((int)h.coef.size() / 2 >= UD_MAXCOEF) return PLG_ERR_UNSUPPORTED;
          it = cix.insert({{t.re, t.im}, (int)h.coef.size() / 2}).first;
          h.coef.push_back(t.re);
          h.coef.push_back(t.im);
        }
        v.push_back((uint16_t)(sl.second | (it->second << 8)));
      }
      const int len = (((int)v.size() + 3) / 4 + 3) / 4 * 4;
      h.s.ent_off[c][k] = (int)h.ents.size();
      h.s.ent_len[c][k] = len;
      for (int r = 0; r < 4; r++)
        for (int j = 0; j < len; j++) {
          const size_t e = (size_t)r * len + j;  // contiguous quarters of the list
          h.ents.push_back(e < v.size() ? v[e] : (uint16_t)0);
        }
    }
  h.s.nent = (int)h.ents.size();
  h.s.ncoef = (int)h.coef.size() / 2;
  if (h.s.nent > UD_MAXENT || h.s.ncoef > UD_MAXCOEF || 64 * UD_TPQ > 256) return PLG_ERR_UNSUPPORTED;
  return PLG_OK;
}

// the schedule of a flavour product of the built-in tables
int udsc_schedule(int prod, UdscSchedHost &h) {
  if (prod < 0 || prod >= UDSC_NPROD) return PLG_ERR_INVALID;
  const UdscProd &P = UDSC_PRODS[prod];
  std::vector<int> first(P.nops), count(P.nops);
  for (int k = 0; k < P.nops; k++) {
    first[k] = UDSC_OPS[P.first_op + k].first;
    count[k] = UDSC_OPS[P.first_op + k].count;
  }
  return udsc_schedule_terms(UDSC_TERMS, first.data(), count.data(), P.nops, h);
}

int udsc_products() { return UDSC_NPROD; }
const char *udsc_flavours(int prod) { return UDSC_PRODS[prod].flavours; }
int udsc_nops(int prod) { return UDSC_PRODS[prod].nops; }
const char *udsc_op_name(int prod, int k) { return UDSC_OPS[UDSC_PRODS[prod].first_op + k].name; }

// ------------------------------------------------------------------ kernel
// A workgroup of 4 waves stages the three propagators of G consecutive sites
// (one 128-byte line of every component) into LDS with coalesced loads; each
// wave then contracts G/4 of those sites, one at a time.
// LDS written and read by lanes of one wave only: a wavefront-scope fence
// orders the accesses (a wave's LDS operations are performed in order)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <typename T> struct UdscGeom {
  static constexpr int G = 128 / (int)sizeof(cx<T>);  // sites per line
  static constexpr int SPAD = 433;                    // LDS stride of a site (432 + 1)
};

template <typename T>
__global__ __launch_bounds__(256) void k_udsc(UdscSched S, const PropDev<T> *__restrict__ props,
                                              const RowDev *__restrict__ rows, cx<T> *__restrict__ out, int64_t V3,
                                              int T_lat, int nch_total, int ch0, int groups_per_block) {
  constexpr int G = UdscGeom<T>::G, SPAD = UdscGeom<T>::SPAD;
  __shared__ cx<T> sp[G * SPAD];              // [site][f][m01][c c']
  __shared__ cx<T> sterm[4][64 * UD_TPQ];     // each wave's chunk terms
  __shared__ __align__(16) uint16_t sent[UD_MAXENT];
  __shared__ float scoef[2 * UD_MAXCOEF];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const RowDev row = rows[blockIdx.y];
  const PropDev<T> P0 = props[row.pa], P1 = props[row.pb], P2 = props[row.pc];
  auto plane_base = [&](const PropDev<T> &Q) {
    int plane = row.t - Q.t_first;
    plane += plane < 0 ? T_lat : 0;
    return Q.data + (int64_t)plane * V3;
  };
  const cx<T> *b0 = plane_base(P0), *b1 = plane_base(P1), *b2 = plane_base(P2);
  const T sg = prop_sign(P0, row.t) * prop_sign(P1, row.t) * prop_sign(P2, row.t);
  const int rot_any = P0.rot | P1.rot | P2.rot;
  const int64_t ncols = (int64_t)gridDim.y * nch_total, col0 = (int64_t)blockIdx.y * nch_total + ch0;
  const int kop = lane >> 2, rr = lane & 3;
  const bool op_lane = kop < S.nops;
  // the lane's schedule in registers; the entries and coefficients in LDS
  int qv[UD_MAXC], tpv[UD_MAXC], eoff[UD_MAXC], en[UD_MAXC];
#pragma unroll
  for (int c = 0; c < UD_MAXC; c++) {
    const bool live = c < S.nchunks;
    qv[c] = live ? S.quads[c * 64 + lane] : -1;
    tpv[c] = live ? *(const int *)(S.tp + (c * 64 + lane) * UD_TPQ) : -1;
    en[c] = live && op_lane ? S.ent_len[c][kop] : 0;
    eoff[c] = live && op_lane ? S.ent_off[c][kop] + rr * en[c] : 0;
  }
  for (int i = tid; i < S.nent; i += 256) sent[i] = S.ents[i];
  for (int i = tid; i < 2 * S.ncoef; i += 256) scoef[i] = S.coef[i];
  cx<T> *st = sterm[wave];

  for (int gi = 0; gi < groups_per_block; gi++) {
    const int64_t x0 = ((int64_t)blockIdx.x * groups_per_block + gi) * G;
    if (x0 >= V3) break;  // uniform over the workgroup
    __syncthreads();      // the previous group's sites are done
    // stage: element idx = comp * G + s, lanes of a component read G consecutive sites
    for (int idx = tid; idx < 432 * G; idx += 256) {
      const int comp = idx / G, s = idx - comp * G;
      const int f = comp / 144, cl = comp - f * 144;
      const cx<T> *b = f == 0 ? b0 : (f == 1 ? b1 : b2);
      const int64_t cs = f == 0 ? P0.cs : (f == 1 ? P1.cs : P2.cs);
      const int64_t x = x0 + s;
      sp[s * SPAD + comp] = x < V3 ? gload(b + (int64_t)cl * cs + x) : cmake<T>(0, 0);
    }
    __syncthreads();
    if (rot_any) {  // rotation to the physical basis, per (site, propagator, colour pair)
      for (int k = tid; k < G * 27; k += 256) {
        const int s = k / 27, f = (k % 27) / 9, cc = k % 9;
        const int rot = f == 0 ? P0.rot : (f == 1 ? P1.rot : P2.rot);
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
      cx<T> acc = cmake<T>(0, 0);
#pragma unroll
      for (int c = 0; c < UD_MAXC; c++) {
        if (c >= S.nchunks) break;
        const int q = qv[c];
        if (q >= 0) {
          // D of this lane's quad, in registers
          const cx<T> *A = spx + (q & 15) * 9, *B = spx + 144 + (q >> 4) * 9;
          cx<T> a[9], b[9], d[9];
#pragma unroll
          for (int j = 0; j < 9; j++) {
            a[j] = A[j];
            b[j] = B[j];
          }
#pragma unroll
          for (int cc = 0; cc < 3; cc++) {
            const int a1 = (cc + 1) % 3, a2 = (cc + 2) % 3;
#pragma unroll
            for (int dd = 0; dd < 3; dd++) {
              const int e1 = (dd + 1) % 3, e2 = (dd + 2) % 3;
              // + a[a1][e1] b[a2][e2] - a[a1][e2] b[a2][e1] - a[a2][e1] b[a1][e2] + a[a2][e2] b[a1][e1]
              const cx<T> p0 = a[a1 * 3 + e1], q0 = b[a2 * 3 + e2], p1 = a[a1 * 3 + e2], q1 = b[a2 * 3 + e1];
              const cx<T> p2 = a[a2 * 3 + e1], q2 = b[a1 * 3 + e2], p3 = a[a2 * 3 + e2], q3 = b[a1 * 3 + e1];
              T re = p0.re * q0.re;
              re = fma(-p0.im, q0.im, re);
              re = fma(-p1.re, q1.re, re);
              re = fma(p1.im, q1.im, re);
              re = fma(-p2.re, q2.re, re);
              re = fma(p2.im, q2.im, re);
              re = fma(p3.re, q3.re, re);
              re = fma(-p3.im, q3.im, re);
              T im = p0.re * q0.im;
              im = fma(p0.im, q0.re, im);
              im = fma(-p1.re, q1.im, im);
              im = fma(-p1.im, q1.re, im);
              im = fma(-p2.re, q2.im, im);
              im = fma(-p2.im, q2.re, im);
              im = fma(p3.re, q3.im, im);
              im = fma(p3.im, q3.re, im);
              d[cc * 3 + dd] = cmake<T>(re, im);
            }
          }
          // its distinct terms T = sum_cc' D[c][c'] S3[m4m5][c][c']
#pragma unroll
          for (int e = 0; e < UD_TPQ; e++) {
            const int p = (signed char)(tpv[c] >> (8 * e));
            if (p >= 0) {
              const cx<T> *C3 = spx + 288 + p * 9;
              cx<T> v = cmake<T>(0, 0);
#pragma unroll
              for (int j = 0; j < 9; j++) cfma(v, d[j], C3[j]);
              st[lane * UD_TPQ + e] = v;
            }
          }
        }
        wave_sync();
        // operator entries: 4 lanes per operator, each a quarter of the list,
        // four entries per LDS read
        for (int e = 0; e < en[c]; e += 4) {
          const uint2 w = *(const uint2 *)(sent + eoff[c] + e);
          const int u[4] = {(int)(w.x & 0xffff), (int)(w.x >> 16), (int)(w.y & 0xffff), (int)(w.y >> 16)};
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const int ci = u[j] >> 8;
            cfma(acc, cmake<T>((T)scoef[2 * ci], (T)scoef[2 * ci + 1]), st[u[j] & 255]);
          }
        }
        wave_sync();
      }
      acc.re += __shfl_xor(acc.re, 1);
      acc.im += __shfl_xor(acc.im, 1);
      acc.re += __shfl_xor(acc.re, 2);
      acc.im += __shfl_xor(acc.im, 2);
      const int64_t x = x0 + s;
      if (rr == 0 && op_lane && x < V3)
        gstore(out + ctile_idx(col0 + kop, x, ncols), cmake<T>(sg * acc.re, sg * acc.im));
    }
  }
}

template <typename T>
int launch_udsc(const UdscSched &S, const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out, int64_t V3,
                int T_lat, int nch_total, int ch0, int spw, hipStream_t st) {
  // spw: sites per workgroup, rounded to whole groups
  constexpr int G = UdscGeom<T>::G;
  const int gpb = std::max(1, spw / G);
  const int64_t ngroups = (V3 + G - 1) / G;
  dim3 block(256), grid((unsigned)((ngroups + gpb - 1) / gpb), (unsigned)nrows);
  k_udsc<T><<<grid, block, 0, st>>>(S, props, rows, out, V3, T_lat, nch_total, ch0, gpb);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

template int launch_udsc<double>(const UdscSched &, const PropDev<double> *, const RowDev *, int, cx<double> *,
                                 int64_t, int, int, int, int, hipStream_t);
template int launch_udsc<float>(const UdscSched &, const PropDev<float> *, const RowDev *, int, cx<float> *,
                                int64_t, int, int, int, int, hipStream_t);

}  // namespace plg
