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
// The next site's propagator entries are loaded into registers while the
// current site is contracted.
#include "plg_internal.h"
#include "plg_udsc_tables.h"

#include <algorithm>
#include <map>
#include <vector>

namespace plg {

// ------------------------------------------------------------------ schedule
// host: build the chunked schedule of one flavour product from the tables
int udsc_schedule(int prod, UdscSchedHost &h) {
  if (prod < 0 || prod >= UDSC_NPROD) return PLG_ERR_INVALID;
  const UdscProd &P = UDSC_PRODS[prod];
  if (P.nops > UD_MAXOPS) return PLG_ERR_UNSUPPORTED;
  auto key6 = [](const UdscTerm &t) {
    int k = 0;
    for (int i = 0; i < 6; i++) k = k * 4 + t.m[i];
    return k;
  };
  auto quad_of = [](int k6) {  // m01 | m23 << 4
    const int m0 = (k6 >> 10) & 3, m1 = (k6 >> 8) & 3, m2 = (k6 >> 6) & 3, m3 = (k6 >> 4) & 3;
    return (m0 * 4 + m1) | ((m2 * 4 + m3) << 4);
  };
  // distinct terms grouped by quad
  std::map<int, std::vector<int>> by_quad;  // quad -> distinct key6 terms
  for (int k = 0; k < P.nops; k++) {
    const UdscOp &o = UDSC_OPS[P.first_op + k];
    for (int i = 0; i < o.count; i++) {
      const int k6 = key6(UDSC_TERMS[o.first + i]);
      auto &v = by_quad[quad_of(k6)];
      if (std::find(v.begin(), v.end(), k6) == v.end()) v.push_back(k6);
    }
  }
  std::vector<int> quads;
  for (auto &kv : by_quad) {
    if ((int)kv.second.size() > UD_TPQ) return PLG_ERR_UNSUPPORTED;
    quads.push_back(kv.first);
  }
  std::sort(quads.begin(), quads.end(), [](int a, int b) {
    return (a & 15) != (b & 15) ? (a & 15) < (b & 15) : (a >> 4) < (b >> 4);
  });
  const int nch = ((int)quads.size() + 63) / 64;
  if (nch > UD_MAXC) return PLG_ERR_UNSUPPORTED;
  std::map<int, std::pair<int, int>> slot;  // key6 -> (chunk, chunk-local slot)
  h = UdscSchedHost{};
  h.s.nops = P.nops;
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
    for (int k = 0; k < P.nops; k++) {
      std::vector<uint16_t> v;
      const UdscOp &o = UDSC_OPS[P.first_op + k];
      for (int i = 0; i < o.count; i++) {
        const UdscTerm &t = UDSC_TERMS[o.first + i];
        const auto &sl = slot[key6(t)];
        if (sl.first != c) continue;
        auto it = cix.find({t.re, t.im});
        if (it == cix.end()) {
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

int udsc_products() { return UDSC_NPROD; }
const char *udsc_flavours(int prod) { return UDSC_PRODS[prod].flavours; }
int udsc_nops(int prod) { return UDSC_PRODS[prod].nops; }
const char *udsc_op_name(int prod, int k) { return UDSC_OPS[UDSC_PRODS[prod].first_op + k].name; }

// ------------------------------------------------------------------ kernel
// one wave per workgroup; the wave walks `spw` consecutive sites of one row
template <typename T>
__global__ __launch_bounds__(64) void k_udsc(UdscSched S, const PropDev<T> *__restrict__ props,
                                             const RowDev *__restrict__ rows, cx<T> *__restrict__ out, int64_t V3,
                                             int T_lat, int nch_total, int ch0, int spw) {
  __shared__ cx<T> sp[3 * 144];           // the site's propagators [f][m01][c c']
  __shared__ cx<T> sterm[64 * UD_TPQ];    // the chunk's distinct terms
  __shared__ __align__(16) uint16_t sent[UD_MAXENT];    // operator entries
  __shared__ float scoef[2 * UD_MAXCOEF];
  const int lane = threadIdx.x;
  const RowDev row = rows[blockIdx.y];
  const PropDev<T> P0 = props[row.pa], P1 = props[row.pb], P2 = props[row.pc];
  auto plane_base = [&](const PropDev<T> &Q) {
    int plane = row.t - Q.t_first;
    plane += plane < 0 ? T_lat : 0;
    return Q.data + (int64_t)plane * V3;
  };
  const T sg = prop_sign(P0, row.t) * prop_sign(P1, row.t) * prop_sign(P2, row.t);
  const int rot_any = P0.rot | P1.rot | P2.rot;
  const int64_t ncols = (int64_t)gridDim.y * nch_total, col0 = (int64_t)blockIdx.y * nch_total + ch0;
  const int kop = lane >> 2, rr = lane & 3;
  const bool op_lane = kop < S.nops;
  // the lane's schedule in registers; the entries and coefficients in LDS:
  // no memory loads but the propagators inside the site loop
  int qv[UD_MAXC], tpv[UD_MAXC], eoff[UD_MAXC], en[UD_MAXC];
#pragma unroll
  for (int c = 0; c < UD_MAXC; c++) {
    const bool live = c < S.nchunks;
    qv[c] = live ? S.quads[c * 64 + lane] : -1;
    tpv[c] = live ? *(const int *)(S.tp + (c * 64 + lane) * UD_TPQ) : -1;
    en[c] = live && op_lane ? S.ent_len[c][kop] : 0;
    eoff[c] = live && op_lane ? S.ent_off[c][kop] + rr * en[c] : 0;
  }
  for (int i = lane; i < S.nent; i += 64) sent[i] = S.ents[i];
  for (int i = lane; i < 2 * S.ncoef; i += 64) scoef[i] = S.coef[i];
  // this lane's 7 staging entries e = lane + 64 k: propagator f, component comp
  const cx<T> *src[7];
#pragma unroll
  for (int k = 0; k < 7; k++) {
    const int e = min(lane + 64 * k, 431);
    const int f = e / 144, comp = e - f * 144;
    const PropDev<T> &Q = f == 0 ? P0 : (f == 1 ? P1 : P2);
    src[k] = plane_base(Q) + (int64_t)comp * Q.cs;
  }
  const int64_t x0 = (int64_t)blockIdx.x * spw;
  const int nsite = (int)min((int64_t)spw, V3 - x0);
  cx<T> pre[7];
#pragma unroll
  for (int k = 0; k < 7; k++) pre[k] = gload(src[k] + x0);

  for (int i = 0; i < nsite; i++) {
    const int64_t x = x0 + i;
#pragma unroll
    for (int k = 0; k < 7; k++)
      if (k < 6 || lane < 48) sp[lane + 64 * k] = pre[k];
    // the next site's entries, in flight during this site's contraction
    if (i + 1 < nsite) {
#pragma unroll
      for (int k = 0; k < 7; k++) pre[k] = gload(src[k] + x + 1);
    }
    __syncthreads();
    if (rot_any) {
      if (lane < 27) {  // rotation to the physical basis, per (propagator, colour pair)
        const int f = lane / 9, cc = lane % 9;
        const int rot = f == 0 ? P0.rot : (f == 1 ? P1.rot : P2.rot);
        if (rot) {
          cx<T> s[16];
#pragma unroll
          for (int m = 0; m < 16; m++) s[m] = sp[f * 144 + m * 9 + cc];
          rot_spins(s, rot);
#pragma unroll
          for (int m = 0; m < 16; m++) sp[f * 144 + m * 9 + cc] = s[m];
        }
      }
      __syncthreads();
    }
    cx<T> acc = cmake<T>(0, 0);
#pragma unroll
    for (int c = 0; c < UD_MAXC; c++) {
      if (c >= S.nchunks) break;
      const int q = qv[c];
      if (q >= 0) {
        // D of this lane's quad, in registers
        const cx<T> *A = sp + (q & 15) * 9, *B = sp + 144 + (q >> 4) * 9;
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
            const int b1 = (dd + 1) % 3, b2 = (dd + 2) % 3;
            cx<T> v = a[a1 * 3 + b1] * b[a2 * 3 + b2];
            v = v - a[a1 * 3 + b2] * b[a2 * 3 + b1];
            v = v - a[a2 * 3 + b1] * b[a1 * 3 + b2];
            cfma(v, a[a2 * 3 + b2], b[a1 * 3 + b1]);
            d[cc * 3 + dd] = v;
          }
        }
        // its distinct terms T = sum_cc' D[c][c'] S3[m4m5][c][c']
#pragma unroll
        for (int e = 0; e < UD_TPQ; e++) {
          const int p = (signed char)(tpv[c] >> (8 * e));
          if (p >= 0) {
            const cx<T> *C3 = sp + 288 + p * 9;
            cx<T> v = cmake<T>(0, 0);
#pragma unroll
            for (int j = 0; j < 9; j++) cfma(v, d[j], C3[j]);
            sterm[lane * UD_TPQ + e] = v;
          }
        }
      }
      __syncthreads();
      // operator entries: 4 lanes per operator, each a quarter of the list,
      // four entries per LDS read
      for (int e = 0; e < en[c]; e += 4) {
        const uint2 w = *(const uint2 *)(sent + eoff[c] + e);
        const int u[4] = {(int)(w.x & 0xffff), (int)(w.x >> 16), (int)(w.y & 0xffff), (int)(w.y >> 16)};
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int ci = u[j] >> 8;
          cfma(acc, cmake<T>((T)scoef[2 * ci], (T)scoef[2 * ci + 1]), sterm[u[j] & 255]);
        }
      }
      __syncthreads();
    }
    acc.re += __shfl_xor(acc.re, 1);
    acc.im += __shfl_xor(acc.im, 1);
    acc.re += __shfl_xor(acc.re, 2);
    acc.im += __shfl_xor(acc.im, 2);
    if (rr == 0 && op_lane) gstore(out + ctile_idx(col0 + kop, x, ncols), cmake<T>(sg * acc.re, sg * acc.im));
  }
}

template <typename T>
int launch_udsc(const UdscSched &S, const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out, int64_t V3,
                int T_lat, int nch_total, int ch0, int spw, hipStream_t st) {
  spw = std::max(1, spw);
  dim3 block(64), grid((unsigned)((V3 + spw - 1) / spw), (unsigned)nrows);
  k_udsc<T><<<grid, block, 0, st>>>(S, props, rows, out, V3, T_lat, nch_total, ch0, spw);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

template int launch_udsc<double>(const UdscSched &, const PropDev<double> *, const RowDev *, int, cx<double> *,
                                 int64_t, int, int, int, int, hipStream_t);
template int launch_udsc<float>(const UdscSched &, const PropDev<float> *, const RowDev *, int, cx<float> *,
                                int64_t, int, int, int, int, hipStream_t);

}  // namespace plg
