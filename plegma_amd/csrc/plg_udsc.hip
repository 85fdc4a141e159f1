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
// Quads are processed in chunks of 64 (one per lane), then the chunk's
// distinct terms (lane-parallel), then each operator's entries (4 lanes per
// operator).  D costs 36 complex products per quad, a term 9, an entry 1.
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
  // distinct terms and quads
  std::map<int, int> term_ix;  // key6 -> distinct index
  std::vector<int> terms;
  for (int k = 0; k < P.nops; k++) {
    const UdscOp &o = UDSC_OPS[P.first_op + k];
    for (int i = 0; i < o.count; i++) {
      const int k6 = key6(UDSC_TERMS[o.first + i]);
      if (!term_ix.count(k6)) {
        term_ix[k6] = 0;
        terms.push_back(k6);
      }
    }
  }
  std::vector<int> quads;
  for (int k6 : terms) quads.push_back(quad_of(k6));
  std::sort(quads.begin(), quads.end(), [](int a, int b) {
    return (a & 15) != (b & 15) ? (a & 15) < (b & 15) : (a >> 4) < (b >> 4);
  });
  quads.erase(std::unique(quads.begin(), quads.end()), quads.end());
  const int nch = ((int)quads.size() + 63) / 64;
  if (nch > UD_MAXC) return PLG_ERR_UNSUPPORTED;
  std::map<int, std::pair<int, int>> qpos;  // quad -> (chunk, lane)
  for (size_t i = 0; i < quads.size(); i++) qpos[quads[i]] = {(int)i / 64, (int)i % 64};
  h = UdscSchedHost{};
  h.s.nops = P.nops;
  h.s.nchunks = nch;
  for (int c = 0; c < nch; c++) {
    h.s.quad_off[c] = (int)h.quads.size();
    for (int l = 0; l < 64; l++) {
      const size_t i = (size_t)c * 64 + l;
      h.quads.push_back(i < quads.size() ? (short)quads[i] : (short)-1);
    }
    // the chunk's distinct terms, padded to whole waves
    h.s.term_off[c] = (int)h.terms.size();
    std::map<int, int> local;
    for (int k6 : terms) {
      const int q = quad_of(k6);
      if (qpos[q].first != c) continue;
      local[k6] = (int)(h.terms.size() - h.s.term_off[c]);
      const int p = (k6 >> 2 & 3) * 4 + (k6 & 3);
      h.terms.push_back((short)(qpos[q].second | (p << 6)));
    }
    const int n = (int)h.terms.size() - h.s.term_off[c];
    if (n > UD_MAXT) return PLG_ERR_UNSUPPORTED;
    while ((h.terms.size() - h.s.term_off[c]) % 64) h.terms.push_back((short)-1);
    h.s.term_n[c] = (int)h.terms.size() - h.s.term_off[c];
    // each operator's entries whose term lies in this chunk
    for (int k = 0; k < P.nops; k++) {
      h.s.ent_off[c][k] = (int)h.ents.size();
      const UdscOp &o = UDSC_OPS[P.first_op + k];
      for (int i = 0; i < o.count; i++) {
        const UdscTerm &t = UDSC_TERMS[o.first + i];
        auto it = local.find(key6(t));
        if (it != local.end()) h.ents.push_back(UdscEntry{it->second, t.re, t.im});
      }
      h.s.ent_n[c][k] = (int)h.ents.size() - h.s.ent_off[c][k];
    }
  }
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
  __shared__ cx<T> sp[3 * 144];      // the site's propagators [f][m01][c c']
  __shared__ cx<T> sd[64 * 9];       // D of the chunk's quads
  __shared__ cx<T> sterm[UD_MAXT];   // T of the chunk's terms
  const int lane = threadIdx.x;
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

  for (int i = 0; i < spw; i++) {
    const int64_t x = (int64_t)blockIdx.x * spw + i;
    if (x >= V3) break;  // uniform over the wave
    // stage the three propagators of site x: entry e = f*144 + (m01*9 + c*3 + c')
#pragma unroll
    for (int k = 0; k < 7; k++) {
      const int e = lane + 64 * k;
      if (e < 432) {
        const int f = e / 144, comp = e - f * 144;
        const cx<T> *b = f == 0 ? b0 : (f == 1 ? b1 : b2);
        const int64_t cs = f == 0 ? P0.cs : (f == 1 ? P1.cs : P2.cs);
        sp[e] = gload(b + (int64_t)comp * cs + x);
      }
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
    for (int c = 0; c < S.nchunks; c++) {
      // D of this lane's quad
      const int q = S.quads[S.quad_off[c] + lane];
      if (q >= 0) {
        const cx<T> *A = sp + (q & 15) * 9, *B = sp + 144 + (q >> 4) * 9;
        cx<T> a[9], b[9];
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
            cx<T> d = a[a1 * 3 + b1] * b[a2 * 3 + b2];
            d = d - a[a1 * 3 + b2] * b[a2 * 3 + b1];
            d = d - a[a2 * 3 + b1] * b[a1 * 3 + b2];
            cfma(d, a[a2 * 3 + b2], b[a1 * 3 + b1]);
            sd[lane * 9 + cc * 3 + dd] = d;
          }
        }
      }
      __syncthreads();
      // the chunk's distinct terms
      for (int j = lane; j < S.term_n[c]; j += 64) {
        const int tt = S.terms[S.term_off[c] + j];
        if (tt >= 0) {
          const cx<T> *D = sd + (tt & 63) * 9, *C3 = sp + 288 + (tt >> 6) * 9;
          cx<T> v = cmake<T>(0, 0);
#pragma unroll
          for (int e = 0; e < 9; e++) cfma(v, D[e], C3[e]);
          sterm[j] = v;
        }
      }
      __syncthreads();
      // operator entries: 4 lanes per operator
      if (kop < S.nops) {
        const int n = S.ent_n[c][kop];
        const UdscEntry *E = S.ents + S.ent_off[c][kop];
        for (int e = rr; e < n; e += 4) {
          const UdscEntry u = E[e];
          cfma(acc, cmake<T>((T)u.re, (T)u.im), sterm[u.t]);
        }
      }
      __syncthreads();
    }
    acc.re += __shfl_xor(acc.re, 1);
    acc.im += __shfl_xor(acc.im, 1);
    acc.re += __shfl_xor(acc.re, 2);
    acc.im += __shfl_xor(acc.im, 2);
    if (rr == 0 && kop < S.nops) gstore(out + ctile_idx(col0 + kop, x, ncols), cmake<T>(sg * acc.re, sg * acc.im));
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
