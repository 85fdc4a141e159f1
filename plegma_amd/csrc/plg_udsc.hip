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
  // quads with the most distinct terms first: a wave runs the term loop of a
  // chunk as often as its fullest quad needs, so the later chunks run fewer
  std::sort(quads.begin(), quads.end(), [&](int a, int b) {
    const size_t na = by_quad[a].size(), nb = by_quad[b].size();
    if (na != nb) return na > nb;
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
    h.s.ntp[c] = std::max(h.s.ntp[c], (int)v.size());
    for (size_t e = 0; e < v.size(); e++) {
      h.tp[i * UD_TPQ + e] = (signed char)((v[e] >> 2 & 3) * 4 + (v[e] & 3));
      slot[v[e]] = {c, (int)e * 64 + l};  // term-major: a wave's 64 writes hit distinct banks
    }
  }
  // each operator's entries per chunk: slot | coefficient index << 8
  std::map<std::pair<float, float>, int> cix;
  h.coef = {0.f, 0.f};
  cix[{0.f, 0.f}] = 0;
  h.s.real_coef = 1;
  std::vector<std::vector<std::vector<uint16_t>>> lists(nch, std::vector<std::vector<uint16_t>>(nops));
  for (int k = 0; k < nops; k++)
    for (int i = 0; i < count[k]; i++) {
      const UdscTerm &t = terms[first[k] + i];
      const auto &sl = slot[key6(t)];
      auto it = cix.find({t.re, t.im});
      if (it == cix.end()) {
        if ((int)h.coef.size() / 2 >= UD_MAXCOEF) return PLG_ERR_UNSUPPORTED;
        it = cix.insert({{t.re, t.im}, (int)h.coef.size() / 2}).first;
        h.coef.push_back(t.re);
        h.coef.push_back(t.im);
        if (t.im != 0.f) h.s.real_coef = 0;
      }
      lists[sl.first][k].push_back((uint16_t)(sl.second | (it->second << 8)));
    }
  // lanes per operator: one each, then one at a time to the operator whose
  // extra lane lowers the blocks summed over the chunks the most (a chunk
  // takes as many blocks as its longest lane list)
  auto blocks = [&](const std::vector<int> &L) {
    int tot = 0;
    for (int c = 0; c < nch; c++) {
      int m = 0;
      for (int k = 0; k < nops; k++) {
        const int per = ((int)lists[c][k].size() + L[k] - 1) / L[k];
        m = std::max(m, (per + 3) / 4);
      }
      tot += m;
    }
    return tot;
  };
  std::vector<int> L(nops, 1), tot_ent(nops, 0);
  for (int c = 0; c < nch; c++)
    for (int k = 0; k < nops; k++) tot_ent[k] += (int)lists[c][k].size();
  for (int extra = 0; extra < 64 - nops; extra++) {
    int best_k = 0, best_b = 1 << 30;
    double best_load = -1;
    for (int k = 0; k < nops; k++) {
      L[k]++;
      const int bl = blocks(L);
      const double load = (double)tot_ent[k] / L[k];
      L[k]--;
      if (bl < best_b || (bl == best_b && load > best_load)) {
        best_b = bl;
        best_load = load;
        best_k = k;
      }
    }
    L[best_k]++;
  }
  for (int k = 0, lo = 0; k < nops; k++) {
    h.s.op_lo[k] = lo;
    h.s.op_n[k] = L[k];
    lo += L[k];
  }
  // blocks: [block][lane][4]; operator k's list split into L[k] contiguous
  // shares of near-equal length, one per lane
  int nblk_tot = 0;
  for (int c = 0; c < nch; c++) {
    std::vector<std::vector<uint16_t>> lane(64);
    int nb = 0;
    for (int k = 0; k < nops; k++) {
      const auto &v = lists[c][k];
      const int n = (int)v.size();
      for (int r = 0; r < L[k]; r++) {
        auto &ln = lane[h.s.op_lo[k] + r];
        for (int i = r * n / L[k]; i < (r + 1) * n / L[k]; i++) ln.push_back(v[i]);
        nb = std::max(nb, ((int)ln.size() + 3) / 4);
      }
    }
    h.s.blk0[c] = nblk_tot;
    h.s.nblk[c] = nb;
    nblk_tot += nb;
    if (nblk_tot > UD_MAXBLK) return PLG_ERR_UNSUPPORTED;
    for (int b = 0; b < nb; b++)
      for (int l = 0; l < 64; l++)
        for (int j = 0; j < 4; j++) {
          const size_t e = (size_t)b * 4 + j;
          h.ents.push_back(e < lane[l].size() ? lane[l][e] : (uint16_t)0);
        }
  }
  h.s.nent = (int)h.ents.size();
  h.s.ncoef = (int)h.coef.size() / 2;
  if (h.s.nent > UD_MAXENT || h.s.ncoef > UD_MAXCOEF) return PLG_ERR_UNSUPPORTED;
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
// A workgroup of W waves stages the three propagators of G consecutive sites
// (one 128-byte line of every component) into LDS with coalesced loads; each
// wave then contracts G/W of those sites, one at a time.  The staged sites
// take 55.4 KB of LDS in either precision; FP32 runs 8 waves on them (two
// workgroups of 78 KB fill a CU's 160 KB with 4 waves per SIMD), FP64 4
// waves (its 159 VGPRs allow 3 waves per SIMD, and 8 waves would double
// the term slots to 94 KB, one workgroup per CU).
// LDS written and read by lanes of one wave only: a wavefront-scope fence
// orders the accesses (a wave's LDS operations are performed in order)
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
  __builtin_amdgcn_wave_barrier();
}

template <typename T> struct UdscGeom {
  static constexpr int G = 128 / (int)sizeof(cx<T>);  // sites per line
  static constexpr int SPAD = 433;                    // LDS stride of a site (432 + 1)
  static constexpr int W = sizeof(T) == 4 ? 8 : 4;    // waves per workgroup
  static constexpr int NT = 64 * W;
};

// RC: every coefficient is real (the built-in tables), an entry costs two
// FMAs instead of a complex product
template <typename T, bool RC>
__global__ __launch_bounds__(64 * UdscGeom<T>::W) void k_udsc(UdscSched S, const PropDev<T> *__restrict__ props,
                                              const RowDev *__restrict__ rows, cx<T> *__restrict__ out, int64_t V3,
                                              int T_lat, int nch_total, int ch0, int groups_per_block) {
  constexpr int G = UdscGeom<T>::G, SPAD = UdscGeom<T>::SPAD, W = UdscGeom<T>::W, NT = UdscGeom<T>::NT;
  constexpr int CPP = NT / G;                   // components one pass of the workgroup stages
  constexpr int NPASS = (144 + CPP - 1) / CPP;  // passes over a propagator's 144 components
  __shared__ cx<T> sp[G * SPAD];              // [site][f][m01][c c']
  __shared__ cx<T> sterm[W][64 * UD_TPQ];     // each wave's chunk terms, then its lanes' sums
  __shared__ __align__(16) uint16_t sent[UD_MAXENT];
  __shared__ T scoef[2 * UD_MAXCOEF];         // RC: the real parts only
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
  // the lane's quads and terms in registers; the entries and coefficients in LDS
  int qv[UD_MAXC], tpv[UD_MAXC];
#pragma unroll
  for (int c = 0; c < UD_MAXC; c++) {
    const bool live = c < S.nchunks;
    qv[c] = live ? S.quads[c * 64 + lane] : -1;
    tpv[c] = live ? *(const int *)(S.tp + (c * 64 + lane) * UD_TPQ) : -1;
  }
  for (int i = tid; i < S.nent; i += NT) sent[i] = S.ents[i];
  if (RC)
    for (int i = tid; i < S.ncoef; i += NT) scoef[i] = (T)S.coef[2 * i];
  else
    for (int i = tid; i < 2 * S.ncoef; i += NT) scoef[i] = (T)S.coef[i];
  cx<T> *st = sterm[wave];
  // staging: thread (site s_l, component c_l) loads components c_l + i CPP of
  // the three propagators at site x0 + s_l; the G threads of a component read
  // one 128-byte line
  const int s_l = tid % G, c_l = tid / G;
  const int64_t o0 = (int64_t)c_l * P0.cs + s_l, o1 = (int64_t)c_l * P1.cs + s_l, o2 = (int64_t)c_l * P2.cs + s_l;
  const int64_t d0 = (int64_t)CPP * P0.cs, d1 = (int64_t)CPP * P1.cs, d2 = (int64_t)CPP * P2.cs;
  cx<T> *dst = sp + s_l * SPAD + c_l;

  for (int gi = 0; gi < groups_per_block; gi++) {
    const int64_t x0 = ((int64_t)blockIdx.x * groups_per_block + gi) * G;
    if (x0 >= V3) break;  // uniform over the workgroup
    // all loads of the group in flight at once, issued before the wait for
    // the previous group's sites
    const bool xin = x0 + s_l < V3;
    cx<T> v[3][NPASS];
#pragma unroll
    for (int i = 0; i < NPASS; i++) {
      const bool ok = xin && c_l + i * CPP < 144;
      v[0][i] = ok ? gload(b0 + x0 + o0 + i * d0) : cmake<T>(0, 0);
      v[1][i] = ok ? gload(b1 + x0 + o1 + i * d1) : cmake<T>(0, 0);
      v[2][i] = ok ? gload(b2 + x0 + o2 + i * d2) : cmake<T>(0, 0);
    }
    __syncthreads();  // the previous group's sites are done
#pragma unroll
    for (int i = 0; i < NPASS; i++)
      if (c_l + i * CPP < 144) {
        dst[i * CPP] = v[0][i];
        dst[144 + i * CPP] = v[1][i];
        dst[288 + i * CPP] = v[2][i];
      }
    __syncthreads();
    if (rot_any) {  // rotation to the physical basis, per (site, propagator, colour pair)
      for (int k = tid; k < G * 27; k += NT) {
        const int s = k / 27, f = (k % 27) / 9, cc = k % 9;
        const int rot = f == 0 ? P0.rot : (f == 1 ? P1.rot : P2.rot);
        if (rot) {
          cx<T> *b = sp + s * SPAD + f * 144 + cc;
          cx<T> w[16];
#pragma unroll
          for (int m = 0; m < 16; m++) w[m] = b[m * 9];
          rot_spins(w, rot);
#pragma unroll
          for (int m = 0; m < 16; m++) b[m * 9] = w[m];
        }
      }
      __syncthreads();
    }
    for (int s = wave; s < G; s += W) {
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
            if (e >= S.ntp[c]) break;  // uniform: the chunk's fullest quad is done
            const int p = (signed char)(tpv[c] >> (8 * e));
            if (p >= 0) {
              const cx<T> *C3 = spx + 288 + p * 9;
              cx<T> t = cmake<T>(0, 0);
#pragma unroll
              for (int j = 0; j < 9; j++) cfma(t, d[j], C3[j]);
              st[e * 64 + lane] = t;
            }
          }
        }
        wave_sync();
        // the lane's share of its operator's entries: nblk[c] blocks of four
        const uint16_t *ec = sent + S.blk0[c] * 256 + lane * 4;
        for (int bk = 0; bk < S.nblk[c]; bk++) {
          const uint2 w = *(const uint2 *)(ec + bk * 256);
          const int u[4] = {(int)(w.x & 0xffff), (int)(w.x >> 16), (int)(w.y & 0xffff), (int)(w.y >> 16)};
#pragma unroll
          for (int j = 0; j < 4; j++) {
            const cx<T> t = st[u[j] & 255];
            const int ci = u[j] >> 8;
            if (RC) {
              const T cr = scoef[ci];
              acc.re = fma(cr, t.re, acc.re);
              acc.im = fma(cr, t.im, acc.im);
            } else {
              cfma(acc, cmake<T>(scoef[2 * ci], scoef[2 * ci + 1]), t);
            }
          }
        }
        wave_sync();
      }
      // each operator sums the shares of its lanes
      st[lane] = acc;
      wave_sync();
      const int64_t x = x0 + s;
      if (lane < S.nops && x < V3) {
        const int lo = S.op_lo[lane], n = S.op_n[lane];
        cx<T> sum = cmake<T>(0, 0);
        for (int r = 0; r < n; r++) sum = sum + st[lo + r];
        gstore(out + ctile_idx(col0 + lane, x, ncols), cmake<T>(sg * sum.re, sg * sum.im));
      }
      wave_sync();
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
  dim3 block(UdscGeom<T>::NT), grid((unsigned)((ngroups + gpb - 1) / gpb), (unsigned)nrows);
  if (S.real_coef)
    k_udsc<T, true><<<grid, block, 0, st>>>(S, props, rows, out, V3, T_lat, nch_total, ch0, gpb);
  else
    k_udsc<T, false><<<grid, block, 0, st>>>(S, props, rows, out, V3, T_lat, nch_total, ch0, gpb);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

template int launch_udsc<double>(const UdscSched &, const PropDev<double> *, const RowDev *, int, cx<double> *,
                                 int64_t, int, int, int, int, hipStream_t);
template int launch_udsc<float>(const UdscSched &, const PropDev<float> *, const RowDev *, int, cx<float> *,
                                int64_t, int, int, int, int, hipStream_t);

}  // namespace plg
