// plg_project_sep.hip -- momentum projection as a separable DFT (gfx950).
//
// The projection of a site-correlator column c(x, y, z) on a momentum
// n = (n_x, n_y, n_z) (lib/kernels/PLEGMA_kernel_utils.cuh:820-860, sign -1)
//   F(n) = sum_{x,y,z} exp(-i 2 pi (n_x x / Lx + n_y y / Ly + n_z z / Lz)) c(x, y, z)
// factorises over the directions.  A createMom list (|n|^2 <= Q) has few
// distinct components, so the sums run direction by direction:
//   G[n_x](y, z) = sum_x exp(-i 2 pi n_x x / Lx) c          (per line)
//   H[n_x, n_y](z) = sum_y exp(-i 2 pi n_y y / Ly) G[n_x]   (per plane)
//   F(n) = sum_z exp(-i 2 pi n_z z / Lz) H[n_x, n_y]        (per column)
// For 257 momenta (|n|^2 <= 16) on 32^3 that is about 25 real FMAs per site
// and column instead of 8 x 129 for the paired phase-matrix GEMM, all in FP64
// (the FP32 site correlator is converted exactly), so the pass is bound by
// reading the site correlator from HBM.
//
//  k_sep_xy  workgroup = 16 columns x one z plane.  The plane is staged in
//            chunks of CHL lines (256 FP32 or 128 FP64 sites per column) in
//            LDS; each line is summed by NPART threads (NPART * CHL = 16),
//            each over W = Lx / NPART sites with the phases of its offset
//            i < W read as uniform (scalar) loads, rotated by the part's
//            phase and combined by DPP row exchanges (the line's parts are
//            neighbouring lanes).  The cos and sin sums of
//            a = |n_x| give both n_x = a and n_x = -a.  Thread (column,
//            n_x residue) then accumulates the plane's cos and sin sums over y
//            for every |n_y| class; the (n_x, n_y) pairs the momenta use are
//            written.
//  k_sep_z   thread = column, workgroup row = (n_x, n_y) pair: the sums over
//            z of the pair's momenta, written momentum-major for the
//            finalize kernel.
//
// LDS strides: a chunk element (column c, line l, part p, offset i) sits at
// c RS + l LS + p PS + i; PS = 1 and RS = NPART modulo the lanes of one LDS
// pass (16 for 8-byte reads, 8 for 16-byte reads), so the lanes of a pass
// (consecutive parts, then columns) read distinct banks.
#include "plg_internal.h"
#include "plg_tile.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>

namespace plg {

template <typename T> struct SepT;
template <> struct SepT<float> {
  static constexpr int SCH = 256, UPC = 8, EPU = 2, PASS = 16, ELEMS = 5120;
};
template <> struct SepT<double> {
  static constexpr int SCH = 128, UPC = 16, EPU = 1, PASS = 8, ELEMS = 2560;
};
constexpr int SEP_LDS = 40960;  // bytes: SepT<T>::ELEMS elements of cx<T>

template <typename T, int EPU> struct SepVec;
template <> struct SepVec<float, 2> {
  typedef float v __attribute__((ext_vector_type(4)));
};
template <> struct SepVec<double, 1> {
  typedef double v __attribute__((ext_vector_type(2)));
};

// v += the value of the lanes whose index differs in the low log2(N) bits
template <int N> __device__ __forceinline__ void sep_rowsum(double &v) {
  if (N > 1) v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  if (N > 2) v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  if (N > 4) v += dpp<0x141>(v);  // row_half_mirror
  if (N > 8) v += dpp<0x140>(v);  // row_mirror
}

// small tables: three waves per SIMD
template <typename T, int NPART, int MAXA, int MAXB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(MAXA <= 5 ? 3 : 1))) void k_sep_xy(SepGeom s, const cx<T> *__restrict__ c,
                                                cx<double> *__restrict__ H) {
  using Q = SepT<T>;
  typedef typename SepVec<T, Q::EPU>::v vec;
  constexpr int CHL = 16 / NPART;                  // lines per chunk
  constexpr int KU = Q::SCH * Q::UPC / 256;        // 16-byte units per thread and chunk
  __shared__ __align__(16) unsigned char lds[SEP_LDS];
  cx<T> *st = (cx<T> *)lds;              // staged chunk
  cx<double> *gb = (cx<double> *)lds;    // line sums G[n_x] of the chunk (after the staged data is used)
  const int tid = threadIdx.x;
  const int64_t col0 = (int64_t)blockIdx.x * 16;
  const int z = blockIdx.y;
  const int Lx = s.Lx, SCHa = CHL * Lx, NU = SCHa * Q::UPC;

  // staging role: the same column and LDS offsets in every chunk
  const int sc = (tid / Q::UPC) & 15;
  const bool col_ok = col0 + sc < s.ncols;
  uint32_t loff[KU];  // LDS offsets (< 2^16) of the unit's elements, 16 bits each
#pragma unroll
  for (int k = 0; k < KU; k++) {
    const int u = tid + 256 * k, tt = u / (16 * Q::UPC), j = u % Q::UPC;
    loff[k] = 0;
#pragma unroll
    for (int e = 0; e < Q::EPU; e++) {
      const int site = tt * 16 + j * Q::EPU + e, l = site / Lx, x = site - l * Lx, p = x / s.W;
      loff[k] |= (uint32_t)(sc * s.RS + l * s.LS + p * s.PS + (x - p * s.W)) << (16 * e);
    }
  }
  // line role: part p1 of line l1 of column c1
  const int p1 = tid % NPART, c1 = (tid / NPART) & 15, l1 = tid / (NPART * 16);
  const int base1 = c1 * s.RS + l1 * s.LS + p1 * s.PS;
  // plane role: column c2, n_x residue slot j2; cos and sin sums over y for
  // the |n_y| classes b
  const int c2 = tid & 15, j2 = tid >> 4;
  const bool role2 = j2 < s.NXR;
  cx<double> hc[MAXB], hs[MAXB];
#pragma unroll
  for (int b = 0; b < MAXB; b++) hc[b] = hs[b] = cmake<double>(0, 0);

  const int64_t tile_plane = (int64_t)z * (Lx * s.Ly) / 16;
  const int nchunk = s.Ly / CHL;
  // the chunk's 16-byte units (loading the next chunk during the sums holds
  // 32 more registers: two waves per SIMD instead of three, measured slower)
  vec v[KU];
  auto fetch = [&](int q) {
    const int64_t tile0 = tile_plane + (int64_t)q * (SCHa / 16);
#pragma unroll
    for (int k = 0; k < KU; k++) {
      const int u = tid + 256 * k, tt = u / (16 * Q::UPC), j = u % Q::UPC;
      if (u < NU && col_ok)
        v[k] = *(const __attribute__((address_space(1))) vec *)(c + (tile0 + tt) * s.gstride +
                                                               (col0 + sc) * s.cstride + j * Q::EPU);
      else
        v[k] = vec{};
    }
  };
  for (int q = 0; q < nchunk; q++) {
    fetch(q);
    __syncthreads();  // the previous chunk's sums (which share the LDS) are read
#pragma unroll
    for (int k = 0; k < KU; k++) {
      const int u = tid + 256 * k;
      if (u < NU) {
#pragma unroll
        for (int e = 0; e < Q::EPU; e++) st[(loff[k] >> (16 * e)) & 0xffff] = cmake<T>(v[k][2 * e], v[k][2 * e + 1]);
      }
    }
    __syncthreads();
    // sums over the part's sites: C_a = sum cos(2 pi a i / Lx) c, S_a likewise
    cx<double> C[MAXA], S[MAXA];
#pragma unroll
    for (int a = 0; a < MAXA; a++) C[a] = S[a] = cmake<double>(0, 0);
#pragma unroll 2
    for (int i = 0; i < s.W; i++) {
      const cx<T> w = st[base1 + i];
      const double wr = (double)w.re, wi = (double)w.im;
#pragma unroll
      for (int a = 0; a < MAXA; a++) {  // (classes a >= NA have zero phases)
        const cx<double> ph = s.ph1[i * MAXA + a];
        C[a].re += ph.re * wr;
        C[a].im += ph.re * wi;
        S[a].re += ph.im * wr;
        S[a].im += ph.im * wi;
      }
    }
    if (NPART > 1) {
#pragma unroll
      for (int a = 0; a < MAXA; a++) {
        // offset p W: cos(t + u) = cos t cos u - sin t sin u, sin(t + u) = sin t cos u + cos t sin u
        const cx<double> pp = s.ph1p[p1 * MAXA + a];
        const cx<double> c0 = C[a], s0 = S[a];
        C[a] = cmake<double>(pp.re * c0.re - pp.im * s0.re, pp.re * c0.im - pp.im * s0.im);
        S[a] = cmake<double>(pp.im * c0.re + pp.re * s0.re, pp.im * c0.im + pp.re * s0.im);
      }
      // sums over the line's NPART lanes (the low lane bits): quad
      // exchanges, then the other quad, then the other half row
#pragma unroll
      for (int a = 0; a < MAXA; a++) {
        sep_rowsum<NPART>(C[a].re);
        sep_rowsum<NPART>(C[a].im);
        sep_rowsum<NPART>(S[a].re);
        sep_rowsum<NPART>(S[a].im);
      }
    }
    __syncthreads();  // every thread has read its staged sites
    // G[n_x] of the line: a = n_x gives C - i S, a = -n_x (mod Lx) gives C + i S
#pragma unroll
    for (int a = 0; a < MAXA; a++)
      if (a % NPART == p1) {
        const int sp = s.xs[2 * a], sm = s.xs[2 * a + 1];
        if (sp >= 0) gb[(l1 * s.NXR + sp) * 16 + c1] = cmake<double>(C[a].re + S[a].im, C[a].im - S[a].re);
        if (sm >= 0) gb[(l1 * s.NXR + sm) * 16 + c1] = cmake<double>(C[a].re - S[a].im, C[a].im + S[a].re);
      }
    __syncthreads();
    if (role2) {
#pragma unroll 2
      for (int l = 0; l < CHL; l++) {
        const int y = q * CHL + l;
        const cx<double> g = gb[(l * s.NXR + j2) * 16 + c2];
#pragma unroll
        for (int b = 0; b < MAXB; b++) {  // (classes b >= NB have zero phases)
          const cx<double> ph = s.ph2[y * MAXB + b];
          hc[b].re += ph.re * g.re;
          hc[b].im += ph.re * g.im;
          hs[b].re += ph.im * g.re;
          hs[b].im += ph.im * g.im;
        }
      }
    }
  }
  if (role2 && col0 + c2 < s.ncols) {
    // n_y = b: C - i S; n_y = -b (mod Ly): C + i S
#pragma unroll
    for (int b = 0; b < MAXB; b++) {
      const int sp = s.slot[(j2 * MAXB + b) * 2], sm = s.slot[(j2 * MAXB + b) * 2 + 1];
      if (sp >= 0)
        gstore(H + ((int64_t)z * s.NS + sp) * s.ncols + col0 + c2,
               cmake<double>(hc[b].re + hs[b].im, hc[b].im - hs[b].re));
      if (sm >= 0)
        gstore(H + ((int64_t)z * s.NS + sm) * s.ncols + col0 + c2,
               cmake<double>(hc[b].re - hs[b].im, hc[b].im + hs[b].re));
    }
  }
}

template <int MAXZ>
__global__ __launch_bounds__(256) void k_sep_z(SepGeom s, const cx<double> *__restrict__ H,
                                               cx<double> *__restrict__ part) {
  const int64_t col = (int64_t)blockIdx.x * 256 + threadIdx.x;
  const int sl = blockIdx.y;
  if (col >= s.ncols) return;
  // the pair's momenta, padded to MAXZ (padding: n_z residue 0, momentum -1)
  const int *rz = s.sl_rz + sl * MAXZ, *mm = s.sl_m + sl * MAXZ;
  cx<double> acc[MAXZ];
#pragma unroll
  for (int j = 0; j < MAXZ; j++) acc[j] = cmake<double>(0, 0);
  for (int z = 0; z < s.Lz; z++) {
    const cx<double> h = gload(H + ((int64_t)z * s.NS + sl) * s.ncols + col);
#pragma unroll
    for (int j = 0; j < MAXZ; j++) {  // (cos - i sin) h
      const cx<double> ph = s.ph3[rz[j] * s.Lz + z];
      acc[j].re += ph.re * h.re;
      acc[j].re += ph.im * h.im;
      acc[j].im += ph.re * h.im;
      acc[j].im -= ph.im * h.re;
    }
  }
#pragma unroll
  for (int j = 0; j < MAXZ; j++)
    if (mm[j] >= 0) gstore(part + (int64_t)mm[j] * s.ncols + col, acc[j]);
}

// ------------------------------------------------------------------ host
void sep_build(const int L[3], int M, const std::vector<int> &ms, int64_t ncols, bool f64, SepHost &h) {
  h = SepHost{};
  const int nmom = (int)ms.size() / 3;
  const int Lx = L[0], Ly = L[1], Lz = L[2];
  if (nmom <= 0 || (Lx * Ly) % 16 || ncols <= 0) return;
  const int SCH = f64 ? SepT<double>::SCH : SepT<float>::SCH;
  const int PASS = f64 ? SepT<double>::PASS : SepT<float>::PASS;
  const int ELEMS = f64 ? SepT<double>::ELEMS : SepT<float>::ELEMS;
  // residues of the components (ms holds n_k (M / L_k) mod M)
  std::vector<int> rx(nmom), ry(nmom), rz(nmom);
  for (int m = 0; m < nmom; m++) {
    rx[m] = ms[3 * m] / (M / Lx);
    ry[m] = ms[3 * m + 1] / (M / Ly);
    rz[m] = ms[3 * m + 2] / (M / Lz);
  }
  std::map<int, int> amap, rxmap, rymap, bmap;
  for (int m = 0; m < nmom; m++) {
    amap[std::min(rx[m], Lx - rx[m])] = 0;
    rxmap[rx[m]] = 0;
    rymap[ry[m]] = 0;
    bmap[std::min(ry[m], Ly - ry[m])] = 0;
  }
  int k = 0;
  for (auto &e : amap) e.second = k++;
  k = 0;
  for (auto &e : rxmap) e.second = k++;
  k = 0;
  for (auto &e : rymap) e.second = k++;
  k = 0;
  for (auto &e : bmap) e.second = k++;
  SepGeom &g = h.g;
  g.Lx = Lx;
  g.Ly = Ly;
  g.Lz = Lz;
  g.ncols = ncols;
  g.nmom = nmom;
  g.NA = (int)amap.size();
  g.NXR = (int)rxmap.size();
  g.NYR = (int)rymap.size();
  g.NB = (int)bmap.size();
  if (g.NA > 8 || g.NXR > 16 || g.NB > 8) return;
  // chunk: CHL lines (a power of two dividing Ly and 16), whole tiles, the
  // most lines whose staged sites and line sums fit the LDS
  auto fits = [&](int cl) {
    if (Ly % cl || cl * Lx > SCH || (cl * Lx) % 16 || Lx % (16 / cl)) return false;
    g.NPART = 16 / cl;
    g.W = Lx / g.NPART;
    g.PS = g.W + (((1 - g.W) % PASS) + PASS) % PASS;  // PS = 1 mod PASS
    g.LS = g.NPART * g.PS;
    g.RS = cl * g.LS;
    if (g.NPART < PASS)
      while (g.RS % PASS != g.NPART % PASS) g.RS++;
    return 16 * g.RS <= ELEMS && cl * g.NXR * 16 * (int)sizeof(cx<double>) <= SEP_LDS;
  };
  int CHL = 0;
  for (int cl = 16; cl >= 1 && !CHL; cl >>= 1)
    if (fits(cl)) CHL = cl;
  if (!CHL) return;
  // (n_x, n_y) pairs and their momenta
  std::map<std::pair<int, int>, std::vector<int>> pairs;
  for (int m = 0; m < nmom; m++) pairs[{rxmap[rx[m]], rymap[ry[m]]}].push_back(m);
  g.NS = (int)pairs.size();
  int maxcnt = 0;
  for (auto &e : pairs) maxcnt = std::max(maxcnt, (int)e.second.size());
  if (maxcnt > 16) return;
  g.small = g.NA <= 5 && g.NB <= 5 && maxcnt <= 9;
  // tables padded to the kernel's class and momentum counts
  const int NAP = g.small ? 5 : 8, NBP = g.small ? 5 : 8, NZP = g.small ? 9 : 16;
  std::vector<int> ryres(g.NYR);
  for (auto &e : rymap) ryres[e.second] = e.first;
  std::vector<int> slot(g.NXR * NBP * 2, -1), slm(g.NS * NZP, -1), slrz(g.NS * NZP, 0);
  {
    int sidx = 0;
    for (auto &e : pairs) {
      // the pair's n_y residue as class b, sign + (r = b) or - (r = Ly - b != b)
      const int r = ryres[e.first.second], b = std::min(r, Ly - r);
      slot[(e.first.first * NBP + bmap[b]) * 2 + (r == b ? 0 : 1)] = sidx;
      for (size_t j = 0; j < e.second.size(); j++) {
        slm[sidx * NZP + j] = e.second[j];
        slrz[sidx * NZP + j] = rz[e.second[j]];
      }
      sidx++;
    }
  }
  // residue slots of +a and -a
  std::vector<int> xs(2 * NAP, -1);
  for (auto &e : amap) {
    const int a = e.first;
    auto ip = rxmap.find(a);
    if (ip != rxmap.end()) xs[2 * e.second] = ip->second;
    auto im = rxmap.find((Lx - a) % Lx);
    if (a != 0 && 2 * a != Lx && im != rxmap.end()) xs[2 * e.second + 1] = im->second;
  }
  // phases from exact integer residues
  auto phase = [](int64_t num, int den) {
    const double t = 2.0 * M_PI * (double)(num % den) / (double)den;
    return std::make_pair(std::cos(t), std::sin(t));
  };
  std::vector<double> ph1(2 * g.W * NAP, 0.0), ph1p(2 * g.NPART * NAP, 0.0), ph2(2 * Ly * NBP, 0.0), ph3;
  for (auto &e : amap)
    for (int i = 0; i < g.W; i++) {
      auto cs = phase((int64_t)e.first * i, Lx);
      ph1[2 * (i * NAP + e.second)] = cs.first;
      ph1[2 * (i * NAP + e.second) + 1] = cs.second;
    }
  for (auto &e : amap)
    for (int p = 0; p < g.NPART; p++) {
      auto cs = phase((int64_t)e.first * p * g.W, Lx);
      ph1p[2 * (p * NAP + e.second)] = cs.first;
      ph1p[2 * (p * NAP + e.second) + 1] = cs.second;
    }
  for (auto &e : bmap)
    for (int y = 0; y < Ly; y++) {
      auto cs = phase((int64_t)e.first * y, Ly);
      ph2[2 * (y * NBP + e.second)] = cs.first;
      ph2[2 * (y * NBP + e.second) + 1] = cs.second;
    }
  for (int r = 0; r < Lz; r++)
    for (int z = 0; z < Lz; z++) {
      auto cs = phase((int64_t)r * z, Lz);
      ph3.push_back(cs.first);
      ph3.push_back(cs.second);
    }
  // one blob, 256-byte aligned pieces
  auto put = [&](const void *src, size_t bytes) {
    const size_t off = (h.blob.size() + 255) / 256 * 256;
    h.blob.resize(off + std::max<size_t>(bytes, 1));
    if (bytes) std::memcpy(h.blob.data() + off, src, bytes);
    return off;
  };
  h.off_ph1 = put(ph1.data(), ph1.size() * sizeof(double));
  h.off_ph1p = put(ph1p.data(), ph1p.size() * sizeof(double));
  h.off_ph2 = put(ph2.data(), ph2.size() * sizeof(double));
  h.off_ph3 = put(ph3.data(), ph3.size() * sizeof(double));
  h.off_xs = put(xs.data(), xs.size() * sizeof(int));
  h.off_slot = put(slot.data(), slot.size() * sizeof(int));
  h.off_slm = put(slm.data(), slm.size() * sizeof(int));
  h.off_slrz = put(slrz.data(), slrz.size() * sizeof(int));
  h.ok = true;
}

SepGeom sep_geom(const SepHost &h, const void *d_blob) {
  SepGeom g = h.g;
  g.gstride = 16 * g.ncols;  // the tile-major site correlator
  g.cstride = 16;
  const unsigned char *b = (const unsigned char *)d_blob;
  g.ph1 = (const cx<double> *)(b + h.off_ph1);
  g.ph1p = (const cx<double> *)(b + h.off_ph1p);
  g.ph2 = (const cx<double> *)(b + h.off_ph2);
  g.ph3 = (const cx<double> *)(b + h.off_ph3);
  g.xs = (const int *)(b + h.off_xs);
  g.slot = (const int *)(b + h.off_slot);
  g.sl_m = (const int *)(b + h.off_slm);
  g.sl_rz = (const int *)(b + h.off_slrz);
  return g;
}

template <typename T, int MAXA, int MAXB>
static int launch_xy(const SepGeom &g, const cx<T> *c, cx<double> *H, hipStream_t st) {
  dim3 block(256), grid((unsigned)((g.ncols + 15) / 16), (unsigned)g.Lz);
  switch (g.NPART) {
  case 1: k_sep_xy<T, 1, MAXA, MAXB><<<grid, block, 0, st>>>(g, c, H); break;
  case 2: k_sep_xy<T, 2, MAXA, MAXB><<<grid, block, 0, st>>>(g, c, H); break;
  case 4: k_sep_xy<T, 4, MAXA, MAXB><<<grid, block, 0, st>>>(g, c, H); break;
  case 8: k_sep_xy<T, 8, MAXA, MAXB><<<grid, block, 0, st>>>(g, c, H); break;
  case 16: k_sep_xy<T, 16, MAXA, MAXB><<<grid, block, 0, st>>>(g, c, H); break;
  default: return PLG_ERR_UNSUPPORTED;
  }
  return PLG_OK;
}

int launch_project_sep(const SepGeom &g, const void *c, int c_prec, double *part, hipStream_t st) {
  cx<double> *P = (cx<double> *)part;
  cx<double> *H = P + (int64_t)g.nmom * g.ncols;
  int rc;
  if (c_prec == PLG_F64)
    rc = g.small ? launch_xy<double, 5, 5>(g, (const cx<double> *)c, H, st)
                 : launch_xy<double, 8, 8>(g, (const cx<double> *)c, H, st);
  else
    rc = g.small ? launch_xy<float, 5, 5>(g, (const cx<float> *)c, H, st)
                 : launch_xy<float, 8, 8>(g, (const cx<float> *)c, H, st);
  if (rc) return rc;
  dim3 block(256), grid((unsigned)((g.ncols + 255) / 256), (unsigned)g.NS);
  if (g.small) k_sep_z<9><<<grid, block, 0, st>>>(g, H, P);
  else k_sep_z<16><<<grid, block, 0, st>>>(g, H, P);
  return hipGetLastError() == hipSuccess ? PLG_OK : PLG_ERR_HIP;
}

}  // namespace plg
