// plg_internal.h -- declarations shared by the translation units of
// libplegma_amd.so (not installed).
#pragma once
#include "plegma_amd.h"
#include "plg_common.h"

#include <array>
#include <string>
#include <vector>

namespace plg {

// channel sets of the meson kernel
enum {
  SET_M16 = 0,       // 16 GAMMAS, one propagator
  SET_M10 = 1,       // contractMesons order, one propagator
  SET_M10_PAIR = 2,  // contractMesonsNew, S1 x conj(S2)
  SET_MALL = 3       // contractMesonsAll, 34 channels
};

template <typename T> struct PropDev {
  const cx<T> *data;
  int64_t cs;  // component stride (complex)
  int t_first;
  int nt;
  int rot;   // 0, or the sign of the rotation to the physical basis applied on load
  int bct0;  // entries at global time t < bct0 are negated on load
};
// sign of a propagator's entries at global time t after the boundary fix
template <typename T> __device__ __forceinline__ T prop_sign(const PropDev<T> &P, int t) {
  return t < P.bct0 ? (T)-1 : (T)1;
}

// one contraction row: a time slice of a job and its propagators
struct RowDev {
  int t;
  int pa, pb;
  int job;
  int pc;  // third propagator (udsc baryons)
};

// a finalize row: computed row, source, output base (complex elements).
// Output channel k of the row reads computed channel ch_base + csel[k] and
// lands at base + ch_off[k] + m * m_stride.
struct OutRowDev {
  int crow;      // row of the site correlator / partial sums
  int ch_base;
  int sx, sy, sz;
  int nch_row;   // channels of this row (0: the launch-wide count)
  int64_t base;  // H offset of (m = 0, k = 0)
  // result scale 2^(ka*ea[pla] + kb*eb[plb] + kc*ec[plc]) undoes the
  // propagators' plane exponents (ea/eb/ec NULL: no scaling)
  const int *ea, *eb;
  int pla, plb, ka, kb;
  const int *ec;
  int plc, kc;
  const int *ed;
  int pld, kd;
};
__device__ __forceinline__ double row_scale(const OutRowDev &R) {
  int e = 0;
  if (R.ea) e += R.ka * R.ea[R.pla];
  if (R.eb) e += R.kb * R.eb[R.plb];
  if (R.ec) e += R.kc * R.ec[R.plc];
  if (R.ed) e += R.kd * R.ed[R.pld];
  return ldexp(1.0, e);
}

// udsc baryons: chunked schedule of one flavour product's term tables.
// Chunk c holds up to 64 (m0m1, m2m3) quads, one per lane; lane l's quad has
// up to UD_TPQ distinct terms, kept in the chunk-local slots e*64 + l.
// Operator k owns the lanes [op_lo[k], op_lo[k] + op_n[k]) (all 64 lanes are
// owned, more lanes to operators with more entries); in chunk c each lane sums
// its share of its operator's entries, nblk[c] blocks of 4 entries, laid out
// as [block][lane][4] from block blk0[c] (zero-coefficient padding).
constexpr int UD_MAXC = 4;        // chunks
constexpr int UD_TPQ = 4;         // distinct terms per quad
constexpr int UD_MAXOPS = 16;     // operators of a product
constexpr int UD_MAXBLK = 16;     // entry blocks of a product (4 per lane each)
constexpr int UD_MAXENT = UD_MAXBLK * 256;
constexpr int UD_MAXCOEF = 64;    // distinct coefficients of a product
struct UdscSched {
  int nops, nchunks, nent, ncoef;
  int real_coef;                    // every coefficient has a zero imaginary part
  int ntp[UD_MAXC];                 // most distinct terms of a quad in chunk c
  int blk0[UD_MAXC], nblk[UD_MAXC];
  int op_lo[UD_MAXOPS], op_n[UD_MAXOPS];
  const short *quads;       // [chunk][lane]: m0m1 | m2m3 << 4, -1 = none
  const signed char *tp;    // [chunk][lane][UD_TPQ]: m4m5 of each term, -1 = none
  const uint16_t *ents;     // chunk-local term slot | coefficient index << 8
  const float *coef;        // [ncoef][2] coefficients (float, as the reference's tables); 0 is zero
};
struct UdscSchedHost {
  UdscSched s{};
  std::vector<short> quads;
  std::vector<signed char> tp;
  std::vector<uint16_t> ents;
  std::vector<float> coef;
};
int udsc_schedule(int prod, UdscSchedHost &h);
// one term of a table: six spin indices and a coefficient (the layout of the
// generated tables, plg_udsc_tables.h)
struct UdscTerm {
  signed char m[6];
  float re, im;
};
int udsc_schedule_terms(const UdscTerm *terms, const int *first, const int *count, int nops, UdscSchedHost &h);

// four-propagator table contractions (tetraquark table format)
constexpr int TETRA_MAXOPS = 64;
// A term, 32 bits: the (P1, P3) spin pair p = 16 s1 + s3 (bits 0-7), the
// (P2, P4) pair q = 16 s2 + s4 (bits 8-15) and the coefficient index (bits
// 16-31), with the 4-bit spin pairs s1 = 4 i0 + i1, s2 = 4 i2 + i3,
// s3 = 4 (i5+2) + (i4+2), s4 = 4 (i7+2) + (i6+2) (indices mod 4).  An
// operator's terms are its direct ones (<P1,P3><P2,P4>) first, then its
// crossed ones (Tr[(P3^dag P1)(P4^dag P2)]).
typedef uint32_t TetraTerm;
constexpr int TETRA_MAXCOEF = 65536;
struct TetraSched {
  int nops;
  int first[TETRA_MAXOPS], count[TETRA_MAXOPS], ndirect[TETRA_MAXOPS];
  const TetraTerm *terms;
  const float *coef;  // re, im of each distinct coefficient
  int ncoef;
  int accumulate;     // add the operators to the site correlator instead of writing them
};
struct RowDev4 {
  int t;
  int p[4];
};
// The terms of a table in two sets.  Set 0 pairs the colours of P1 with P3
// and of P2 with P4 (the two structures above); set 1 pairs P1 with P4 and
// P2 with P3 (<P1,P4><P2,P3> and Tr[(P4^dag P1)(P3^dag P2)]), which are the
// same structures on the views P1, P2, P4, P3, and is packed for those views.
struct TetraPacked {
  std::vector<TetraTerm> terms[2];
  std::vector<int> first[2], count[2], ndirect[2];
  std::vector<float> coef;  // shared by both sets
};
int tetra_pack(const int16_t *idxs, const int16_t *cols, const float *vals, const int16_t *counts, int nops,
               TetraPacked &out);
template <typename T>
int launch_tetra(const TetraSched &S, const PropDev<T> *props, const RowDev4 *rows, int nrows, cx<T> *out,
                 int64_t V3, int T_lat, int nch_total, int spw, hipStream_t st);
int udsc_products();
const char *udsc_flavours(int prod);
int udsc_nops(int prod);
const char *udsc_op_name(int prod, int k);
template <typename T>
int launch_udsc(const UdscSched &S, const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out, int64_t V3,
                int T_lat, int nch_total, int ch0, int spw, hipStream_t st);

template <typename T>
int launch_meson(int set, const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out,
                 int64_t V3, int T_lat, int nch_total, int ch0, hipStream_t st);
template <typename T>
int launch_mesmes(const PropDev<T> *props, const RowDev *rows, int nrows, cx<T> *out, int64_t V3, int T_lat,
                  int nch_total, int ch0, int tiles_per_block, hipStream_t st);
template <typename T>
int launch_baryon(bool two, bool mes, int gmask, const PropDev<T> *props, const RowDev *rows,
                  int nrows, cx<T> *out, int64_t V3, int T_lat, int nch_total, int ch0,
                  int tiles_per_block, hipStream_t st);

// projection: partial sums over each K split's site range of
// sum_x exp(-i 2pi n.x / L) c[n][x].  Both paths (FP64 and mixed) write the
// paired layout part[split][2n + {re, im}][2j + {cos, sin}], real double,
// j = pair of momenta (n, -n) (see part_value); npairs == 0 (no momentum
// list) leaves the complex layout part[split][n][m].
struct ProjGeom {
  int64_t V3;
  int L[3];
  int M;          // common phase denominator lcm(Lx,Ly,Lz)
  int nmom;
  int ncols;      // n = row*nch + ch
  int nsplit;
  const int *mom_scaled;   // device, nmom*3: n_k * (M / L_k) mod M
  const double *cos_tab;   // device, M entries cos(2 pi k / M)
  const double *sin_tab;
  int npairs = 0;              // momentum pairs of the paired layout
  const int *pair_ms = nullptr;   // device, npairs*3: scaled representative momenta
  const int *mom_pair = nullptr;  // device, nmom: 2j (+1 for the -n member)
  const uint16_t *pidx = nullptr;  // device, phase indices n_j.x (M/L) mod M of the pairs
  int zero_pair = 0;  // 1: the last pair is n = 0, summed directly instead of a GEMM row
  // 1: the partials hold sums of c - a (a: per-column shift stored after
  // them, see proj_shift); the mixed projection, whose FP32 accumulation of
  // a column's mean value would otherwise leave a coherent error at n != 0
  int shifted = 0;
  // layout of the K x N operand: 1 = site correlator, tile-major (below);
  // 0 = a field, column-major (column n, site x at n*V3 + x).  The mixed
  // projection reads the tile-major layout only.
  int ctile;
  // 1: part holds the separable projection's sums, momentum-major
  // (part[m][n], one complex double each, see plg_project_sep.hip)
  int sep = 0;
};
// The site correlator is stored tile-major: the 16 sites of a site tile are
// contiguous per column, and the columns of a tile are contiguous, so that a
// projection step (16 sites x many columns) reads one contiguous block and a
// contraction workgroup writes one.  Element (column n, site x):
__host__ __device__ __forceinline__ int64_t ctile_idx(int64_t n, int64_t x, int64_t ncols) {
  return ((x >> 4) * ncols + n) * 16 + (x & 15);
}
__host__ __device__ __forceinline__ int64_t cidx(const ProjGeom &g, int64_t n, int64_t x) {
  return g.ctile ? ctile_idx(n, x, g.ncols) : n * g.V3 + x;
}
// sites of a site-correlator column rounded up to whole tiles
__host__ __device__ __forceinline__ int64_t ctile_sites(int64_t V3) { return (V3 + 15) / 16 * 16; }
int launch_project_f64(const ProjGeom &g, const cx<double> *c, cx<double> *part, int mtile,
                       hipStream_t st);
// mixed projection in the paired layout, on the phase indices g.pidx and
// wtab (M x 6 packed words, the first three cos_p | sin_p << 16 for the bf16
// pieces p of cos/sin(2 pi k / M)); code: tile code 1000 + 100 WR + 10 WC
// + GR (see plg_project_bf16.hip)
int launch_project_bf16(const ProjGeom &g, const cx<float> *c, double *part, int split_terms,
                        const uint32_t *wtab, int code, hipStream_t st);
// phase indices n_j . x (M / L) mod M of g's pairs
size_t phase_index_bytes(int64_t V3, int npairs);
int launch_phase_index(const ProjGeom &g, uint16_t *tab, hipStream_t st);
// column-major FP32 field (column n at n * V3) -> tile-major layout (ctile_idx)
int launch_to_tiles(const cx<float> *f, int64_t ncols, int64_t V3, cx<float> *t, hipStream_t st);
// the same for an FP64 field (the separable projection of plg_ft_apply)
// (npairs here: the pairs of the GEMM, without a zero-momentum pair)
int64_t proj_workgroups(int code, int64_t ncols, int npairs);
int pick_proj(int npairs);
std::vector<int> proj_candidates(int npairs);
// pairs (n, -n) of scaled momenta; returns the number of pairs.  A zero
// momentum pair is moved to the end (*zero_last = 1).
int make_pairs(const std::vector<int> &ms, int M, std::vector<int> &pair_ms, std::vector<int> &mom_pair,
               int *zero_last);
// Per-column shift of the mixed projection: complex FP32 a[n] after the
// paired partials.  The projection sums phase x (c - a); since the phases of
// a momentum n != 0 (mod L) sum to zero over the volume, only momenta whose
// scaled components are all zero get V3 a added back (part_value).
__host__ __device__ __forceinline__ cx<float> *proj_shift(const ProjGeom &g, double *part) {
  return (cx<float> *)(part + (int64_t)g.nsplit * 2 * g.ncols * 2 * g.npairs);
}
__host__ __device__ __forceinline__ const cx<float> *proj_shift(const ProjGeom &g, const double *part) {
  return (const cx<float> *)(part + (int64_t)g.nsplit * 2 * g.ncols * 2 * g.npairs);
}
// the projection sum_s part[s] of column n, momentum m, in either layout
__device__ __forceinline__ cx<double> part_value(const ProjGeom &g, const double *part, int64_t n, int m) {
  double re = 0, im = 0;
  if (g.sep) {
    const double *p = part + 2 * ((int64_t)m * g.ncols + n);
    return cmake<double>(p[0], p[1]);
  }
  if (g.npairs > 0) {
    const int pm = g.mom_pair[m];
    const int R = 2 * g.npairs, rc = pm & ~1;
    const int64_t nreal = 2 * (int64_t)g.ncols;
    double cr = 0, ci = 0, sr = 0, si = 0;
    for (int s = 0; s < g.nsplit; s++) {
      const double *p = part + ((int64_t)s * nreal + 2 * n) * R + rc;
      cr += p[0];
      sr += p[1];
      ci += p[R];
      si += p[R + 1];
    }
    const int *pj = g.pair_ms + 3 * (pm >> 1);
    if (g.shifted && pj[0] == 0 && pj[1] == 0 && pj[2] == 0) {
      const cx<float> a = proj_shift(g, part)[n];
      cr += (double)g.V3 * (double)a.re;
      ci += (double)g.V3 * (double)a.im;
    }
    // P(n) = (Cr + Si) + i (Ci - Sr), P(-n) = (Cr - Si) + i (Ci + Sr)
    if (pm & 1) {
      re = cr - si;
      im = ci + sr;
    } else {
      re = cr + si;
      im = ci - sr;
    }
  } else {
    for (int s = 0; s < g.nsplit; s++) {
      const double *p = part + 2 * (((int64_t)s * g.ncols + n) * g.nmom + m);
      re += p[0];
      im += p[1];
    }
  }
  return cmake<double>(re, im);
}
// bytes of the partial sums of a projection
inline size_t proj_part_bytes(int nsplit, int64_t ncols, int nmom, int npairs) {
  return npairs > 0 ? (size_t)nsplit * ncols * npairs * 32 + (size_t)ncols * sizeof(cx<float>)
                    : (size_t)nsplit * ncols * nmom * 16;
}
// Separable projection (plg_project_sep.hip): the phase of a momentum
// factorises over the three directions, so the sums run direction by
// direction in FP64 -- over x per line and over y per plane (k_sep_xy), then
// over z (k_sep_z) -- instead of as one phase-matrix GEMM over the volume.
constexpr int PROJ_SEP = 1;  // projection code of the separable form (both precisions)
struct SepGeom {
  int Lx, Ly, Lz;
  int64_t ncols;
  // element strides of the input: of a group of 16 sites and of a column.
  // The tile-major site correlator: 16 * ncols and 16; a field in its own
  // [column][site] layout: 16 and V3
  int64_t gstride, cstride;
  int nmom;
  int NA;     // distinct |n_x| classes a (cos and sin sums over x)
  int NXR;    // distinct n_x residues
  int NYR;    // distinct n_y residues
  int NB;     // distinct |n_y| classes b (cos and sin sums over y)
  int NS;     // distinct (n_x, n_y) residue pairs: planes of the per-plane sums
  int NPART;  // threads per line of a chunk (the line's sites split in parts)
  int W;      // sites of a part (Lx / NPART)
  int PS, LS, RS;  // LDS strides of a staged chunk: part, line, column (elements)
  // 1: NA <= 5, NB <= 5, at most 9 momenta per (n_x, n_y) pair; the
  // tables are padded to these counts (else 8, 8, 16) with zero phases and
  // slots -1, so that the kernels' loops run without bounds
  int small;
  const cx<double> *ph1;   // [W][NAP]: cos, sin (2 pi a i / Lx)
  const cx<double> *ph1p;  // [NPART][NAP]: cos, sin (2 pi a p W / Lx)
  const cx<double> *ph2;   // [Ly][NBP]: cos, sin (2 pi b y / Ly)
  const cx<double> *ph3;   // [Lz][Lz]: cos, sin (2 pi r z / Lz) at r * Lz + z
  const int *xs;           // [NAP][2]: residue slot of +a and of -a (or -1)
  const int *slot;         // [NXR][NBP][2]: pair index of (n_x, +b) and (n_x, -b), or -1
  const int *sl_m, *sl_rz; // [NS][NZP]: momentum index (or -1) and n_z residue of a pair's momenta
};
// host description of a plan's separable projection
struct SepHost {
  bool ok = false;
  SepGeom g{};                     // scalars (pointers set by sep_geom)
  std::vector<unsigned char> blob;  // the device tables, uploaded as one piece
  size_t off_ph1 = 0, off_ph1p = 0, off_ph2 = 0, off_ph3 = 0, off_xs = 0, off_slot = 0, off_slm = 0,
         off_slrz = 0;
};
// tables of the separable projection of scaled momenta ms (3 per momentum)
// on L; false (h.ok) when the momenta or the lattice do not suit it
void sep_build(const int L[3], int M, const std::vector<int> &ms, int64_t ncols, bool f64, SepHost &h);
SepGeom sep_geom(const SepHost &h, const void *d_blob);
// bytes of part for the separable projection: the sums part[m][n], then
// the per-plane sums of the (n_x, n_y) pairs
inline size_t sep_part_bytes(const SepGeom &g) {
  return ((size_t)g.nmom + (size_t)g.Lz * g.NS) * (size_t)g.ncols * sizeof(cx<double>);
}
int launch_project_sep(const SepGeom &g, const void *c, int c_prec, double *part, hipStream_t st);
int launch_ft_out(const ProjGeom &g, const double *part, int dof, int nt, int accumulate, void *out,
                  int out_prec, hipStream_t st);
// sum over splits, source phase, scatter into the H layout
// nch_c: channels per computed row; nout: output channels per out row
int launch_finalize(const ProjGeom &g, const double *part, const OutRowDev *orows, int norows,
                    int nch_c, int nout, const int *csel, const int64_t *ch_off, int64_t m_stride,
                    void *out, int out_prec, hipStream_t st);
// position space: c[crow][ch][x] -> out[base + ch_off[k] + x*x_stride] (H position layout)
int launch_scatter_pos(const void *c, int c_prec, int64_t V3, int64_t ncols, const OutRowDev *orows, int norows,
                       int nch_c, int nout, const int *csel, const int64_t *ch_off, int64_t x_stride,
                       void *out, int out_prec, hipStream_t st);

// inputs
int launch_synth(const plg_prop &dst, const int L[4], uint64_t seed, uint32_t sa, uint32_t sb,
                 double eps, int t0, double lambda, hipStream_t st);
int launch_cast(const plg_prop &dst, const plg_prop &src, int64_t V3, int T_lat, hipStream_t st);

// propagator preparation in place (all planes of a view)
int launch_prop_rotate(const plg_prop &p, int64_t V3, int sign, hipStream_t st);
int launch_prop_negate(const plg_prop &p, int64_t V3, const int *d_planes, int nplanes, hipStream_t st);

// error reporting
void set_error(const std::string &msg);
int fail(int code, const std::string &msg);

}  // namespace plg
