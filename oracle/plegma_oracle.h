/*
 * plegma_oracle.h -- CPU oracle for the two-point correlator path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py may load it.
 * The product (libplegma_amd.so) never links, loads or calls it.
 *
 * It restates, in plain FP64 C with OpenMP, the per-site bodies and the
 * momentum projection of PLEGMA's 2pt path:
 *   mesons      lib/kernels/PLEGMA_mesons.cuh:4-65
 *   mesonsNew   lib/kernels/PLEGMA_mesonsNew.cuh:5-60
 *   mesonsAll   lib/kernels/PLEGMA_mesonsAll.cuh:7-95
 *   proton      lib/kernels/PLEGMA_baryons_NtoN.cu:3-46
 *   deltas      lib/kernels/PLEGMA_baryons_deltas.cu:3-110
 *   DFT         lib/kernels/PLEGMA_kernel_utils.cuh:820-860
 *   momenta     lib/PLEGMA_FT.cu:50-72
 *   windows     include/PLEGMA_Correlator.h:53-79, PLEGMA_mesons.cuh:17-23
 *   layout      PLEGMA_mesons.cuh:113-124, PLEGMA_baryons.cuh:146-160,196-200
 * The index/sign tables those bodies iterate over are generated here from the
 * gamma matrices (basis of lib/kernels/PLEGMA_gammas.cuh:6-41) rather than
 * typed in; tests compare them with the reference's tables.
 */
#ifndef PLEGMA_ORACLE_H
#define PLEGMA_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* correlator kinds (channel sets) */
enum {
  ORC_MESONS = 0,      /* contractMesons: 10 ch, datasets {S1,S2}          */
  ORC_MESONS16 = 1,    /* same formula over the 16 GAMMAS, datasets {S1,S2} */
  ORC_MESONS_NEW = 2,  /* contractMesonsNew: 10 ch, S1 x conj(S2), 1 dataset */
  ORC_MESONS_ALL = 3,  /* contractMesonsAll: 34 ch, 1 dataset               */
  ORC_BARYONS = 4,     /* contractBaryons: groups selected by mask, 16 ch, 2 datasets */
  ORC_MESON_MESON = 5, /* four-propagator meson-meson (defined by formula)   */
  ORC_TWOPT = 6        /* one propagator: 16 mesons then baryon groups, 1 dataset */
};
/* baryon group bits, in the reference's group order */
enum {
  ORC_G_NUCL = 1 << 0,       /* baryons/nucl_nucl                */
  ORC_G_DELTA12_11 = 1 << 4, /* baryons/deltap_deltaz_11 ..       */
  ORC_G_DELTA12_22 = 1 << 5,
  ORC_G_DELTA12_33 = 1 << 6,
  ORC_G_DELTA32_11 = 1 << 7, /* baryons/deltapp_deltamm_11 ..     */
  ORC_G_DELTA32_22 = 1 << 8,
  ORC_G_DELTA32_33 = 1 << 9
};

/* propagator view: SoA, element (comp, plane, x) complex at
 *   data[((comp*comp_stride) + plane*V3 + x)*2 + {0,1}]
 * plane = (t - t_first) mod T must be < nt.  The reference layout
 * (lib/PLEGMA_Propagator.cu:39-44) is comp_stride = V, t_first = 0, nt = T. */
typedef struct {
  const double *data;
  int64_t comp_stride;
  int t_first;
  int nt;
} orc_prop;

int orc_nmom(int q2max);
/* writes nmom*3 ints (nx,ny,nz) in createMom order */
int orc_momlist(int q2max, int *out);
/* reference window rule for a single process (local = global). */
void orc_window(const int L[4], const int src[4], int totalT, int *startT,
                int *endT, int *localT);
/* lattice time of buffer row j */
int orc_row_time(const int L[4], const int src[4], int totalT, int j);

int orc_kind_nch(int kind, int group_mask, int *ndatasets, int *ngroups);

/* Per-site bodies.  S1,S2 point to one site as double[4][4][3][3][2].
 * out receives the reference accum[] array (complex, interleaved). */
void orc_site_mesons10(const double *S1, const double *S2, double *out);   /* 20 */
void orc_site_mesons16(const double *S1, const double *S2, double *out);   /* 32 */
void orc_site_mesons_new(const double *S1, const double *S2, double *out); /* 10 */
void orc_site_mesons_all(const double *S1, const double *S2, double *out); /* 34 */
void orc_site_nton(const double *S1, const double *S2, double *out);       /* 32 */
void orc_site_delta32(int k, const double *S1, const double *S2, double *out); /* 32 */
void orc_site_delta12(int k, const double *S1, const double *S2, double *out); /* 32 */
void orc_site_meson_meson(const double *SL, const double *SH, double *out);    /* 2*NMM */

/* Propagator preparation.
 * orc_rotate_site: rotation to the physical basis of one site (144 complex,
 *   component ((mu*4+nu)*3+a)*3+b), sign +-1:
 *   rotateToPhysicalBase_kernel, lib/kernels/PLEGMA_propagator_utils.cuh:89-124.
 * orc_prop_rotate: the same over a view (nt planes of V3 sites), in place.
 * orc_prop_boundaries: negates the planes of a view whose global time is
 *   below t0: apply_boundaries_kernel, lib/kernels/PLEGMA_propagator_utils.cuh:64-80. */
void orc_rotate_site(const double *S, int sign, double *out);
void orc_prop_rotate(double *data, int64_t comp_stride, int nt, int64_t V3, int sign);
void orc_prop_boundaries(double *data, int64_t comp_stride, int t_first, int nt, int T, int64_t V3, int t0);

/* Full correlator in the reference H_elem layout
 *   H[((((g*nd + d)*localT + j)*nmom + m)*nch + c)*2 + ri]
 * Rows listed in `rows` (nrows>0) are computed, others left untouched;
 * nrows<=0 computes every row.  nthreads<=0 uses the OpenMP default.
 * Returns 0 on success. */
int orc_correlator(int kind, int group_mask, const int L[4], const int src[4],
                   int totalT, int q2max, const orc_prop *p1, const orc_prop *p2,
                   double *H, const int *rows, int nrows, int nthreads);

/* Position-space site correlator for one row: out[(x*nch_total + c)*2+ri]
 * with nch_total = nd*ng*nch (dataset-major like the kernel's accum[]). */
int orc_site_correlator(int kind, int group_mask, const int L[4], int t,
                        const orc_prop *p1, const orc_prop *p2, double *out,
                        int nthreads);

/* udsc baryons (lib/kernels/PLEGMA_baryons_udsc.cu:86-121): operators k of a
 * term table (idx 6 ints, val re/im per term; operator k = terms
 * [first[k], first[k]+count[k])) on one site, out[2*k + ri] */
void orc_site_udsc(const double *S1, const double *S2, const double *S3, const int *idx, const double *val,
                   const int *first, const int *count, int nops, double *out);
/* the same over the V3 sites of time slice t: out[(x*nops + k)*2 + ri] */
int orc_udsc_field(const int L[4], int t, const orc_prop *p1, const orc_prop *p2, const orc_prop *p3,
                   const int *idx, const double *val, const int *first, const int *count, int nops, double *out,
                   int nthreads);

/* four-propagator tables (lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:26-49):
 * 8 spin indices and 8 colour slots per term, operator k = terms
 * [first[k], first[k]+count[k]); one site, out[2*k + ri] */
void orc_site_tetra(const double *S1, const double *S2, const double *S3, const double *S4, const int *idx,
                    const int *col, const double *val, const int *first, const int *count, int nops, double *out);
/* the same over the V3 sites of time slice t: out[(x*nops + k)*2 + ri] */
int orc_tetra_field(const int L[4], int t, const orc_prop *p1, const orc_prop *p2, const orc_prop *p3,
                    const orc_prop *p4, const int *idx, const int *col, const double *val, const int *first,
                    const int *count, int nops, double *out, int nthreads);

/* Naive DFT of a site field (reference fourier_transform_3D semantics, sign -1):
 * f[(x*ncomp + c)*2+ri] over V3 sites -> out[(m*ncomp + c)*2+ri]. */
void orc_dft(const int L[4], const int src3[3], int q2max, const double *f,
             int ncomp, double *out, int nthreads);

/* Seeded synthetic propagator (Philox4x32-10 + Box-Muller), identical to the
 * generator kernel of the product up to libm rounding.  Writes planes
 * t_first..t_first+nt-1 (mod T) of a SoA view with comp_stride.
 * mode 0: plain; scaling exp(-lambda*|t - t0|) without periodic wrap. */
void orc_synth_prop(const int L[4], uint64_t seed, uint32_t stream, int t0,
                    double lambda, double *data, int64_t comp_stride, int t_first,
                    int nt, int nthreads);
/* raw N(0,1) pair for (seed, stream, site, comp) -- for generator tests */
void orc_normal_pair(uint64_t seed, uint32_t stream, uint32_t site, uint32_t comp,
                     double *re, double *im);

/* Philox4x32-10 block (Salmon et al., SC'11), for known-answer tests */
void orc_philox(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* tables exported for the tests that compare them with the reference's */
int orc_meson_table(int gamma_idx, int *idx /*16x4: beta,gamma,delta,alpha*/,
                    double *val /*16x2*/);
int orc_diquark_table(int which /*0: Cg5, 1..3: Cgk*/, int *idx /*16x4*/,
                      double *val /*16x2*/);
void orc_gamma(int op, double *m /*4x4x2*/);

#ifdef __cplusplus
}
#endif
#endif
