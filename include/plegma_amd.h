/*
 * plegma_amd.h -- C ABI of the MI355X (gfx950) two-point correlator library.
 *
 * Plain C: device pointers, sizes and status codes; no HIP or torch types
 * (streams are passed as void*, a hipStream_t or NULL for the null stream).
 * Every function returns PLG_OK (0) or a negative status; plg_last_error()
 * gives the message of the last failure on the calling thread.
 *
 * Each entry point names the PLEGMA interface it replaces (paths relative to
 * the PLEGMA source tree).  Single GPU, no MPI: the reference's spatial
 * all-reduce (lib/kernels/PLEGMA_mesons.cuh:161-164) is the identity here.
 */
#ifndef PLEGMA_AMD_H
#define PLEGMA_AMD_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------ */
enum plg_status {
  PLG_OK = 0,
  PLG_ERR_INVALID = -1,     /* bad argument (reference: PLEGMA_error)          */
  PLG_ERR_NOT_INIT = -2,    /* plg_init not called                             */
  PLG_ERR_HIP = -3,         /* HIP runtime error                               */
  PLG_ERR_ALLOC = -4,       /* device allocation failed                        */
  PLG_ERR_SITE_SIZE = -5,   /* site size mismatch (PLEGMA_mesons.cuh:140-141)  */
  PLG_ERR_UNSUPPORTED = -6, /* combination not provided by this library        */
  PLG_ERR_RANGE = -7        /* propagator view does not cover a needed slice   */
};

/* ---- enums ------------------------------------------------------------- */
enum plg_precision { PLG_F64 = 0, PLG_F32 = 1 };
/* arithmetic path: FP64 throughout, or FP32 contraction + split-bf16 MFMA
 * projection with FP32 accumulation and FP64 cross-block / time-slice sums */
enum plg_mode { PLG_MODE_F64 = 0, PLG_MODE_MIXED = 1 };
/* bf16 split products in the mixed projection (ignored in PLG_MODE_F64) */
enum plg_split { PLG_SPLIT_BF16X3 = 3, PLG_SPLIT_BF16X6 = 6 };
/* include/global/PLEGMA_enums.h:7 */
enum plg_space { PLG_POSITION_SPACE = 0, PLG_MOMENTUM_SPACE = 1 };
/* channel sets (correlator kinds) */
enum plg_kind {
  PLG_MESONS = 0,      /* contractMesons: 10 ch, datasets {S1, S2}            */
  PLG_MESONS16 = 1,    /* same formula over the 16 GAMMAS, datasets {S1, S2}  */
  PLG_MESONS_NEW = 2,  /* contractMesonsNew: 10 ch, S1 x conj(S2), 1 dataset  */
  PLG_MESONS_ALL = 3,  /* contractMesonsAll: 34 ch, 1 dataset                 */
  PLG_BARYONS = 4,     /* contractBaryons: groups by mask, 16 ch, 2 datasets  */
  PLG_MESON_MESON = 5, /* S1 = light, S2 = heavy: 16 mesons Tr[G S1 g5 G S2^dag
                          g5] (GAMMAS order), then (Tr A_G)^2 - Tr[A_G A_G]
                          for G = g5, g1, g2, g3 (A_G the 12x12 product);
                          20 ch, 1 dataset.  Counterpart of contractTetraquarks
                          (lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:163-212) */
  PLG_TWOPT = 6        /* one propagator: group 0 = 16 mesons, then baryon
                          groups by mask, 16 ch each, 1 dataset               */
};
/* baryon group bits in the reference's group order (PLEGMA_Correlator.cu:123-128) */
#define PLG_G_NUCL (1 << 0)
#define PLG_G_DELTAPP_11 (1 << 7)
#define PLG_G_DELTAPP_22 (1 << 8)
#define PLG_G_DELTAPP_33 (1 << 9)
#define PLG_G_DELTAPP_ALL (PLG_G_DELTAPP_11 | PLG_G_DELTAPP_22 | PLG_G_DELTAPP_33)

/* Propagator view (input of the path).  Complex interleaved SoA:
 *   element (component i, time t, site x) at ((i*comp_stride) + plane*V3 + x)
 *   complex units, plane = (t - t_first) mod T, plane < nt,
 *   i = ((mu*4 + nu)*3 + a)*3 + b.
 * PLEGMA's layout (lib/PLEGMA_Propagator.cu:39-44, lib/PLEGMA_Field.cu:225-264)
 * is comp_stride = V, t_first = 0, nt = T, plane_exp = NULL.
 * plane_exp (optional, device, nt ints): the stored entries of plane p are the
 * propagator times 2^-plane_exp[p] (an exact power-of-two rescaling that keeps
 * FP32 data of fast-decaying propagators inside the normal range).  Results
 * are returned unscaled.
 * rotate_sign (0, +1 or -1) and bc_t0 (0: none) make the contractions read
 * the view as if plg_prop_rotate_physical(rotate_sign) and then
 * plg_prop_apply_boundaries(bc_t0) had been applied to it, without changing
 * it: the preparation PLEGMA's drivers do in place before contracting
 * (plegma/mesons_2pt.cpp:158-165, plegma/PDFs_nucleon.cpp:382-385), fused
 * into the contraction's load stage.  Zero-initialised views prepare nothing. */
typedef struct {
  const void *data;    /* device pointer */
  int precision;       /* plg_precision */
  int t_first;
  int nt;
  int pad_;
  int64_t comp_stride; /* complex elements */
  const int *plane_exp;
  int rotate_sign;     /* 0: none; +-1: twisted-mass rotation to the physical basis */
  int bc_t0;           /* entries at global time t < bc_t0 enter negated */
} plg_prop;

/* ---- runtime ------------------------------------------------------------ */
/* replaces PLEGMA_init (lib/PLEGMA.cu:19-212) for one process on one GPU:
 * L = {Lx, Ly, Lz, Lt}; selects `device`, builds the phase tables. */
int plg_init(const int L[4], int device);
int plg_finalize(void);
const char *plg_last_error(void);
int plg_version(void);

/* ---- conventions -------------------------------------------------------- */
/* momentum list of PLEGMA_FT::createMom (lib/PLEGMA_FT.cu:50-72), 3D:
 * writes min(n, cap) triples (nx,ny,nz) to xyz (may be NULL), returns n in *nmom */
int plg_momlist(int q2max, int *nmom, int *xyz, int cap);
/* PLEGMA_Correlator::startT/endT/localT (include/PLEGMA_Correlator.h:53-79) */
int plg_window(const int source[4], int totalT, int *startT, int *endT, int *localT);
/* lattice time of buffer row j (lib/kernels/PLEGMA_mesons.cuh:22) */
int plg_row_time(const int source[4], int totalT, int j, int *t);
/* shape of a kind: groups, datasets, channels per group (getSiteSize =
 * ng*nd*nch, include/PLEGMA_Correlator.h:113-118) */
int plg_kind_shape(int kind, int group_mask, int *ngroups, int *ndatasets, int *nch);

/* ---- the hot path ------------------------------------------------------- */
/* One correlator, replacing contract_mesons / contract_mesons_new /
 * contract_mesons_all / contract_baryons (lib/kernels/PLEGMA_mesons.cuh:134-165,
 * PLEGMA_mesonsNew.cuh:65-97, PLEGMA_mesonsAll.cuh:61-92,
 * PLEGMA_baryons.cuh:169-206) with the fused DFT of
 * PLEGMA_kernel_utils.cuh:820-860 and the host block sums.
 *   moms: host array nmom*3 ints (a createMom list or any integer list)
 *   out : device buffer, H_elem layout of the reference
 *         momentum space: [g][d][j][m][c] complex, (PLEGMA_mesons.cuh:117-124,
 *                         PLEGMA_baryons.cuh:151-160, 196-200)
 *         position space: [g][d][j][x][c] complex
 *         of out_precision (PLG_F64 / PLG_F32).  Overwritten, not accumulated.
 * Synchronous with respect to `stream` ordering only. */
int plg_correlator(int kind, int group_mask, const plg_prop *p1, const plg_prop *p2,
                   const int source[4], int totalT, const int *moms, int nmom, int mode,
                   int space, int out_precision, void *out, void *stream);
/* plg_correlator without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory while they run
 * (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results as plg_correlator; the
 * plan tables, the site correlator and the projection partials of the call
 * live in one static device array of PLG_STATIC_SCRATCH_BYTES per GPU, set up
 * by plg_init and held until the process ends.  A correlator that needs more
 * (site correlator: rows x channels x sites x 8 B in FP32, 16 B in FP64)
 * returns PLG_ERR_UNSUPPORTED before any work is queued.  Tiles come from the
 * tuning cache when it holds the geometry, else from the built-in choice (no
 * tuning run).  Calls of one process are serialised; calls of several
 * processes, or replays of a captured call, that overlap on one GPU race on
 * the array.  Synchronises `stream` before returning. */
#define PLG_STATIC_SCRATCH_BYTES 134217728LL /* 128 MiB */
int plg_correlator_static(int kind, int group_mask, const plg_prop *p1, const plg_prop *p2,
                          const int source[4], int totalT, const int *moms, int nmom, int mode,
                          int space, int out_precision, void *out, void *stream);
/* number of complex elements plg_correlator writes (getTotalSize) */
int plg_correlator_size(int kind, int group_mask, const int source[4], int totalT, int nmom,
                        int space, int64_t *elems);

/* ---- udsc baryons ------------------------------------------------------
 * Replaces contractBaryonsUDSC / contract_baryons_udsc
 * (lib/PLEGMA_Correlator.cu:136-188, lib/kernels/PLEGMA_baryons_udsc.cu:267-300):
 * the two-point functions of the udsc baryon interpolators for every flavour
 * product ("ucc", "dcc", ..., "sss") whose propagators are given (NULL =
 * absent), optionally only products holding s (only_st) or c (only_ch).
 * Datasets: the product's operators in the reference's order, products in
 * the reference's order; names from plg_udsc_op_name ("Lambda/Pp_Cg5_Cg5").
 * out: device, [dataset][j][m] complex (momentum space) or [dataset][j][x]
 * (position space) of out_precision; mode as plg_correlator (PLG_MODE_MIXED
 * takes FP32 propagators).  Term tables: generated from the interpolating
 * fields (python/baryons.py of the reference makes the same tables). */
int plg_udsc_count(int *nprod);
/* flavour letters (3 chars + NUL into flavours[4]) and operator count */
int plg_udsc_product(int prod, char *flavours, int *nops);
int plg_udsc_op_name(int prod, int k, char *buf, int cap);
/* the products (indices) plg_baryons_udsc runs for these arguments */
int plg_baryons_udsc_select(const plg_prop *up, const plg_prop *dn, const plg_prop *st, const plg_prop *ch,
                            int only_st, int only_ch, int *prods, int cap, int *nprods, int *ndatasets);
int plg_baryons_udsc_size(const plg_prop *up, const plg_prop *dn, const plg_prop *st, const plg_prop *ch,
                          int only_st, int only_ch, const int source[4], int totalT, int nmom, int space,
                          int64_t *elems);
int plg_baryons_udsc(const plg_prop *up, const plg_prop *dn, const plg_prop *st, const plg_prop *ch, int only_st,
                     int only_ch, const int source[4], int totalT, const int *moms, int nmom, int mode, int space,
                     int out_precision, void *out, void *stream);
/* A three-propagator table contraction in the format of the reference's
 * generated tables (contract_props, lib/kernels/PLEGMA_baryons_udsc.cu:86-121,
 * tables BP_prop_prods_idxs / _vals / _count of
 * lib/kernels/PLEGMA_baryons_udsc_arrays.cuh):
 *   O_k = sum_t vals_t sum_{eps, eps'} sgn sgn' S1[i0][i1][a][a'] S2[i2][i3][b][b'] S3[i4][i5][c][c']
 * with operator k the next counts[k] terms; idxs: 6 spin indices per term,
 * vals: re, im (float) per term.  Limits: 16 operators, 4 distinct terms per
 * (i0 i1, i2 i3), 64 distinct coefficients (PLG_ERR_UNSUPPORTED beyond).
 * out: [k][j][m] (or [k][j][x]) as plg_baryons_udsc. */
int plg_baryons_table(const plg_prop *p1, const plg_prop *p2, const plg_prop *p3, const int16_t *idxs,
                      const float *vals, const int16_t *counts, int nops, const int source[4], int totalT,
                      const int *moms, int nmom, int mode, int space, int out_precision, void *out, void *stream);
/* A four-propagator table contraction in the format of the reference's
 * tetraquark tables (contract_props,
 * lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:6-62, tables
 * TETRA_prop_prods_idxs / _col_contr / _vals / _count of
 * PLEGMA_heavy_light_tetraquarks_arrays.cuh):
 *   O_k = sum_t vals_t sum_{a0..a3} P1[i0][i1][a[c0]][a[c1]] P2[i2][i3][a[c2]][a[c3]]
 *         conj(P3[i5+2][i4+2][a[c5]][a[c4]]) conj(P4[i7+2][i6+2][a[c7]][a[c6]])
 * (spin indices + 2 mod 4), operator k the next counts[k] terms; idxs i0..i7
 * and col_contr c0..c7: 8 int16 per term; vals: re, im (float) per term.
 * Supported: up to 64 operators; c0..c3 a permutation of 0..3 closing as
 * (c0,c1) = (c5,c4), (c2,c3) = (c7,c6); as c0 = c5, c1 = c6, c2 = c7,
 * c3 = c4 (the forms of the heavy-light tables); or as those two with P3 and
 * P4 exchanged, (c0,c1) = (c7,c6), (c2,c3) = (c5,c4) and c0 = c7, c1 = c4,
 * c2 = c5, c3 = c6 (the further forms of the bcud tables);
 * PLG_ERR_UNSUPPORTED else.
 * out: [k][j][m] (or [k][j][x]) as plg_baryons_udsc. */
int plg_tetraquark_table(const plg_prop *p1, const plg_prop *p2, const plg_prop *p3, const plg_prop *p4,
                         const int16_t *idxs, const int16_t *col_contr, const float *vals, const int16_t *counts,
                         int nops, const int source[4], int totalT, const int *moms, int nmom, int mode, int space,
                         int out_precision, void *out, void *stream);
/* ---- heavy-light tetraquarks --------------------------------------------
 * Replaces contractTetraquarks / contract_tetraquarks
 * (lib/PLEGMA_Correlator.cu:192-250,
 * lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:163-212): the two-point
 * functions of the heavy-light tetraquark operators (meson-meson,
 * meson-diquark, diquark-meson and diquark-diquark forms; 16 per product) for
 * every flavour product ("udbb", "usbb") whose propagators are given (NULL =
 * absent): u and d read lt, s reads st, c reads ch, b reads bt; optionally
 * only products holding s (only_st) or c (only_ch).  Datasets: the product's
 * operators in the reference's order ("bbud/g5_gj_g5_gj_ll", ...), products
 * in the reference's order; out as plg_baryons_udsc.  Term tables: generated
 * from the interpolating fields (python/tetraquarks_v3.py of the reference
 * makes the same tables). */
int plg_tetra_count(int *nprod);
/* flavour letters (4 chars + NUL into flavours[5]) and operator count */
int plg_tetra_product(int prod, char *flavours, int *nops);
int plg_tetra_op_name(int prod, int k, char *buf, int cap);
/* the products (indices) plg_tetraquarks runs for these arguments */
int plg_tetraquarks_select(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt,
                           int only_st, int only_ch, int *prods, int cap, int *nprods, int *ndatasets);
int plg_tetraquarks_size(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt,
                         int only_st, int only_ch, const int source[4], int totalT, int nmom, int space,
                         int64_t *elems);
int plg_tetraquarks(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt, int only_st,
                    int only_ch, const int source[4], int totalT, const int *moms, int nmom, int mode, int space,
                    int out_precision, void *out, void *stream);
/* ---- bcud tetraquarks ---------------------------------------------------
 * Replaces contractTetraquarksBCUD / contract_tetraquarks_bcud
 * (lib/PLEGMA_Correlator.cu:252-310,
 * lib/kernels/PLEGMA_bcud_tetraquarks.cu:159-212): the J = 1 (48) and J = 0
 * (9) two-point functions of the bcud tetraquark, flavour content
 * (bbar u cbar d) - (bbar d cbar u), product "udcb" (u, d read lt, c reads ch,
 * b reads bt; selection as plg_tetraquarks).  Datasets in the reference's
 * order ("bcud_J1/g5_gj_g5_gj_ll", ..., "bcud_J0/g5C_Cg5_g5C_Cg5_ll"); out as
 * plg_baryons_udsc.  Term tables: generated from the interpolating fields
 * (python/tetraquarks_v4.py of the reference makes the same tables). */
int plg_tetra_bcud_count(int *nprod);
int plg_tetra_bcud_product(int prod, char *flavours, int *nops);
int plg_tetra_bcud_op_name(int prod, int k, char *buf, int cap);
int plg_tetraquarks_bcud_select(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt,
                                int only_st, int only_ch, int *prods, int cap, int *nprods, int *ndatasets);
int plg_tetraquarks_bcud_size(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt,
                              int only_st, int only_ch, const int source[4], int totalT, int nmom, int space,
                              int64_t *elems);
int plg_tetraquarks_bcud(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt,
                         int only_st, int only_ch, const int source[4], int totalT, const int *moms, int nmom,
                         int mode, int space, int out_precision, void *out, void *stream);
/* the same, with HIP events between the stages of every product: nstages
 * stage names (32 bytes each) and their milliseconds summed over the products */
int plg_baryons_udsc_profile(const plg_prop *up, const plg_prop *dn, const plg_prop *st, const plg_prop *ch,
                             int only_st, int only_ch, const int source[4], int totalT, const int *moms, int nmom,
                             int mode, int space, int out_precision, void *out, void *stream, int cap, int *nstages,
                             double *ms, char *names);

/* Standalone momentum projection of a field, replacing PLEGMA_FT::apply with
 * FT_GEMV (lib/PLEGMA_FT.cu:144-162, lib/kernels/PLEGMA_FT.cuh:74-100):
 *   out[t][i][m] (+)= sum_x f_i(x, t) exp(sign i 2pi n_m . x / L)
 * with x the lattice coordinates (no source shift: PLEGMA_FT's
 * mulConstMomentumPhases applies that).  field: device, PLEGMA_Field layout,
 * dof components of nt*V3 complex (nt = T for a 4D field, 1 for a 3D field);
 * precision PLG_F64 or PLG_F32.  Where the momenta suit them (and the field
 * is 16-byte aligned) the separable sums over x, y and z run in FP64 on the
 * field in place; else the phase-matrix GEMM (FP64 MFMA, or split-bf16 MFMA
 * with FP64 sums for an FP32 field); PLG_PROJ_SEP=0 pins the GEMM.
 * moms: host nmom*3 ints.  out: device [nt][dof][nmom] complex of
 * out_precision, overwritten, or added to when accumulate != 0 (IsAccum).
 * The call synchronises the stream; the device tables of its momentum list
 * are kept until a call with another list, field shape or precision, or
 * plg_finalize. */
int plg_ft_apply(const void *field, int precision, int dof, int nt, const int *moms, int nmom,
                 int sign, int accumulate, void *out, int out_precision, void *stream);

/* ---- batched path (no single reference counterpart: the drivers loop over
 * sources calling contractMesons/contractBaryons, e.g. plegma/mesons_2pt.cpp) */
typedef struct {
  int source[4];
  int totalT;
  int prop;            /* index into the props array: S1 */
  int prop2;           /* S2 for kinds that read two propagators (mesons S2
                          dataset, contractMesonsNew/All, baryons, meson-meson
                          heavy flavour); -1: the same as prop */
  int64_t out_offset;  /* complex elements into the output buffer */
} plg_job;
typedef struct plg_batch_s *plg_batch;
#define PLG_BATCH_GRAPH 1   /* capture the launch sequence in a hipGraph, replay on run */
#define PLG_BATCH_TUNE 2    /* pick tiles with the autotuner (cached) */
int plg_batch_create(int kind, int group_mask, const plg_prop *props, int nprops,
                     const plg_job *jobs, int njobs, const int *moms, int nmom, int mode,
                     int split, int flags, plg_batch *out);
/* writes each job's H_elem block (layout of plg_correlator, momentum space,
 * out_precision F64) at its out_offset */
int plg_batch_run(plg_batch b, void *out, void *stream);
int plg_batch_info(plg_batch b, int64_t *out_elems, int64_t *nsites, int *nrows);
/* one eager pass with a HIP event after every stage on `stream`: per-stage
 * milliseconds and stage names (32 chars each) for up to `cap` stages */
int plg_batch_profile(plg_batch b, void *out, void *stream, int cap, int *nstages, double *ms,
                      char *names);
int plg_batch_destroy(plg_batch b);

/* ---- inputs ------------------------------------------------------------- */
/* seeded synthetic propagator: Philox4x32-10 + Box-Muller N(0,1) entries,
 * scaled by exp(-lambda*|t - t0|); counter (site + V3*t, component, stream).
 * With dst->plane_exp set, e = round(-lambda |t - t0| / ln 2) is written per
 * plane and entries are stored times 2^-e. */
int plg_synth_prop(const plg_prop *dst, uint64_t seed, uint32_t stream_id, int t0, double lambda,
                   void *stream);
/* base + eps * perturbation, both seeded, then the same scaling */
int plg_synth_prop_perturbed(const plg_prop *dst, uint64_t seed, uint32_t base_stream,
                             uint32_t pert_stream, double eps, int t0, double lambda, void *stream);
/* round-to-nearest FP64 -> FP32 copy over dst's planes; when dst->plane_exp
 * is set, it is written first (e = floor(log2 max |entry|) of each plane) and
 * the copy stores entry * 2^-e */
int plg_cast_prop(const plg_prop *dst, const plg_prop *src, void *stream);

/* Propagator preparation, in place on every plane of a view (the view's
 * rotate_sign and bc_t0 are ignored here).
 * plg_prop_rotate_physical: S -> U S U with U = (1 + i sign g5) / sqrt(2),
 *   S'[mu][nu] = (S[mu][nu] - S[mu+2][nu+2] + i sign (S[mu][nu+2] + S[mu+2][nu])) / 2
 *   (spin indices mod 4, every colour pair), sign = +-1; replaces
 *   PLEGMA_Propagator::rotateToPhysicalBase_device
 *   (lib/PLEGMA_Propagator.cu:103-106, lib/kernels/PLEGMA_propagator_utils.cuh:89-124).
 * plg_prop_apply_boundaries: negates every entry at global time t < t0;
 *   replaces PLEGMA_Propagator::applyBoundaries_device
 *   (lib/PLEGMA_Propagator.cu:98-100, lib/kernels/PLEGMA_propagator_utils.cuh:64-80). */
int plg_prop_rotate_physical(const plg_prop *p, int sign, void *stream);
int plg_prop_apply_boundaries(const plg_prop *p, int t0, void *stream);

/* ---- learned part (no counterpart in the reference): a multilayer
 * perceptron x -> [Linear -> SiLU] x (nlayers - 1) -> Linear -> y with BF16
 * weights/activations, FP32 accumulation and FP32 master weights, trained
 * with Adam on mean squared error.  dims[0] inputs ... dims[nlayers] outputs.
 * Parameter vector (FP32): for each layer W [out][in] row-major, then b [out].
 * x, t, y are device pointers to row-major [batch][features] FP32 arrays. */
typedef struct plg_mlp_s *plg_mlp;
int plg_mlp_create(const int *dims, int nlayers, int max_batch, plg_mlp *out);
int plg_mlp_nparams(plg_mlp net, int64_t *n);
/* host copies of the FP32 parameters; set also resets the optimiser state */
int plg_mlp_set_params(plg_mlp net, const float *host);
int plg_mlp_get_params(plg_mlp net, float *host);
/* inference: y = net(x) */
int plg_mlp_forward(plg_mlp net, const float *x, int batch, float *y, void *stream);
/* loss = mean (net(x) - t)^2 and its gradient w.r.t. the parameters
 * (grad_host: nparams floats, may be NULL) */
int plg_mlp_gradient(plg_mlp net, const float *x, const float *t, int batch, double *loss,
                     float *grad_host, void *stream);
/* `steps` full-batch Adam steps (beta 0.9/0.999, eps 1e-8); loss after the last */
int plg_mlp_train(plg_mlp net, const float *x, const float *t, int batch, int steps, float lr,
                  double *loss, void *stream);
int plg_mlp_destroy(plg_mlp net);

#ifdef __cplusplus
}
#endif
#endif
