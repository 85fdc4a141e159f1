/*
 * plegma_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY; see the header).
 *
 * FP64 restatement of PLEGMA's two-point correlator path.  Each body follows
 * the loop structure of the reference kernel it names (table-driven sums over
 * spin terms and epsilon colour permutations), so that the GPU path is
 * checked against the reference's own algorithm rather than against the
 * factorised forms the GPU kernels use.
 */
#include "plegma_oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;
#define NS 4
#define NC 3
#define NPROP 144
/* reference constant, include/PLEGMA_global.h:39 */
#define REF_PI 3.141592653589793

/* ------------------------------------------------------------------ */
/* gamma basis: lib/kernels/PLEGMA_gammas.cuh:6-41 (GAMMAS order,      */
/* include/global/PLEGMA_enums.h:22).  Built from gamma_1..gamma_4.    */
/* ------------------------------------------------------------------ */
static cplx GM[16][4][4];
static int gm_ready = 0;

static void mat_mul(cplx a[4][4], cplx b[4][4], cplx r[4][4]) {
  cplx t[4][4];
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      cplx s = 0;
      for (int k = 0; k < 4; k++) s += a[i][k] * b[k][j];
      t[i][j] = s;
    }
  memcpy(r, t, sizeof(t));
}

static void sigma(cplx a[4][4], cplx b[4][4], cplx r[4][4]) {
  /* -(i/2)[a,b] */
  cplx ab[4][4], ba[4][4];
  mat_mul(a, b, ab);
  mat_mul(b, a, ba);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) r[i][j] = -0.5 * I * (ab[i][j] - ba[i][j]);
}

static void build_gammas_once(void) {
  cplx g[5][4][4];
  memset(g, 0, sizeof(g));
  /* gamma_1 */
  g[1][0][3] = I; g[1][1][2] = I; g[1][2][1] = -I; g[1][3][0] = -I;
  /* gamma_2 */
  g[2][0][3] = 1; g[2][1][2] = -1; g[2][2][1] = -1; g[2][3][0] = 1;
  /* gamma_3 */
  g[3][0][2] = I; g[3][1][3] = -I; g[3][2][0] = -I; g[3][3][1] = I;
  /* gamma_4 */
  g[4][0][0] = 1; g[4][1][1] = 1; g[4][2][2] = -1; g[4][3][3] = -1;
  cplx g5[4][4];
  mat_mul(g[1], g[2], g5);
  mat_mul(g5, g[3], g5);
  mat_mul(g5, g[4], g5);
  memset(GM, 0, sizeof(GM));
  for (int i = 0; i < 4; i++) GM[0][i][i] = 1;
  for (int k = 1; k <= 4; k++) memcpy(GM[k], g[k], sizeof(g5));
  memcpy(GM[5], g5, sizeof(g5));
  for (int k = 1; k <= 4; k++) mat_mul(g5, g[k], GM[5 + k]);
  sigma(g[1], g[2], GM[10]);
  sigma(g[1], g[3], GM[11]);
  sigma(g[2], g[3], GM[12]);
  sigma(g[4], g[1], GM[13]);
  sigma(g[4], g[2], GM[14]);
  sigma(g[4], g[3], GM[15]);
  gm_ready = 1;
}
static pthread_once_t gm_once = PTHREAD_ONCE_INIT;
static void build_gammas(void) { pthread_once(&gm_once, build_gammas_once); }

void orc_gamma(int op, double *m) {
  build_gammas();
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) {
      m[(i * 4 + j) * 2 + 0] = creal(GM[op][i][j]);
      m[(i * 4 + j) * 2 + 1] = cimag(GM[op][i][j]);
    }
}

/* ------------------------------------------------------------------ */
/* meson term tables.  The reference (PLEGMA_mesons.cuh:6-8) stores,  */
/* per channel, 16 terms (beta,gamma,delta,alpha; value) for           */
/*   accum += value * S[alpha][beta] . conj(S[delta][gamma])           */
/* which is Tr[G S g5 G S^dag g5]: value = X[delta][alpha] X[beta][gamma] */
/* with X = g5 G.  Generated here from the matrices.                  */
/* ------------------------------------------------------------------ */
typedef struct {
  int b, g, d, a;
  cplx v;
} mterm;

static void meson_terms_build(int op, mterm t[16]) {
  cplx X[4][4];
  mat_mul(GM[5], GM[op], X);
  int n = 0;
  for (int alpha = 0; alpha < 4; alpha++)
    for (int beta = 0; beta < 4; beta++) {
      int delta = -1, gam = -1;
      for (int r = 0; r < 4; r++)
        if (cabs(X[r][alpha]) > 0.5) delta = r;
      for (int c = 0; c < 4; c++)
        if (cabs(X[beta][c]) > 0.5) gam = c;
      t[n].a = alpha; t[n].b = beta; t[n].d = delta; t[n].g = gam;
      t[n].v = X[delta][alpha] * X[beta][gam];
      n++;
    }
}

static void meson_terms(int op, mterm t[16]);
int orc_meson_table(int op, int *idx, double *val) {
  mterm t[16];
  if (op < 0 || op > 15) return -1;
  meson_terms(op, t);
  for (int i = 0; i < 16; i++) {
    idx[i * 4 + 0] = t[i].b; idx[i * 4 + 1] = t[i].g;
    idx[i * 4 + 2] = t[i].d; idx[i * 4 + 3] = t[i].a;
    val[i * 2 + 0] = creal(t[i].v); val[i * 2 + 1] = cimag(t[i].v);
  }
  return 16;
}

/* diquark tables: entries (nu,lu,lup,nup; value) with                */
/* value = A[nu][lu] * Abar[lup][nup], A = C Gamma, Abar = g4 A^dag g4 */
/* (PLEGMA_baryons_NtoN.cu:3-4 for Cg5, PLEGMA_baryons_deltas.cu:3-4   */
/* for Cg1..Cg3).  C = i g2 g4.                                        */
typedef struct {
  int nu, lu, lup, nup;
  cplx v;
} dterm;

static int diquark_terms_build(int which, dterm t[16]) {
  cplx C[4][4], A[4][4], Ad[4][4], Ab[4][4];
  mat_mul(GM[2], GM[4], C);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) C[i][j] *= I;
  int op = which == 0 ? 5 : which; /* Cg5 or Cg_k */
  mat_mul(C, GM[op], A);
  for (int i = 0; i < 4; i++)
    for (int j = 0; j < 4; j++) Ad[i][j] = conj(A[j][i]);
  mat_mul(GM[4], Ad, Ab);
  mat_mul(Ab, GM[4], Ab);
  int n = 0;
  for (int nu = 0; nu < 4; nu++)
    for (int lu = 0; lu < 4; lu++) {
      if (cabs(A[nu][lu]) < 0.5) continue;
      for (int lup = 0; lup < 4; lup++)
        for (int nup = 0; nup < 4; nup++) {
          if (cabs(Ab[lup][nup]) < 0.5) continue;
          t[n].nu = nu; t[n].lu = lu; t[n].lup = lup; t[n].nup = nup;
          t[n].v = A[nu][lu] * Ab[lup][nup];
          n++;
        }
    }
  return n;
}

static mterm MT[16][16];
static dterm DT[4][16];
static pthread_once_t tab_once = PTHREAD_ONCE_INIT;
static void build_tables_once(void) {
  build_gammas();
  for (int op = 0; op < 16; op++) meson_terms_build(op, MT[op]);
  for (int w = 0; w < 4; w++) diquark_terms_build(w, DT[w]);
}
static void build_tables(void) { pthread_once(&tab_once, build_tables_once); }
static void meson_terms(int op, mterm t[16]) {
  build_tables();
  memcpy(t, MT[op], sizeof(MT[op]));
}
static int diquark_terms(int which, dterm t[16]) {
  build_tables();
  memcpy(t, DT[which], sizeof(DT[which]));
  return 16;
}

int orc_diquark_table(int which, int *idx, double *val) {
  dterm t[16];
  if (which < 0 || which > 3) return -1;
  int n = diquark_terms(which, t);
  for (int i = 0; i < n; i++) {
    idx[i * 4 + 0] = t[i].nu; idx[i * 4 + 1] = t[i].lu;
    idx[i * 4 + 2] = t[i].lup; idx[i * 4 + 3] = t[i].nup;
    val[i * 2 + 0] = creal(t[i].v); val[i * 2 + 1] = cimag(t[i].v);
  }
  return n;
}

/* epsilon permutations, lib/kernels/PLEGMA_kernel_utils.cuh:49-56 */
static const int EPS[6][3] = {{0, 1, 2}, {2, 0, 1}, {1, 2, 0},
                              {2, 1, 0}, {0, 2, 1}, {1, 0, 2}};
static const double SGN_EPS[6] = {+1, +1, +1, -1, -1, -1};

/* site accessor: S[mu][nu][a][b] of a 144-complex site block */
#define PS(S, mu, nu, a, b) ((S)[(((mu)*NS + (nu)) * NC + (a)) * NC + (b)])

static void load_site(const double *src, cplx *S) {
  for (int i = 0; i < NPROP; i++) S[i] = src[2 * i] + I * src[2 * i + 1];
}
static void store(cplx v, double *out) {
  out[0] = creal(v);
  out[1] = cimag(v);
}

/* ------------------------------------------------------------------ */
/* meson bodies                                                        */
/* ------------------------------------------------------------------ */
/* channel order of contractMesons (PLEGMA_Correlator.cu:47-50):
 * pseudoscalar, scalar, g5g1, g5g2, g5g3, g5g4, g1, g2, g3, g4        */
static const int MESON10_OPS[10] = {5, 0, 6, 7, 8, 9, 1, 2, 3, 4};

static cplx meson_channel(int op, const cplx *A, const cplx *B) {
  mterm t[16];
  meson_terms(op, t);
  cplx acc = 0;
  for (int is = 0; is < 16; is++)
    for (int a = 0; a < NC; a++)
      for (int b = 0; b < NC; b++)
        acc += t[is].v * PS(A, t[is].a, t[is].b, a, b) * conj(PS(B, t[is].d, t[is].g, a, b));
  return acc;
}

/* PLEGMA_mesons.cuh:25-52: accum[ip] from S1, accum[10+ip] from S2 */
void orc_site_mesons10(const double *S1r, const double *S2r, double *out) {
  cplx S1[NPROP], S2[NPROP];
  load_site(S1r, S1);
  load_site(S2r, S2);
  for (int ip = 0; ip < 10; ip++) {
    store(meson_channel(MESON10_OPS[ip], S1, S1), out + 2 * ip);
    store(meson_channel(MESON10_OPS[ip], S2, S2), out + 2 * (10 + ip));
  }
}

/* same formula over the 16 GAMMAS in enum order */
void orc_site_mesons16(const double *S1r, const double *S2r, double *out) {
  cplx S1[NPROP], S2[NPROP];
  load_site(S1r, S1);
  load_site(S2r, S2);
  for (int ip = 0; ip < 16; ip++) {
    store(meson_channel(ip, S1, S1), out + 2 * ip);
    store(meson_channel(ip, S2, S2), out + 2 * (16 + ip));
  }
}

/* PLEGMA_mesonsNew.cuh:25-48: S1 . conj(S2) */
void orc_site_mesons_new(const double *S1r, const double *S2r, double *out) {
  cplx S1[NPROP], S2[NPROP];
  load_site(S1r, S1);
  load_site(S2r, S2);
  for (int ip = 0; ip < 10; ip++) store(meson_channel(MESON10_OPS[ip], S1, S2), out + 2 * ip);
}

/* PLEGMA_mesonsAll.cuh:27-83: PS, S, then Tr[Gmu S1 Gnu S2^dag] for
 * Gmu,Gnu in {g1..g4} (p=0) and {g5g1..g5g4} (p=1),
 * channel 2 + ip1*4 + ip2 + 16p */
void orc_site_mesons_all(const double *S1r, const double *S2r, double *out) {
  cplx S1[NPROP], S2[NPROP];
  load_site(S1r, S1);
  load_site(S2r, S2);
  build_gammas();
  for (int ip = 0; ip < 2; ip++) {
    store(meson_channel(MESON10_OPS[ip], S1, S2), out + 2 * ip);
    for (int ip1 = 0; ip1 < 4; ip1++)
      for (int ip2 = 0; ip2 < 4; ip2++) {
        int op1 = 1 + ip1 + 5 * ip, op2 = 1 + ip2 + 5 * ip;
        cplx acc = 0;
        for (int delta = 0; delta < 4; delta++)
          for (int alpha = 0; alpha < 4; alpha++) {
            cplx v1 = GM[op1][delta][alpha];
            if (cabs(v1) < 0.5) continue;
            for (int beta = 0; beta < 4; beta++)
              for (int zeta = 0; zeta < 4; zeta++) {
                cplx v2 = GM[op2][beta][zeta];
                if (cabs(v2) < 0.5) continue;
                for (int a = 0; a < NC; a++)
                  for (int b = 0; b < NC; b++)
                    acc += v1 * PS(S1, alpha, beta, a, b) * v2 * conj(PS(S2, delta, zeta, a, b));
              }
          }
        store(acc, out + 2 * (2 + ip1 * 4 + ip2 + 16 * ip));
      }
  }
}

/* ------------------------------------------------------------------ */
/* baryon bodies                                                       */
/* ------------------------------------------------------------------ */
/* PLEGMA_baryons_NtoN.cu:12-45 */
static void nton_body(const cplx *P1, const cplx *P2, cplx acc[16]) {
  dterm t[16];
  diquark_terms(0, t);
  for (int mu = 0; mu < 4; mu++)
    for (int mup = 0; mup < 4; mup++) {
      cplx tmp = 0;
      for (int idx = 0; idx < 16; idx++) {
        int nu = t[idx].nu, lu = t[idx].lu, lup = t[idx].lup, nup = t[idx].nup;
        for (int cc1 = 0; cc1 < 6; cc1++) {
          int a = EPS[cc1][0], b = EPS[cc1][1], c = EPS[cc1][2];
          for (int cc2 = 0; cc2 < 6; cc2++) {
            int a1 = EPS[cc2][0], b1 = EPS[cc2][1], c1 = EPS[cc2][2];
            cplx factor = SGN_EPS[cc1] * SGN_EPS[cc2] * t[idx].v;
            tmp += factor * PS(P2, lu, lup, b, b1) *
                   (PS(P1, nu, nup, a, a1) * PS(P1, mu, mup, c, c1) -
                    PS(P1, nu, mup, a, c1) * PS(P1, mu, nup, c, a1));
          }
        }
      }
      acc[mu * 4 + mup] = tmp;
    }
}

void orc_site_nton(const double *S1r, const double *S2r, double *out) {
  cplx S1[NPROP], S2[NPROP], acc[16];
  load_site(S1r, S1);
  load_site(S2r, S2);
  nton_body(S1, S2, acc);
  for (int i = 0; i < 16; i++) store(acc[i], out + 2 * i);
  nton_body(S2, S1, acc);
  for (int i = 0; i < 16; i++) store(acc[i], out + 2 * (16 + i));
}

/* PLEGMA_baryons_deltas.cu:62-109 (iso 3/2: all three quarks from P) */
static void delta32_body(int k, const cplx *P, cplx acc[16]) {
  dterm t[16];
  diquark_terms(k + 1, t);
  for (int mu = 0; mu < 4; mu++)
    for (int mup = 0; mup < 4; mup++) {
      cplx tmp = 0;
      for (int idx = 0; idx < 16; idx++) {
        int nu = t[idx].nu, lu = t[idx].lu, lup = t[idx].lup, nup = t[idx].nup;
        for (int cc1 = 0; cc1 < 6; cc1++) {
          int a = EPS[cc1][0], b = EPS[cc1][1], c = EPS[cc1][2];
          for (int cc2 = 0; cc2 < 6; cc2++) {
            int a1 = EPS[cc2][0], b1 = EPS[cc2][1], c1 = EPS[cc2][2];
            cplx factor = SGN_EPS[cc1] * SGN_EPS[cc2] * t[idx].v;
            tmp += factor * (PS(P, nu, lup, a, b1) * PS(P, lu, mup, b, c1) * PS(P, mu, nup, c, a1)
                           - PS(P, nu, mup, a, c1) * PS(P, lu, lup, b, b1) * PS(P, mu, nup, c, a1)
                           + PS(P, nu, mup, a, c1) * PS(P, lu, nup, b, a1) * PS(P, mu, lup, c, b1)
                           - PS(P, nu, nup, a, a1) * PS(P, lu, mup, b, c1) * PS(P, mu, lup, c, b1)
                           - PS(P, nu, lup, a, b1) * PS(P, lu, nup, b, a1) * PS(P, mu, mup, c, c1)
                           + PS(P, nu, nup, a, a1) * PS(P, lu, lup, b, b1) * PS(P, mu, mup, c, c1));
          }
        }
      }
      acc[mu * 4 + mup] = tmp;
    }
}

void orc_site_delta32(int k, const double *S1r, const double *S2r, double *out) {
  cplx S1[NPROP], S2[NPROP], acc[16];
  load_site(S1r, S1);
  load_site(S2r, S2);
  delta32_body(k, S1, acc);
  for (int i = 0; i < 16; i++) store(acc[i], out + 2 * i);
  delta32_body(k, S2, acc);
  for (int i = 0; i < 16; i++) store(acc[i], out + 2 * (16 + i));
}

/* PLEGMA_baryons_deltas.cu:8-58 (iso 1/2: uud with P1 doubly represented) */
static void delta12_body(int k, const cplx *P1, const cplx *P2, cplx acc[16]) {
  dterm t[16];
  diquark_terms(k + 1, t);
  for (int mu = 0; mu < 4; mu++)
    for (int mup = 0; mup < 4; mup++) {
      cplx tmp = 0;
      for (int idx = 0; idx < 16; idx++) {
        int nu = t[idx].nu, lu = t[idx].lu, lup = t[idx].lup, nup = t[idx].nup;
        for (int cc1 = 0; cc1 < 6; cc1++) {
          int a = EPS[cc1][0], b = EPS[cc1][1], c = EPS[cc1][2];
          for (int cc2 = 0; cc2 < 6; cc2++) {
            int a1 = EPS[cc2][0], b1 = EPS[cc2][1], c1 = EPS[cc2][2];
            cplx factor = (1. / 3.) * SGN_EPS[cc1] * SGN_EPS[cc2] * t[idx].v;
            tmp += factor * (-4. * PS(P1, nu, mup, a, c1) * PS(P2, lu, lup, b, b1) * PS(P1, mu, nup, c, a1)
                             + 2. * PS(P1, nu, lup, a, b1) * PS(P2, lu, mup, b, c1) * PS(P1, mu, nup, c, a1)
                             + 2. * PS(P1, nu, mup, a, c1) * PS(P1, lu, nup, b, a1) * PS(P2, mu, lup, c, b1)
                             - 2. * PS(P1, nu, nup, a, a1) * PS(P1, lu, mup, b, c1) * PS(P2, mu, lup, c, b1)
                             - 2. * PS(P1, nu, nup, a, a1) * PS(P2, lu, mup, b, c1) * PS(P1, mu, lup, c, b1)
                             - 1. * PS(P1, nu, lup, a, b1) * PS(P1, lu, nup, b, a1) * PS(P2, mu, mup, c, c1)
                             + 1. * PS(P1, nu, nup, a, a1) * PS(P1, lu, lup, b, b1) * PS(P2, mu, mup, c, c1)
                             + 4. * PS(P1, nu, nup, a, a1) * PS(P2, lu, lup, b, b1) * PS(P1, mu, mup, c, c1));
          }
        }
      }
      acc[mu * 4 + mup] = tmp;
    }
}

void orc_site_delta12(int k, const double *S1r, const double *S2r, double *out) {
  cplx S1[NPROP], S2[NPROP], acc[16];
  load_site(S1r, S1);
  load_site(S2r, S2);
  delta12_body(k, S1, S2, acc);
  for (int i = 0; i < 16; i++) store(acc[i], out + 2 * i);
  delta12_body(k, S2, S1, acc);
  for (int i = 0; i < 16; i++) store(acc[i], out + 2 * (16 + i));
}

/* ------------------------------------------------------------------ */
/* four-propagator meson-meson block (no single reference function;    */
/* counterpart of lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:5-61). */
/* Defined by formula on 12x12 spin-colour matrices:                   */
/*   A_G = G S_L g5 G S_H^dag g5                                       */
/*   meson(G)      = Tr A_G                 (16 GAMMAS, enum order)    */
/*   mesmes(G)     = (Tr A_G)^2 - Tr[A_G A_G]  for G in g5,g1,g2,g3   */
/* ------------------------------------------------------------------ */
#define NMM_MESON 16
#define NMM_MM 4
static const int MM_OPS[NMM_MM] = {5, 1, 2, 3};

static void amat(int op, const cplx *SL, const cplx *SH, cplx A[12][12]) {
  build_gammas();
  cplx X[4][4];
  mat_mul(GM[5], GM[op], X); /* g5 G */
  /* A = G SL (g5 G) SH^dag g5 ; index (spin,colour) -> spin*3+colour */
  static __thread cplx T1[12][12], T2[12][12];
  /* T1 = G SL */
  for (int mu = 0; mu < 4; mu++)
    for (int a = 0; a < 3; a++)
      for (int nu = 0; nu < 4; nu++)
        for (int b = 0; b < 3; b++) {
          cplx s = 0;
          for (int k = 0; k < 4; k++) s += GM[op][mu][k] * PS(SL, k, nu, a, b);
          T1[mu * 3 + a][nu * 3 + b] = s;
        }
  /* T2 = (g5 G) SH^dag g5 : SH^dag[(nu,b),(rho,c)] = conj(SH[rho][nu][c][b]) */
  for (int nu = 0; nu < 4; nu++)
    for (int b = 0; b < 3; b++)
      for (int rho = 0; rho < 4; rho++)
        for (int c = 0; c < 3; c++) {
          cplx s = 0;
          for (int k = 0; k < 4; k++)
            for (int l = 0; l < 4; l++) {
              cplx g5lr = GM[5][l][rho];
              if (cabs(g5lr) < 0.5 || cabs(X[nu][k]) < 0.5) continue;
              s += X[nu][k] * conj(PS(SH, l, k, c, b)) * g5lr;
            }
          T2[nu * 3 + b][rho * 3 + c] = s;
        }
  for (int i = 0; i < 12; i++)
    for (int j = 0; j < 12; j++) {
      cplx s = 0;
      for (int k = 0; k < 12; k++) s += T1[i][k] * T2[k][j];
      A[i][j] = s;
    }
}

void orc_site_meson_meson(const double *SLr, const double *SHr, double *out) {
  cplx SL[NPROP], SH[NPROP];
  cplx A[12][12];
  load_site(SLr, SL);
  load_site(SHr, SH);
  for (int op = 0; op < 16; op++) {
    amat(op, SL, SH, A);
    cplx tr = 0;
    for (int i = 0; i < 12; i++) tr += A[i][i];
    store(tr, out + 2 * op);
    for (int k = 0; k < NMM_MM; k++)
      if (MM_OPS[k] == op) {
        cplx tr2 = 0;
        for (int i = 0; i < 12; i++)
          for (int j = 0; j < 12; j++) tr2 += A[i][j] * A[j][i];
        store(tr * tr - tr2, out + 2 * (NMM_MESON + k));
      }
  }
}

/* ------------------------------------------------------------------ */
/* momenta, windows, layout                                            */
/* ------------------------------------------------------------------ */
/* lib/PLEGMA_FT.cu:50-72 (3D) */
int orc_momlist(int q2max, int *out) {
  int n = 0;
  for (int iQ = 0; iQ <= q2max; iQ++)
    for (int nx = iQ; nx >= -iQ; nx--)
      for (int ny = iQ; ny >= -iQ; ny--)
        for (int nz = iQ; nz >= -iQ; nz--)
          if (nx * nx + ny * ny + nz * nz == iQ) {
            if (out) { out[3 * n] = nx; out[3 * n + 1] = ny; out[3 * n + 2] = nz; }
            n++;
          }
  return n;
}
int orc_nmom(int q2max) { return q2max < 0 ? -1 : orc_momlist(q2max, NULL); }

/* include/PLEGMA_Correlator.h:53-79 with one process (procPosition = 0) */
void orc_window(const int L[4], const int src[4], int totalT, int *startT,
                int *endT, int *localT) {
  int T = L[3];
  int start = (T - src[3]) % T;
  if (start >= totalT) start = 0;
  int end = start > 0 ? (totalT < T + start ? totalT : T + start) : 0;
  int ts = src[3] % T;
  int lt = end - start + (totalT < T - ts ? totalT : T - ts);
  *startT = start;
  *endT = end;
  *localT = lt;
}

/* PLEGMA_mesons.cuh:22: t = j; if (t >= maxT) t = t0 % L_T + t - maxT */
int orc_row_time(const int L[4], const int src[4], int totalT, int j) {
  int s, e, lt;
  orc_window(L, src, totalT, &s, &e, &lt);
  int maxT = e - s;
  int t = j;
  if (t >= maxT) t = (src[3] % L[3]) + t - maxT;
  return t;
}

int orc_kind_nch(int kind, int mask, int *nd, int *ng) {
  switch (kind) {
  case ORC_MESONS: *nd = 2; *ng = 1; return 10;
  case ORC_MESONS16: *nd = 2; *ng = 1; return 16;
  case ORC_MESONS_NEW: *nd = 1; *ng = 1; return 10;
  case ORC_MESONS_ALL: *nd = 1; *ng = 1; return 34;
  case ORC_BARYONS: {
    int g = 0;
    for (int b = 0; b < 10; b++)
      if (mask & (1 << b)) g++;
    *nd = 2; *ng = g; return 16;
  }
  case ORC_MESON_MESON: *nd = 1; *ng = 1; return NMM_MESON + NMM_MM;
  case ORC_TWOPT: {
    int g = 1;
    for (int b = 0; b < 10; b++)
      if (mask & (1 << b)) g++;
    *nd = 1; *ng = g; return 16;
  }
  }
  return -1;
}

static const double *site_ptr(const orc_prop *p, int T, int64_t V3, int t, int64_t x,
                              int comp) {
  int plane = ((t - p->t_first) % T + T) % T;
  if (plane >= p->nt) return NULL;
  return p->data + ((int64_t)comp * p->comp_stride + (int64_t)plane * V3 + x) * 2;
}

static int gather_site(const orc_prop *p, int T, int64_t V3, int t, int64_t x, double *S) {
  for (int i = 0; i < NPROP; i++) {
    const double *q = site_ptr(p, T, V3, t, x, i);
    if (!q) return -1;
    S[2 * i] = q[0];
    S[2 * i + 1] = q[1];
  }
  return 0;
}

/* per-site accum[] of the reference kernel for this kind:
 * dataset-major, then group, then channel */
static int site_accum(int kind, int mask, const double *S1, const double *S2, double *acc) {
  switch (kind) {
  case ORC_MESONS: orc_site_mesons10(S1, S2, acc); return 20;
  case ORC_MESONS16: orc_site_mesons16(S1, S2, acc); return 32;
  case ORC_MESONS_NEW: orc_site_mesons_new(S1, S2, acc); return 10;
  case ORC_MESONS_ALL: orc_site_mesons_all(S1, S2, acc); return 34;
  case ORC_MESON_MESON: orc_site_meson_meson(S1, S2, acc); return NMM_MESON + NMM_MM;
  case ORC_TWOPT: {
    /* the reference bodies on one propagator, first dataset only:
     * group 0 = mesons (GAMMAS order), then each selected baryon group */
    cplx S[NPROP], r[16];
    load_site(S1, S);
    for (int ip = 0; ip < 16; ip++) store(meson_channel(ip, S, S), acc + 2 * ip);
    int g = 1;
    for (int b = 0; b < 10; b++) {
      if (!(mask & (1 << b))) continue;
      if (b == 0) nton_body(S, S, r);
      else if (b >= 4 && b <= 6) delta12_body(b - 4, S, S, r);
      else if (b >= 7 && b <= 9) delta32_body(b - 7, S, r);
      else return -1;
      for (int c = 0; c < 16; c++) store(r[c], acc + 2 * (16 * g + c));
      g++;
    }
    return 16 * g;
  }
  case ORC_BARYONS: {
    /* reference runs one pass per group (PLEGMA_baryons.cuh:192-203);
     * accum per group is [dataset][16].  Here: acc[(d*ng + g)*16 + c]. */
    int nd, ng;
    orc_kind_nch(kind, mask, &nd, &ng);
    double tmp[64];
    int g = 0;
    for (int b = 0; b < 10; b++) {
      if (!(mask & (1 << b))) continue;
      if (b == 0) orc_site_nton(S1, S2, tmp);
      else if (b >= 4 && b <= 6) orc_site_delta12(b - 4, S1, S2, tmp);
      else if (b >= 7 && b <= 9) orc_site_delta32(b - 7, S1, S2, tmp);
      else return -1; /* nucl2 interpolators are not part of this path */
      for (int d = 0; d < 2; d++)
        for (int c = 0; c < 16; c++) {
          acc[((d * ng + g) * 16 + c) * 2] = tmp[(d * 16 + c) * 2];
          acc[((d * ng + g) * 16 + c) * 2 + 1] = tmp[(d * 16 + c) * 2 + 1];
        }
      g++;
    }
    return 2 * ng * 16;
  }
  }
  return -1;
}

int orc_site_correlator(int kind, int mask, const int L[4], int t, const orc_prop *p1,
                        const orc_prop *p2, double *out, int nthreads) {
  int nd, ng;
  int nch = orc_kind_nch(kind, mask, &nd, &ng);
  if (nch < 0) return -1;
  int ntot = nd * ng * nch;
  int64_t V3 = (int64_t)L[0] * L[1] * L[2];
  int T = L[3];
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(| : err)
#endif
  for (int64_t x = 0; x < V3; x++) {
    double S1[2 * NPROP], S2[2 * NPROP];
    if (gather_site(p1, T, V3, t, x, S1) || gather_site(p2, T, V3, t, x, S2)) {
      err |= 1;
      continue;
    }
    site_accum(kind, mask, S1, S2, out + x * ntot * 2);
  }
  return err ? -2 : 0;
}

/* PLEGMA_kernel_utils.cuh:820-846 phase, sign -1, x - x0 not reduced */
static void phase_of(const int L[4], const int src3[3], const int *mom, int64_t x,
                     double *c, double *s) {
  int id[3] = {(int)(x % L[0]), (int)((x / L[0]) % L[1]), (int)((x / L[0] / L[1]) % L[2])};
  for (int i = 0; i < 3; i++) id[i] -= src3[i];
  double phase = mom[0] * id[0] / ((double)L[0]) + mom[1] * id[1] / ((double)L[1]) +
                 mom[2] * id[2] / ((double)L[2]);
  phase *= 2. * REF_PI;
  *c = cos(phase);
  *s = -1. * sin(phase);
}

void orc_dft(const int L[4], const int src3[3], int q2max, const double *f, int ncomp,
             double *out, int nthreads) {
  int nmom = orc_nmom(q2max);
  int *moms = malloc(sizeof(int) * 3 * nmom);
  orc_momlist(q2max, moms);
  int64_t V3 = (int64_t)L[0] * L[1] * L[2];
  memset(out, 0, sizeof(double) * 2 * nmom * ncomp);
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel
#endif
  {
    double *part = calloc((size_t)2 * nmom * ncomp, sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int64_t x = 0; x < V3; x++)
      for (int m = 0; m < nmom; m++) {
        double c, s;
        phase_of(L, src3, moms + 3 * m, x, &c, &s);
        for (int ip = 0; ip < ncomp; ip++) {
          double fr = f[(x * ncomp + ip) * 2], fi = f[(x * ncomp + ip) * 2 + 1];
          part[(m * ncomp + ip) * 2] += fr * c - fi * s;
          part[(m * ncomp + ip) * 2 + 1] += fr * s + fi * c;
        }
      }
#ifdef _OPENMP
#pragma omp critical
#endif
    for (int i = 0; i < 2 * nmom * ncomp; i++) out[i] += part[i];
    free(part);
  }
  free(moms);
}

/* Correlator driver: per row j, per site the kernel body then the naive
 * DFT (fused as in the reference kernel), summed over sites in FP64. */
int orc_correlator(int kind, int mask, const int L[4], const int src[4], int totalT,
                   int q2max, const orc_prop *p1, const orc_prop *p2, double *H,
                   const int *rows, int nrows, int nthreads) {
  int nd, ng;
  int nch = orc_kind_nch(kind, mask, &nd, &ng);
  if (nch < 0 || q2max < 0) return -1;
  int startT, endT, localT;
  orc_window(L, src, totalT, &startT, &endT, &localT);
  int nmom = orc_nmom(q2max);
  int ntot = nd * ng * nch;
  int64_t V3 = (int64_t)L[0] * L[1] * L[2];
  int *moms = malloc(sizeof(int) * 3 * nmom);
  orc_momlist(q2max, moms);
  int nr = nrows > 0 ? nrows : localT;
  double *red = malloc(sizeof(double) * 2 * nmom * ntot);
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
  for (int r = 0; r < nr && !err; r++) {
    int j = nrows > 0 ? rows[r] : r;
    if (j < 0 || j >= localT) { err = -3; break; }
    int t = orc_row_time(L, src, totalT, j);
    memset(red, 0, sizeof(double) * 2 * nmom * ntot);
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
      double *part = calloc((size_t)2 * nmom * ntot, sizeof(double));
      double S1[2 * NPROP], S2[2 * NPROP], acc[2 * 128];
      int lerr = 0;
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
      for (int64_t x = 0; x < V3; x++) {
        if (gather_site(p1, L[3], V3, t, x, S1) || gather_site(p2, L[3], V3, t, x, S2)) {
          lerr = 1;
          continue;
        }
        site_accum(kind, mask, S1, S2, acc);
        for (int m = 0; m < nmom; m++) {
          double c, s;
          phase_of(L, src, moms + 3 * m, x, &c, &s);
          for (int ip = 0; ip < ntot; ip++) {
            part[(m * ntot + ip) * 2] += acc[2 * ip] * c - acc[2 * ip + 1] * s;
            part[(m * ntot + ip) * 2 + 1] += acc[2 * ip] * s + acc[2 * ip + 1] * c;
          }
        }
      }
#ifdef _OPENMP
#pragma omp critical
#endif
      {
        for (int i = 0; i < 2 * nmom * ntot; i++) red[i] += part[i];
        if (lerr) err = -2;
      }
      free(part);
    }
    /* H layout: [g][d][j][m][c]; accum index ip = (d*ng + g)*nch + c */
    for (int d = 0; d < nd; d++)
      for (int g = 0; g < ng; g++)
        for (int m = 0; m < nmom; m++)
          for (int c = 0; c < nch; c++) {
            int ip = (d * ng + g) * nch + c;
            int64_t h = ((((int64_t)g * nd + d) * localT + j) * nmom + m) * nch + c;
            H[2 * h] = red[(m * ntot + ip) * 2];
            H[2 * h + 1] = red[(m * ntot + ip) * 2 + 1];
          }
  }
  free(red);
  free(moms);
  return err;
}

/* ------------------------------------------------------------------ */
/* synthetic propagators: Philox4x32-10 (Salmon et al., SC'11)         */
/* ------------------------------------------------------------------ */
void orc_philox(const uint32_t ctr_in[4], const uint32_t key_in[2], uint32_t out[4]) {
  uint32_t ctr[4] = {ctr_in[0], ctr_in[1], ctr_in[2], ctr_in[3]};
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u;
  const uint32_t W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key_in[0], k1 = key_in[1];
  for (int r = 0; r < 10; r++) {
    uint64_t p0 = (uint64_t)M0 * c0, p1 = (uint64_t)M1 * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    k0 += W0; k1 += W1;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

void orc_normal_pair(uint64_t seed, uint32_t stream, uint32_t site, uint32_t comp,
                     double *re, double *im) {
  uint32_t ctr[4] = {site, comp, stream, 0x504C4731u};
  uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
  uint32_t r[4];
  orc_philox(ctr, key, r);
  uint64_t a = (((uint64_t)r[0] << 32) | r[1]) >> 11;
  uint64_t b = (((uint64_t)r[2] << 32) | r[3]) >> 11;
  double u1 = ((double)a + 0.5) * (1.0 / 9007199254740992.0);
  double u2 = ((double)b + 0.5) * (1.0 / 9007199254740992.0);
  double rad = sqrt(-2.0 * log(u1));
  double th = 2.0 * REF_PI * u2;
  *re = rad * cos(th);
  *im = rad * sin(th);
}

void orc_synth_prop(const int L[4], uint64_t seed, uint32_t stream, int t0, double lambda,
                    double *data, int64_t comp_stride, int t_first, int nt, int nthreads) {
  int64_t V3 = (int64_t)L[0] * L[1] * L[2];
  int T = L[3];
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) collapse(2)
#endif
  for (int pl = 0; pl < nt; pl++)
    for (int64_t x = 0; x < V3; x++) {
      int t = (t_first + pl) % T;
      double scale = exp(-lambda * fabs((double)(t - t0)));
      uint32_t site = (uint32_t)(x + V3 * t);
      for (int c = 0; c < NPROP; c++) {
        double re, im;
        orc_normal_pair(seed, stream, site, (uint32_t)c, &re, &im);
        int64_t o = ((int64_t)c * comp_stride + (int64_t)pl * V3 + x) * 2;
        data[o] = re * scale;
        data[o + 1] = im * scale;
      }
    }
}

/* ------------------------------------------------------------------ propagator preparation */
/* PLEGMA_propagator_utils.cuh:103-118: per colour pair (c1, c2),
 * PT[mu][nu] = 0.5 (P[mu][nu] - P[mu+2][nu+2] + (i sign) (P[mu][nu+2] + P[mu+2][nu])),
 * spin indices mod 4; the product with i sign is the complex product of
 * Float2 (re = 0 re - sign im, im = 0 im + sign re). */
void orc_rotate_site(const double *S, int sign, double *out) {
  for (int c = 0; c < 9; c++)
    for (int mu = 0; mu < 4; mu++)
      for (int nu = 0; nu < 4; nu++) {
        const double *p = S + 2 * ((mu * 4 + nu) * 9 + c);
        const double *q = S + 2 * ((((mu + 2) % 4) * 4 + (nu + 2) % 4) * 9 + c);
        const double *u = S + 2 * ((mu * 4 + (nu + 2) % 4) * 9 + c);
        const double *v = S + 2 * ((((mu + 2) % 4) * 4 + nu) * 9 + c);
        const double sr = u[0] + v[0], si = u[1] + v[1];
        const double tr = 0.0 * sr - sign * si, ti = 0.0 * si + sign * sr;
        double *o = out + 2 * ((mu * 4 + nu) * 9 + c);
        o[0] = 0.5 * ((p[0] - q[0]) + tr);
        o[1] = 0.5 * ((p[1] - q[1]) + ti);
      }
}

void orc_prop_rotate(double *data, int64_t comp_stride, int nt, int64_t V3, int sign) {
#pragma omp parallel for schedule(static)
  for (int64_t e = 0; e < (int64_t)nt * V3; e++) {
    double S[2 * NPROP], R[2 * NPROP];
    for (int i = 0; i < NPROP; i++) {
      S[2 * i] = data[2 * (i * comp_stride + e)];
      S[2 * i + 1] = data[2 * (i * comp_stride + e) + 1];
    }
    orc_rotate_site(S, sign, R);
    for (int i = 0; i < NPROP; i++) {
      data[2 * (i * comp_stride + e)] = R[2 * i];
      data[2 * (i * comp_stride + e) + 1] = R[2 * i + 1];
    }
  }
}

/* PLEGMA_propagator_utils.cuh:64-80: if (t < t0) every component *= -1 */
void orc_prop_boundaries(double *data, int64_t comp_stride, int t_first, int nt, int T, int64_t V3, int t0) {
  for (int pl = 0; pl < nt; pl++) {
    if ((t_first + pl) % T >= t0) continue;
    for (int i = 0; i < NPROP; i++)
      for (int64_t x = 0; x < V3; x++) {
        double *v = data + 2 * (i * comp_stride + (int64_t)pl * V3 + x);
        v[0] *= -1.;
        v[1] *= -1.;
      }
  }
}

/* ------------------------------------------------------------------ */
/* udsc baryons: lib/kernels/PLEGMA_baryons_udsc.cu:86-121             */
/*   O_k = sum_t v_t sum_{cc1,cc2} sgn sgn' S1[m0][m1][a][a1]          */
/*         S2[m2][m3][b][b1] S3[m4][m5][c][c1]                         */
/* with (a,b,c) = eps[cc1], (a1,b1,c1) = eps[cc2]; terms of operator k */
/* are idx[6*t..], val[2*t..] for t in [first[k], first[k]+count[k]).  */
/* ------------------------------------------------------------------ */
void orc_site_udsc(const double *S1r, const double *S2r, const double *S3r, const int *idx, const double *val,
                   const int *first, const int *count, int nops, double *out) {
  cplx S1[NPROP], S2[NPROP], S3[NPROP];
  load_site(S1r, S1);
  load_site(S2r, S2);
  load_site(S3r, S3);
  for (int k = 0; k < nops; k++) {
    cplx acc = 0;
    for (int t = first[k]; t < first[k] + count[k]; t++) {
      const int *mu = idx + 6 * t;
      const cplx v = val[2 * t] + I * val[2 * t + 1];
      for (int cc1 = 0; cc1 < 6; cc1++) {
        const int a = EPS[cc1][0], b = EPS[cc1][1], c = EPS[cc1][2];
        for (int cc2 = 0; cc2 < 6; cc2++) {
          const int a1 = EPS[cc2][0], b1 = EPS[cc2][1], c1 = EPS[cc2][2];
          acc += SGN_EPS[cc1] * SGN_EPS[cc2] * v * PS(S1, mu[0], mu[1], a, a1) * PS(S2, mu[2], mu[3], b, b1) *
                 PS(S3, mu[4], mu[5], c, c1);
        }
      }
    }
    store(acc, out + 2 * k);
  }
}

int orc_udsc_field(const int L[4], int t, const orc_prop *p1, const orc_prop *p2, const orc_prop *p3,
                   const int *idx, const double *val, const int *first, const int *count, int nops, double *out,
                   int nthreads) {
  int64_t V3 = (int64_t)L[0] * L[1] * L[2];
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(| : err)
#endif
  for (int64_t x = 0; x < V3; x++) {
    double S1[2 * NPROP], S2[2 * NPROP], S3[2 * NPROP];
    if (gather_site(p1, L[3], V3, t, x, S1) || gather_site(p2, L[3], V3, t, x, S2) ||
        gather_site(p3, L[3], V3, t, x, S3)) {
      err |= 1;
      continue;
    }
    orc_site_udsc(S1, S2, S3, idx, val, first, count, nops, out + x * nops * 2);
  }
  return err ? -2 : 0;
}

/* ------------------------------------------------------------------ */
/* four-propagator tables: lib/kernels/PLEGMA_heavy_light_tetraquarks */
/* .cu:26-49.  Term t of operator k (terms [first[k], first[k] +       */
/* count[k])): spins mu = idx[8t..], colour slots c = col[8t..],       */
/*   v sum_{a0..a3} P1[mu0][mu1][a[c0]][a[c1]] P2[mu2][mu3][a[c2]][a[c3]] */
/*   conj(P3[(mu5+2)%4][(mu4+2)%4][a[c5]][a[c4]])                       */
/*   conj(P4[(mu7+2)%4][(mu6+2)%4][a[c7]][a[c6]])                       */
/* ------------------------------------------------------------------ */
void orc_site_tetra(const double *S1r, const double *S2r, const double *S3r, const double *S4r, const int *idx,
                    const int *col, const double *val, const int *first, const int *count, int nops, double *out) {
  cplx S1[NPROP], S2[NPROP], S3[NPROP], S4[NPROP];
  load_site(S1r, S1);
  load_site(S2r, S2);
  load_site(S3r, S3);
  load_site(S4r, S4);
  for (int k = 0; k < nops; k++) {
    cplx acc = 0;
    for (int t = first[k]; t < first[k] + count[k]; t++) {
      const int *mu = idx + 8 * t, *c = col + 8 * t;
      const cplx v = val[2 * t] + I * val[2 * t + 1];
      int a[4];
      for (a[0] = 0; a[0] < 3; a[0]++)
        for (a[1] = 0; a[1] < 3; a[1]++)
          for (a[2] = 0; a[2] < 3; a[2]++)
            for (a[3] = 0; a[3] < 3; a[3]++)
              acc += v * PS(S1, mu[0], mu[1], a[c[0]], a[c[1]]) * PS(S2, mu[2], mu[3], a[c[2]], a[c[3]]) *
                     conj(PS(S3, (mu[5] + 2) % 4, (mu[4] + 2) % 4, a[c[5]], a[c[4]])) *
                     conj(PS(S4, (mu[7] + 2) % 4, (mu[6] + 2) % 4, a[c[7]], a[c[6]]));
    }
    store(acc, out + 2 * k);
  }
}

int orc_tetra_field(const int L[4], int t, const orc_prop *p1, const orc_prop *p2, const orc_prop *p3,
                    const orc_prop *p4, const int *idx, const int *col, const double *val, const int *first,
                    const int *count, int nops, double *out, int nthreads) {
  int64_t V3 = (int64_t)L[0] * L[1] * L[2];
  int err = 0;
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(static) reduction(| : err)
#endif
  for (int64_t x = 0; x < V3; x++) {
    double S1[2 * NPROP], S2[2 * NPROP], S3[2 * NPROP], S4[2 * NPROP];
    if (gather_site(p1, L[3], V3, t, x, S1) || gather_site(p2, L[3], V3, t, x, S2) ||
        gather_site(p3, L[3], V3, t, x, S3) || gather_site(p4, L[3], V3, t, x, S4)) {
      err |= 1;
      continue;
    }
    orc_site_tetra(S1, S2, S3, S4, idx, col, val, first, count, nops, out + x * nops * 2);
  }
  return err ? -2 : 0;
}
