// plg_api.hip -- the C ABI of libplegma_amd.so (declared in include/plegma_amd.h).
//
// A plan maps a correlator kind onto
//   contraction launches -> site correlator c[row][ch][x]      (plg_contract.hip)
//   projection GEMM      -> per-split partials part[s][n][m]   (plg_project.hip)
//   finalize             -> FP64 split sums, source phase, H_elem layout
// and is built either per call (plg_correlator, the reference's per-source
// contract* calls) or once for many sources (plg_batch_*, replayed from a
// captured hipGraph).
#include "plg_internal.h"
#include "plg_tetra_tables.h"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <list>
#include <memory>
#include <map>
#include <mutex>
#include <numeric>
#include <sstream>
#include <vector>

namespace plg {

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }
int fail(int code, const std::string &m) {
  g_err = m;
  return code;
}
#define HIPCHK(x)                                                                      \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) return fail(PLG_ERR_HIP, std::string(#x ": ") + hipGetErrorString(e_)); \
  } while (0)

// ------------------------------------------------------------------ context
struct Context {
  bool init = false;
  int L[4] = {0, 0, 0, 0};
  int64_t V3 = 0;
  int device = 0;
  int M = 1;                     // lcm(Lx,Ly,Lz)
  double *d_cos = nullptr, *d_sin = nullptr;
  uint32_t *d_wtab = nullptr;    // [M][6] packed bf16 operand pairs (see plg_init)
  // udsc baryon schedules on the device, built on first use per product
  std::map<int, UdscSched> udsc;
  std::vector<void *> udsc_mem;
  // grow-only workspace
  void *ws_c = nullptr;
  size_t ws_c_bytes = 0;
  void *ws_part = nullptr;
  size_t ws_part_bytes = 0;
};
static Context g_ctx;
// g_mu guards the context (tables, workspace sizes, udsc schedules).
// g_ws_mu is held by a single call (plg_correlator, the udsc and table
// contractions, plg_ft_apply) from the moment it sizes the shared workspace
// until its stream has finished, so two threads' single calls never share
// or free each other's workspace.  g_static_mu (below) is held by
// plg_correlator_static for the whole call.  plg_init and plg_finalize take
// all three, so the context is never rebuilt under a running call; lock
// order g_static_mu, g_ws_mu, g_mu.
static std::mutex g_mu;
static std::mutex g_ws_mu;

// Device scratch of plg_correlator_static: one file-scope array per device,
// held until the process ends.  Host code never passes `scratch` itself (its
// host-side name is no device address): plg_init takes the array's device
// address on every device it sets up (g_scratch_dev), and only that address
// is used.  g_static_mu serialises the calls of one process; calls whose work
// overlaps on one GPU from several processes, or replays of a captured call,
// race on the array.
constexpr int MAX_DEVICES = 64;
static __device__ __align__(256) unsigned char scratch[PLG_STATIC_SCRATCH_BYTES];
static void *g_scratch_dev[MAX_DEVICES];
static std::string g_scratch_err[MAX_DEVICES];
static std::mutex g_static_mu;

static uint16_t host_bf16_rn(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
static float host_bf16_f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  std::memcpy(&f, &u, 4);
  return f;
}
// three bf16 pieces of a double: p0 + p1 + p2 ~ v to about 2^-24 relative
static void split3_host(double v, uint16_t *p) {
  double r = v;
  for (int i = 0; i < 3; i++) {
    p[i] = host_bf16_rn((float)r);
    r -= (double)host_bf16_f(p[i]);
  }
}

static int ensure(void **buf, size_t *have, size_t need) {
  if (need <= *have) return PLG_OK;
  if (*buf) hipFree(*buf);
  *buf = nullptr;
  *have = 0;
  hipError_t e = hipMalloc(buf, need);
  if (e != hipSuccess) {
    *buf = nullptr;
    return fail(PLG_ERR_ALLOC, "device allocation of " + std::to_string(need) + " bytes failed");
  }
  *have = need;
  return PLG_OK;
}

// ------------------------------------------------------------------ conventions
static void momlist_host(int q2max, std::vector<int> &out) {
  out.clear();
  for (int iQ = 0; iQ <= q2max; iQ++)
    for (int nx = iQ; nx >= -iQ; nx--)
      for (int ny = iQ; ny >= -iQ; ny--)
        for (int nz = iQ; nz >= -iQ; nz--)
          if (nx * nx + ny * ny + nz * nz == iQ) {
            out.push_back(nx);
            out.push_back(ny);
            out.push_back(nz);
          }
}

struct Window {
  int startT, endT, localT, maxT;
};
// include/PLEGMA_Correlator.h:53-79 with procPosition = 0, localL = totalL
static Window window_of(const int L[4], const int src[4], int totalT) {
  const int T = L[3];
  Window w;
  int start = (T - src[3]) % T;
  if (start >= totalT) start = 0;
  w.startT = start;
  w.endT = start > 0 ? std::min(totalT, T + start) : 0;
  const int ts = src[3] % T;
  w.localT = w.endT - w.startT + std::min(totalT, T - ts);
  w.maxT = w.endT - w.startT;
  return w;
}
static int row_time(const int L[4], const int src[4], const Window &w, int j) {
  int t = j;
  if (t >= w.maxT) t = (src[3] % L[3]) + t - w.maxT;
  return t;
}

static int popcount_groups(int mask) {
  int n = 0;
  for (int b = 0; b < 10; b++) n += (mask >> b) & 1;
  return n;
}
static bool baryon_mask_ok(int mask) {
  const int allowed = PLG_G_NUCL | PLG_G_DELTAPP_ALL;
  return mask != 0 && (mask & ~allowed) == 0;
}

static int kind_shape(int kind, int mask, int &ng, int &nd, int &nch) {
  switch (kind) {
  case PLG_MESONS: ng = 1; nd = 2; nch = 10; return PLG_OK;
  case PLG_MESONS16: ng = 1; nd = 2; nch = 16; return PLG_OK;
  case PLG_MESONS_NEW: ng = 1; nd = 1; nch = 10; return PLG_OK;
  case PLG_MESONS_ALL: ng = 1; nd = 1; nch = 34; return PLG_OK;
  case PLG_BARYONS:
    if (!baryon_mask_ok(mask)) return fail(PLG_ERR_UNSUPPORTED, "baryon groups: nucl_nucl and deltapp_deltamm_kk only");
    ng = popcount_groups(mask); nd = 2; nch = 16; return PLG_OK;
  case PLG_TWOPT:
    if (mask != 0 && !baryon_mask_ok(mask)) return fail(PLG_ERR_UNSUPPORTED, "twopt groups: nucl_nucl and deltapp_deltamm_kk only");
    ng = 1 + popcount_groups(mask); nd = 1; nch = 16; return PLG_OK;
  case PLG_MESON_MESON: ng = 1; nd = 1; nch = 20; return PLG_OK;
  }
  return fail(PLG_ERR_INVALID, "unknown kind " + std::to_string(kind));
}

// ------------------------------------------------------------------ plans
struct Launch {
  enum { MESON, BARYON, MESMES, UDSC, TETRA } what;
  int set;          // meson set
  bool two, mes;    // baryon
  int gmask;
  int ch0;
};

struct JobIn {
  int src[4];
  int totalT;
  int pa, pb;       // prop indices
  int64_t out_off;
};

struct Plan {
  int kind = 0, mask = 0, mode = PLG_MODE_F64, terms = 6, space = PLG_MOMENTUM_SPACE;
  int T = 1;        // compute precision of the contraction (PLG_F64 / PLG_F32)
  int nch_c = 0;    // computed channels per row
  int nmom = 0;
  std::vector<int> moms;
  std::vector<int> ms;                 // scaled momenta n_k (M / L_k) mod M
  int npairs = 0;                      // (n, -n) pairs of the mixed projection
  int zero_pair = 0;                   // the last pair is n = 0 (summed, not a GEMM row)
  std::vector<int> pair_ms, mom_pair;
  std::vector<RowDev> rows;
  std::vector<OutRowDev> orows;
  std::vector<int> csel;
  std::vector<int64_t> ch_off;
  int64_t m_stride = 0;
  std::vector<Launch> launches;
  std::vector<plg_prop> props;
  int64_t out_elems = 0;
  int64_t nsites = 0;
  // tuning
  int mtile = 64, mbw = 9, nsplit = 1, tpb = 1;
  int gemm_code = 0;  // the GEMM tile picked before the separable form (its fallback)
  // device tables (owned by the plan)
  void *d_props = nullptr;
  RowDev *d_rows = nullptr;
  OutRowDev *d_orows = nullptr;
  int *d_csel = nullptr;
  int64_t *d_choff = nullptr;
  int *d_ms = nullptr, *d_pair_ms = nullptr, *d_mom_pair = nullptr;
  uint16_t *d_pidx = nullptr;  // phase indices of the mixed projection
  // separable projection (both precisions): host tables and their device copy
  SepHost sep;
  void *d_sep = nullptr;
  // udsc baryons: the flavour product's schedule (device arrays owned by the context)
  UdscSched udsc{};
  // four-propagator tables: terms (owned by the caller of the plan) and rows
  TetraSched tetra{}, tetra2{};  // the table's two term sets (tetra2: views P1, P2, P4, P3)
  std::vector<RowDev4> rows4;
  RowDev4 *d_rows4 = nullptr;
  // batch-owned workspace (single calls use the context workspace)
  void *c = nullptr, *part = nullptr;
  size_t c_bytes = 0, part_bytes = 0;
  bool own_ws = false;
  // graph
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  void *graph_out = nullptr;
  hipStream_t cap_stream = nullptr;

  ~Plan() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    if (cap_stream) hipStreamDestroy(cap_stream);
    for (void *p : {(void *)d_props, (void *)d_rows, (void *)d_orows, (void *)d_csel, (void *)d_choff, (void *)d_ms,
                    (void *)d_pair_ms, (void *)d_mom_pair, (void *)d_pidx, (void *)d_rows4, d_sep})
      if (p) hipFree(p);
    if (own_ws) {
      if (c) hipFree(c);
      if (part) hipFree(part);
    }
  }
};

static bool same_view(const plg_prop &a, const plg_prop &b) {
  return a.data == b.data && a.precision == b.precision && a.t_first == b.t_first && a.nt == b.nt &&
         a.comp_stride == b.comp_stride;
}

static int check_prop(const plg_prop &p, int want_prec, int t) {
  if (!p.data) return fail(PLG_ERR_INVALID, "propagator data is NULL");
  if (p.precision != want_prec) return fail(PLG_ERR_INVALID, "propagator precision does not match the mode");
  const int T = g_ctx.L[3];
  if (p.nt <= 0 || p.nt > T || p.t_first < 0 || p.t_first >= T)
    return fail(PLG_ERR_INVALID, "propagator view: bad t_first/nt");
  if (p.comp_stride < (int64_t)p.nt * g_ctx.V3) return fail(PLG_ERR_INVALID, "propagator view: comp_stride < nt*V3");
  int plane = ((t - p.t_first) % T + T) % T;
  if (plane >= p.nt)
    return fail(PLG_ERR_RANGE, "propagator view does not hold time slice " + std::to_string(t));
  return PLG_OK;
}

// n_k (M / L_k) mod M for a momentum list; sign -1 negates the momenta
static std::vector<int> scaled_moms(const int *moms, int nmom, int sign) {
  std::vector<int> ms;
  for (int m = 0; m < nmom; m++)
    for (int i = 0; i < 3; i++) {
      long v = (long)sign * moms[3 * m + i] * (g_ctx.M / g_ctx.L[i]);
      v %= g_ctx.M;
      if (v < 0) v += g_ctx.M;
      ms.push_back((int)v);
    }
  return ms;
}

// build the host side of a plan for a list of jobs
static int build_plan(Plan &P, int kind, int mask, const std::vector<plg_prop> &props,
                      const std::vector<JobIn> &jobs, const int *moms, int nmom, int mode, int terms,
                      int space) {
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (mode != PLG_MODE_F64 && mode != PLG_MODE_MIXED) return fail(PLG_ERR_INVALID, "bad mode");
  if (mode == PLG_MODE_MIXED && terms != 3 && terms != 6) return fail(PLG_ERR_INVALID, "split must be 3 or 6");
  if (space != PLG_MOMENTUM_SPACE && space != PLG_POSITION_SPACE) return fail(PLG_ERR_INVALID, "bad space");
  if (space == PLG_MOMENTUM_SPACE && (nmom <= 0 || !moms)) return fail(PLG_ERR_INVALID, "empty momentum list");
  int ng, nd, nch;
  int rc = kind_shape(kind, mask, ng, nd, nch);
  if (rc) return rc;
  P.kind = kind; P.mask = mask; P.mode = mode; P.terms = terms; P.space = space;
  P.T = mode == PLG_MODE_F64 ? PLG_F64 : PLG_F32;
  P.props = props;
  P.nmom = space == PLG_MOMENTUM_SPACE ? nmom : 0;
  P.moms.assign(moms, moms + 3 * P.nmom);
  P.ms = scaled_moms(moms, P.nmom, 1);
  if (P.nmom > 0) P.npairs = make_pairs(P.ms, g_ctx.M, P.pair_ms, P.mom_pair, &P.zero_pair);
  const int64_t V3 = g_ctx.V3;
  const int64_t inner = space == PLG_MOMENTUM_SPACE ? (int64_t)nmom : V3;  // per-row extent

  // channels computed per row, launches, output-channel maps
  bool pair_same = false;
  const int DEL = PLG_G_DELTAPP_ALL;
  int gcomp = 0;  // baryon groups computed
  std::vector<int> gsel;  // computed group index of each output baryon group
  auto baryon_groups = [&](int m, int &gm) {
    gm = (m == PLG_G_NUCL) ? PLG_G_NUCL : ((m & ~DEL) == 0 ? DEL : (PLG_G_NUCL | DEL));
    std::vector<int> bits;
    for (int b : {0, 7, 8, 9})
      if (gm & (1 << b)) bits.push_back(b);
    gcomp = (int)bits.size();
    gsel.clear();
    for (int b : {0, 7, 8, 9})
      if (m & (1 << b)) gsel.push_back((int)(std::find(bits.begin(), bits.end(), b) - bits.begin()));
  };
  std::vector<int> obits;  // reference group bit of each output baryon group
  for (int b : {0, 7, 8, 9})
    if (mask & (1 << b)) obits.push_back(b);
  // result scale of an out row: the contraction is multilinear in the
  // propagators, ka/kb = how many factors of each view it holds
  auto set_scale = [&](OutRowDev &o, const plg_prop &a, const plg_prop &b, int t, int ka, int kb) {
    const int T = g_ctx.L[3];
    o.ea = ka ? a.plane_exp : nullptr;
    o.eb = kb ? b.plane_exp : nullptr;
    o.pla = ((t - a.t_first) % T + T) % T;
    o.plb = ((t - b.t_first) % T + T) % T;
    o.ka = ka;
    o.kb = kb;
  };
  P.rows.clear();
  P.orows.clear();
  P.launches.clear();
  P.nsites = 0;
  int64_t out_total = 0;

  // decide per-kind structure (same for every job)
  switch (kind) {
  case PLG_MESONS:
  case PLG_MESONS16:
    P.nch_c = nch;
    P.launches.push_back({Launch::MESON, kind == PLG_MESONS ? SET_M10 : SET_M16, false, false, 0, 0});
    break;
  case PLG_MESONS_NEW:
    P.nch_c = 10;
    P.launches.push_back({Launch::MESON, SET_M10_PAIR, false, false, 0, 0});
    break;
  case PLG_MESONS_ALL:
    P.nch_c = 34;
    P.launches.push_back({Launch::MESON, SET_MALL, false, false, 0, 0});
    break;
  case PLG_BARYONS: {
    int gm;
    baryon_groups(mask, gm);
    P.launches.push_back({Launch::BARYON, 0, true, false, gm, 0});  // 'two' fixed per job below
    P.nch_c = 2 * gcomp * 16;
    break;
  }
  case PLG_MESON_MESON:
    P.nch_c = 20;
    P.launches.push_back({Launch::MESMES, 0, true, false, 0, 0});
    break;
  case PLG_TWOPT: {
    int gm = 0;
    if (mask) baryon_groups(mask, gm);
    if (mode == PLG_MODE_MIXED) {
      P.launches.push_back({Launch::BARYON, 0, false, true, gm, 0});
    } else {
      P.launches.push_back({Launch::MESON, SET_M16, false, false, 0, 0});
      if (gm) P.launches.push_back({Launch::BARYON, 0, false, false, gm, 16});
    }
    P.nch_c = 16 + gcomp * 16;
    break;
  }
  }

  // output channel maps (identical for every out row)
  P.csel.clear();
  P.ch_off.clear();
  // per-job H sizes depend on localT; ch_off uses group stride, which depends
  // on localT too -> encode group stride into base per (g) by making ch_off
  // relative and adding per-row group offsets is not possible with one map,
  // so out rows are emitted per (g, d, j) when ng > 1.
  bool two_props_baryons = true;
  for (size_t jb = 0; jb < jobs.size(); jb++) {
    const JobIn &J = jobs[jb];
    const Window w = window_of(g_ctx.L, J.src, J.totalT);
    const int lt = w.localT;
    const plg_prop &pa = props[J.pa];
    const plg_prop &pb = props[J.pb];
    pair_same = same_view(pa, pb);
    const int crow0 = (int)P.rows.size();
    const int64_t gstride = (int64_t)nd * lt * inner * nch;  // H group stride
    for (int j = 0; j < lt; j++) {
      const int t = row_time(g_ctx.L, J.src, w, j);
      if ((rc = check_prop(pa, P.T, t))) return rc;
      if ((rc = check_prop(pb, P.T, t))) return rc;
    }
    switch (kind) {
    case PLG_MESONS:
    case PLG_MESONS16: {
      const int nd_c = pair_same ? 1 : 2;
      for (int d = 0; d < nd_c; d++)
        for (int j = 0; j < lt; j++)
          P.rows.push_back({row_time(g_ctx.L, J.src, w, j), d == 0 ? J.pa : J.pb, d == 0 ? J.pa : J.pb, (int)jb});
      for (int d = 0; d < 2; d++)
        for (int j = 0; j < lt; j++) {
          OutRowDev o{};
          o.crow = crow0 + (pair_same ? j : d * lt + j);
          o.ch_base = 0;
          o.sx = J.src[0]; o.sy = J.src[1]; o.sz = J.src[2];
          o.base = J.out_off + ((int64_t)d * lt + j) * inner * nch;
          const plg_prop &pd = d == 0 ? pa : pb;
          set_scale(o, pd, pd, row_time(g_ctx.L, J.src, w, j), 2, 0);
          P.orows.push_back(o);
        }
      break;
    }
    case PLG_MESONS_NEW:
    case PLG_MESONS_ALL:
      for (int j = 0; j < lt; j++) {
        P.rows.push_back({row_time(g_ctx.L, J.src, w, j), J.pa, J.pb, (int)jb});
        OutRowDev o{};
        o.crow = crow0 + j;
        o.sx = J.src[0]; o.sy = J.src[1]; o.sz = J.src[2];
        o.base = J.out_off + (int64_t)j * inner * nch;
        set_scale(o, pa, pb, row_time(g_ctx.L, J.src, w, j), 1, 1);
        P.orows.push_back(o);
      }
      break;
    case PLG_BARYONS:
      if (jb == 0) two_props_baryons = !pair_same;
      else if (two_props_baryons != !pair_same)
        return fail(PLG_ERR_UNSUPPORTED, "batch mixes jobs with one and two distinct propagators");
      for (int j = 0; j < lt; j++) {
        P.rows.push_back({row_time(g_ctx.L, J.src, w, j), J.pa, J.pb, (int)jb});
        for (size_t g = 0; g < gsel.size(); g++)
          for (int d = 0; d < 2; d++) {
            OutRowDev o{};
            o.crow = crow0 + j;
            const int dc = pair_same ? 0 : d;  // computed dataset
            o.ch_base = (dc * gcomp + gsel[g]) * 16;
            o.sx = J.src[0]; o.sy = J.src[1]; o.sz = J.src[2];
            o.base = J.out_off + (int64_t)g * gstride + ((int64_t)d * lt + j) * inner * nch;
            // proton of dataset d: the d-th propagator twice and the other once;
            // Delta of dataset d: the d-th propagator three times
            const bool nucl = obits[g] == 0;
            const int k1 = nucl ? (d == 0 ? 2 : 1) : (d == 0 ? 3 : 0);
            const int k2 = nucl ? (d == 0 ? 1 : 2) : (d == 0 ? 0 : 3);
            set_scale(o, pa, pb, row_time(g_ctx.L, J.src, w, j), k1, k2);
            P.orows.push_back(o);
          }
      }
      break;
    case PLG_MESON_MESON:
      // two out rows per time slice: the 16 mesons (one factor of each
      // propagator) and the 4 meson-meson channels (two of each)
      for (int j = 0; j < lt; j++) {
        const int t = row_time(g_ctx.L, J.src, w, j);
        P.rows.push_back({t, J.pa, J.pb, (int)jb});
        for (int part = 0; part < 2; part++) {
          OutRowDev o{};
          o.crow = crow0 + j;
          o.ch_base = part == 0 ? 0 : 16;
          o.nch_row = part == 0 ? 16 : 4;
          o.sx = J.src[0]; o.sy = J.src[1]; o.sz = J.src[2];
          o.base = J.out_off + (int64_t)j * inner * nch + o.ch_base;
          set_scale(o, pa, pb, t, part == 0 ? 1 : 2, part == 0 ? 1 : 2);
          P.orows.push_back(o);
        }
      }
      break;
    case PLG_TWOPT:
      for (int j = 0; j < lt; j++) {
        P.rows.push_back({row_time(g_ctx.L, J.src, w, j), J.pa, J.pa, (int)jb});
        for (int g = 0; g < ng; g++) {
          OutRowDev o{};
          o.crow = crow0 + j;
          o.ch_base = g == 0 ? 0 : 16 + gsel[g - 1] * 16;
          o.sx = J.src[0]; o.sy = J.src[1]; o.sz = J.src[2];
          o.base = J.out_off + (int64_t)g * gstride + (int64_t)j * inner * nch;
          set_scale(o, pa, pa, row_time(g_ctx.L, J.src, w, j), g == 0 ? 2 : 3, 0);
          P.orows.push_back(o);
        }
      }
      break;
    }
    P.nsites += (int64_t)lt * V3;
    out_total = std::max(out_total, J.out_off + (int64_t)ng * nd * lt * inner * nch);
  }
  if (kind == PLG_BARYONS) {
    P.launches[0].two = two_props_baryons;
    if (!two_props_baryons) P.nch_c = gcomp * 16;
  }
  // every out row reads nout channels contiguous from ch_base
  const int nout = (kind == PLG_BARYONS || kind == PLG_TWOPT || kind == PLG_MESON_MESON) ? 16 : nch;
  for (int k = 0; k < nout; k++) {
    P.csel.push_back(k);
    P.ch_off.push_back(space == PLG_MOMENTUM_SPACE ? k : k);
  }
  P.m_stride = nch;  // momentum (or site) stride in H
  P.out_elems = out_total;
  return PLG_OK;
}

template <typename T> static void fill_propdev(const std::vector<plg_prop> &in, std::vector<PropDev<T>> &out) {
  out.clear();
  for (auto &p : in)
    out.push_back(PropDev<T>{(const cx<T> *)p.data, p.comp_stride, p.t_first, p.nt, p.rotate_sign, p.bc_t0});
}

// plan tables are small; copied synchronously at plan build time
template <typename A> static int upload(A **dst, const std::vector<A> &v, hipStream_t) {
  if (*dst) hipFree(*dst);
  *dst = nullptr;
  size_t n = std::max<size_t>(v.size(), 1) * sizeof(A);
  HIPCHK(hipMalloc((void **)dst, n));
  if (!v.empty()) HIPCHK(hipMemcpy(*dst, v.data(), v.size() * sizeof(A), hipMemcpyHostToDevice));
  return PLG_OK;
}

static void choose_tiles(Plan &P) {
  const int64_t ncols = (int64_t)P.rows.size() * P.nch_c;
  int64_t tiles;
  if (P.T == PLG_F64) {
    const int R = 2 * std::max(P.npairs, 1);  // real rows {cos, sin} per momentum pair
    P.mtile = R <= 96 ? 32 : 64;
    tiles = ((ncols + 63) / 64) * ((R + P.mtile - 1) / P.mtile);
  } else {
    const int gp = std::max(P.npairs - P.zero_pair, 1);
    P.mbw = pick_proj(gp);
    // PLG_PROJ_CODE pins the projection tile (tile studies and tests)
    if (const char *e = std::getenv("PLG_PROJ_CODE")) P.mbw = std::atoi(e);
    tiles = proj_workgroups(P.mbw, ncols, gp);
  }
  int64_t ns = (1024 + tiles - 1) / tiles;
  ns = std::max<int64_t>(1, std::min<int64_t>(ns, g_ctx.V3 / 512));
  P.nsplit = (int)std::max<int64_t>(1, ns);
  // the separable projection where the momenta suit it (PLG_PROJ_SEP=0
  // keeps the phase-matrix GEMM; a pinned PLG_PROJ_CODE wins for FP32)
  P.sep = SepHost{};
  const char *se = std::getenv("PLG_PROJ_SEP");
  if (P.space == PLG_MOMENTUM_SPACE && P.nmom > 0 && !(se && std::string(se) == "0"))
    sep_build(g_ctx.L, g_ctx.M, P.ms, ncols, P.T == PLG_F64, P.sep);
  P.gemm_code = P.T == PLG_F64 ? P.mtile : P.mbw;
  if (P.sep.ok && !(P.T == PLG_F32 && std::getenv("PLG_PROJ_CODE"))) {
    if (P.T == PLG_F64) P.mtile = PROJ_SEP;
    else P.mbw = PROJ_SEP;
  }
  const int64_t btiles = (g_ctx.V3 + 15) / 16;
  int64_t tpb = (btiles * (int64_t)P.rows.size()) / 8192;
  P.tpb = (int)std::max<int64_t>(1, std::min<int64_t>(tpb, 64));
}

static int upload_plan(Plan &P, hipStream_t st) {
  int rc;
  if (P.T == PLG_F64) {
    std::vector<PropDev<double>> pd;
    fill_propdev(P.props, pd);
    if ((rc = upload((PropDev<double> **)&P.d_props, pd, st))) return rc;
  } else {
    std::vector<PropDev<float>> pd;
    fill_propdev(P.props, pd);
    if ((rc = upload((PropDev<float> **)&P.d_props, pd, st))) return rc;
  }
  if ((rc = upload(&P.d_rows, P.rows, st))) return rc;
  if ((rc = upload(&P.d_orows, P.orows, st))) return rc;
  if ((rc = upload(&P.d_csel, P.csel, st))) return rc;
  if ((rc = upload(&P.d_choff, P.ch_off, st))) return rc;
  if ((rc = upload(&P.d_ms, P.ms, st))) return rc;
  if ((rc = upload(&P.d_pair_ms, P.pair_ms, st))) return rc;
  if ((rc = upload(&P.d_mom_pair, P.mom_pair, st))) return rc;
  if (P.sep.ok) {
    unsigned char *d = nullptr;
    if ((rc = upload(&d, P.sep.blob, st))) return rc;
    if (P.d_sep) hipFree(P.d_sep);
    P.d_sep = d;
  }
  if (P.npairs > 0 && P.T == PLG_F32) {  // phase indices of the mixed projection
    if (P.d_pidx) hipFree(P.d_pidx);
    P.d_pidx = nullptr;
    HIPCHK(hipMalloc((void **)&P.d_pidx, phase_index_bytes(g_ctx.V3, P.npairs)));
    ProjGeom g;
    g.V3 = g_ctx.V3;
    for (int i = 0; i < 3; i++) g.L[i] = g_ctx.L[i];
    g.M = g_ctx.M;
    g.npairs = P.npairs;
    g.pair_ms = P.d_pair_ms;
    if ((rc = launch_phase_index(g, P.d_pidx, st))) return fail(rc, "phase index launch failed");
  }
  return PLG_OK;
}

static size_t c_bytes(const Plan &P) {
  return (size_t)P.rows.size() * P.nch_c * ctile_sites(g_ctx.V3) * (P.T == PLG_F64 ? 16 : 8);
}
static int proj_code(const Plan &P) { return P.T == PLG_F64 ? P.mtile : P.mbw; }
static bool plan_sep(const Plan &P) { return P.sep.ok && proj_code(P) == PROJ_SEP; }
static size_t part_bytes(const Plan &P) {
  if (plan_sep(P)) return sep_part_bytes(P.sep.g);
  return proj_part_bytes(P.nsplit, (int64_t)P.rows.size() * P.nch_c, std::max(P.nmom, 1), P.npairs);
}
// room for every candidate of a tuning run: the GEMM with the most K splits
// and the separable form
static size_t part_bytes_tune(const Plan &P) {
  const int ns = (int)std::max<int64_t>(P.nsplit, std::min<int64_t>(32, g_ctx.V3 / 256));
  size_t b = proj_part_bytes(ns, (int64_t)P.rows.size() * P.nch_c, std::max(P.nmom, 1), P.npairs);
  if (P.sep.ok) b = std::max(b, sep_part_bytes(P.sep.g));
  return b;
}

// projection geometry of a plan
static ProjGeom plan_geom(const Plan &P) {
  ProjGeom g;
  g.V3 = g_ctx.V3;
  for (int i = 0; i < 3; i++) g.L[i] = g_ctx.L[i];
  g.M = g_ctx.M;
  g.nmom = P.nmom;
  g.ncols = (int)P.rows.size() * P.nch_c;
  g.nsplit = P.nsplit;
  g.mom_scaled = P.d_ms;
  g.ctile = 1;
  g.cos_tab = g_ctx.d_cos;
  g.sin_tab = g_ctx.d_sin;
  g.npairs = P.npairs;
  g.pair_ms = P.d_pair_ms;
  g.mom_pair = P.d_mom_pair;
  g.pidx = P.d_pidx;
  g.zero_pair = P.zero_pair;
  g.shifted = P.T == PLG_F32 && P.npairs > 0;
  return g;
}
// the geometry the finalize reads: the separable projection's sums are
// momentum-major, one piece, unshifted
static ProjGeom sep_proj_geom(ProjGeom g) {
  g.sep = 1;
  g.npairs = 0;
  g.nsplit = 1;
  g.shifted = 0;
  return g;
}

// enqueue the whole plan on a stream (capturable: no allocation, no memset)
// optional stage profiling: an event is recorded after every stage
struct StageLog {
  std::vector<hipEvent_t> ev;
  std::vector<std::string> name;
  void mark(const std::string &n, hipStream_t st) {
    hipEvent_t e;
    hipEventCreate(&e);
    hipEventRecord(e, st);
    ev.push_back(e);
    name.push_back(n);
  }
};

static std::string launch_name(const Plan &P, const Launch &l) {
  std::string t = P.T == PLG_F64 ? "f64" : "f32";
  if (l.what == Launch::MESON) return "contract_mesons_" + t;
  if (l.what == Launch::MESMES) return "contract_mesmes_" + t;
  if (l.what == Launch::UDSC) return "contract_udsc_" + t;
  if (l.what == Launch::TETRA) return "contract_tetra_" + t;
  return std::string(l.mes ? "contract_twopt_" : "contract_baryons_") + t;
}

template <typename T> static int launch_contract_t(const Plan &P, const Launch &l, void *c, hipStream_t st) {
  auto *pd = (const PropDev<T> *)P.d_props;
  const int nrows = (int)P.rows.size();
  switch (l.what) {
  case Launch::MESON:
    return launch_meson<T>(l.set, pd, P.d_rows, nrows, (cx<T> *)c, g_ctx.V3, g_ctx.L[3], P.nch_c, l.ch0, st);
  case Launch::MESMES:
    return launch_mesmes<T>(pd, P.d_rows, nrows, (cx<T> *)c, g_ctx.V3, g_ctx.L[3], P.nch_c, l.ch0, P.tpb, st);
  case Launch::UDSC:
    return launch_udsc<T>(P.udsc, pd, P.d_rows, nrows, (cx<T> *)c, g_ctx.V3, g_ctx.L[3], P.nch_c, l.ch0, P.tpb, st);
  case Launch::TETRA:  // set 1 runs on the second half of the rows (views P1, P2, P4, P3)
    return launch_tetra<T>(l.set ? P.tetra2 : P.tetra, pd, P.d_rows4 + (l.set ? nrows : 0), nrows, (cx<T> *)c,
                           g_ctx.V3, g_ctx.L[3], P.nch_c, P.tpb, st);
  default:
    return launch_baryon<T>(l.two, l.mes, l.gmask, pd, P.d_rows, nrows, (cx<T> *)c, g_ctx.V3, g_ctx.L[3],
                            P.nch_c, l.ch0, P.tpb, st);
  }
}
static int launch_contract(const Plan &P, const Launch &l, void *c, hipStream_t st) {
  return P.T == PLG_F64 ? launch_contract_t<double>(P, l, c, st) : launch_contract_t<float>(P, l, c, st);
}

static int enqueue(Plan &P, void *c, void *part, void *out, int out_prec, hipStream_t st,
                   StageLog *log = nullptr) {
  const int nrows = (int)P.rows.size();
  int rc;
  if (log) log->mark("start", st);
  for (const Launch &l : P.launches) {
    rc = launch_contract(P, l, c, st);
    if (rc) return fail(rc, "contraction launch failed: " + std::string(hipGetErrorString(hipGetLastError())));
    if (log) log->mark(launch_name(P, l), st);
  }
  const int nout = (int)P.csel.size();
  if (P.space == PLG_POSITION_SPACE) {
    rc = launch_scatter_pos(c, P.T, g_ctx.V3, (int64_t)P.rows.size() * P.nch_c, P.d_orows, (int)P.orows.size(),
                            P.nch_c, nout, P.d_csel, P.d_choff,
                            P.m_stride, out, out_prec, st);
    return rc ? fail(rc, "position-space scatter failed") : PLG_OK;
  }
  const bool sep = plan_sep(P);
  const ProjGeom g = sep ? sep_proj_geom(plan_geom(P)) : plan_geom(P);
  if (sep) rc = launch_project_sep(sep_geom(P.sep, P.d_sep), c, P.T, (double *)part, st);
  else if (P.T == PLG_F64) rc = launch_project_f64(g, (const cx<double> *)c, (cx<double> *)part, P.mtile, st);
  else rc = launch_project_bf16(g, (const cx<float> *)c, (double *)part, P.terms, g_ctx.d_wtab, P.mbw, st);
  if (rc) return fail(rc, "projection launch failed");
  if (log)
    log->mark(sep ? (P.T == PLG_F64 ? "project_sep_f64" : "project_sep_f32")
                  : (P.T == PLG_F64 ? "project_f64" : (P.terms == 6 ? "project_bf16x6" : "project_bf16x3")),
              st);
  rc = launch_finalize(g, (const double *)part, P.d_orows, (int)P.orows.size(), P.nch_c, nout, P.d_csel,
                       P.d_choff, P.m_stride, out, out_prec, st);
  if (log) log->mark("finalize", st);
  return rc ? fail(rc, "finalize launch failed") : PLG_OK;
}

// ------------------------------------------------------------------ autotuner
// Times the candidate tiles of a plan's projection (tile code / row tile and
// K splits) and its contraction (tiles per workgroup) on the plan's own
// buffers, keeps the fastest, and remembers it in a text cache keyed by
// device architecture, lattice, momenta, columns and precision.
static std::string tune_cache_path() {
  const char *e = getenv("PLG_TUNE_CACHE");
  if (e && *e) return e;
  const char *h = getenv("HOME");
  return std::string(h && *h ? h : "/tmp") + "/.plegma_amd_tunecache";
}

static std::string tune_key(const Plan &P) {
  hipDeviceProp_t prop;
  hipGetDeviceProperties(&prop, g_ctx.device);
  std::ostringstream k;
  k << prop.gcnArchName << " L=" << g_ctx.L[0] << "x" << g_ctx.L[1] << "x" << g_ctx.L[2] << "x" << g_ctx.L[3]
    << " nmom=" << P.nmom << " cols=" << P.rows.size() * P.nch_c << " rows=" << P.rows.size() << " mode=" << P.mode
    << " npairs=" << P.npairs << " terms=" << P.terms << " kind=" << P.kind << " mask=" << P.mask
    << " proj=pair32z.sep" << (P.sep.ok ? 1 : 0);
  for (auto &l : P.launches) k << " L" << l.what << "." << l.set << "." << l.two << l.mes << "." << l.gmask;
  std::string s = k.str();
  for (char &ch : s)
    if (ch == '\n') ch = ' ';
  return s;
}

static bool tune_lookup(const std::string &key, int &a, int &b, int &c) {
  std::ifstream f(tune_cache_path());
  std::string line;
  bool found = false;
  while (std::getline(f, line)) {
    const size_t bar = line.find(" | ");
    if (bar == std::string::npos || line.substr(0, bar) != key) continue;
    std::istringstream v(line.substr(bar + 3));
    if (v >> a >> b >> c) found = true;  // last entry wins
  }
  return found;
}

static void tune_store(const std::string &key, int a, int b, int c, double ms) {
  std::ofstream f(tune_cache_path(), std::ios::app);
  if (f) f << key << " | " << a << " " << b << " " << c << " " << ms << "\n";
}

template <typename F> static double time_ms(F &&launch, hipStream_t st, int reps = 2) {
  hipEvent_t e0, e1;
  hipEventCreate(&e0);
  hipEventCreate(&e1);
  launch();  // warm
  float best = 1e30f;
  for (int r = 0; r < reps; r++) {
    hipEventRecord(e0, st);
    launch();
    hipEventRecord(e1, st);
    hipEventSynchronize(e1);
    float t = 0;
    hipEventElapsedTime(&t, e0, e1);
    best = std::min(best, t);
  }
  hipEventDestroy(e0);
  hipEventDestroy(e1);
  return best;
}

static int autotune(Plan &P, hipStream_t st) {
  const std::string key = tune_key(P);
  int a, b, c;
  if (tune_lookup(key, a, b, c)) {
    if (P.T == PLG_F64) P.mtile = a;
    else P.mbw = a;
    P.nsplit = b;
    P.tpb = c;
    return PLG_OK;
  }
  const int nrows = (int)P.rows.size();
  // contraction: tiles per workgroup
  std::vector<int> tpbs = {1, 2, 4, 8, 16, 32, 64};
  double best_c = 1e30;
  int best_tpb = P.tpb;
  for (int tpb : tpbs) {
    if (tpb > P.tpb * 8 || tpb * 4 < P.tpb) continue;
    P.tpb = tpb;
    const double ms = time_ms(
        [&] {
          for (const Launch &l : P.launches)
            if (l.what != Launch::MESON) launch_contract(P, l, P.c, st);
        },
        st);
    if (ms < best_c) {
      best_c = ms;
      best_tpb = tpb;
    }
  }
  P.tpb = best_tpb;
  // projection: tile code and K splits, timed together with the finalize
  // (which sums the splits) into a scratch output
  ProjGeom g = plan_geom(P);
  void *tune_out = nullptr;
  if (hipMalloc(&tune_out, std::max<size_t>(1, (size_t)P.out_elems) * sizeof(cx<double>)) != hipSuccess)
    return fail(PLG_ERR_HIP, "autotune output allocation failed");
  const int nout = (int)P.csel.size();
  std::vector<int> codes = P.T == PLG_F64 ? std::vector<int>{32, 64} : proj_candidates(std::max(P.npairs - P.zero_pair, 1));
  double best_p = 1e30;
  // the separable form wins only where it was timed
  int best_code = P.gemm_code, best_ns = P.nsplit;
  for (int code : codes)
    for (int ns : {1, 2, 4, 8, 16, 32}) {
      const size_t need = proj_part_bytes(ns, (int64_t)P.rows.size() * P.nch_c, std::max(P.nmom, 1), P.npairs);
      if ((int64_t)ns * 256 > g_ctx.V3 || need > P.part_bytes) continue;
      g.nsplit = ns;
      int lrc = PLG_OK;
      const double ms = time_ms(
          [&] {
            if (P.T == PLG_F64) lrc = launch_project_f64(g, (const cx<double> *)P.c, (cx<double> *)P.part, code, st);
            else
              lrc = launch_project_bf16(g, (const cx<float> *)P.c, (double *)P.part, P.terms, g_ctx.d_wtab, code, st);
            if (lrc == PLG_OK)
              lrc = launch_finalize(g, (const double *)P.part, P.d_orows, (int)P.orows.size(), P.nch_c, nout,
                                    P.d_csel, P.d_choff, P.m_stride, tune_out, PLG_F64, st);
          },
          st);
      if (lrc != PLG_OK) {  // a tile this geometry does not provide (LDS size)
        (void)hipGetLastError();
        continue;
      }
      if (ms < best_p) {
        best_p = ms;
        best_code = code;
        best_ns = ns;
      }
    }
  if (P.sep.ok && sep_part_bytes(P.sep.g) <= P.part_bytes) {
    const ProjGeom gs = sep_proj_geom(plan_geom(P));
    const SepGeom sg = sep_geom(P.sep, P.d_sep);
    int lrc = PLG_OK;
    const double ms = time_ms(
        [&] {
          lrc = launch_project_sep(sg, P.c, P.T, (double *)P.part, st);
          if (lrc == PLG_OK)
            lrc = launch_finalize(gs, (const double *)P.part, P.d_orows, (int)P.orows.size(), P.nch_c, nout,
                                  P.d_csel, P.d_choff, P.m_stride, tune_out, PLG_F64, st);
        },
        st);
    if (lrc != PLG_OK) {
      (void)hipGetLastError();
    } else if (ms < best_p) {
      best_p = ms;
      best_code = PROJ_SEP;
      best_ns = 1;
    }
  }
  if (P.T == PLG_F64) P.mtile = best_code;
  else P.mbw = best_code;
  P.nsplit = best_ns;
  hipStreamSynchronize(st);
  hipFree(tune_out);
  if (hipGetLastError() != hipSuccess) return fail(PLG_ERR_HIP, "autotune launch failed");
  tune_store(key, best_code, best_ns, best_tpb, best_c + best_p);
  return PLG_OK;
}

// ------------------------------------------------------------------ single-call plans
// plg_correlator keeps the plans of its recent calls (device tables, tiles),
// keyed by everything the plan depends on: a driver that contracts source
// after source into the same propagator buffers plans each geometry once.
// Used under g_ws_mu only; the plans' device tables are read by launches the
// call has synchronised before it returns, so evicting one is safe.
static std::list<std::pair<std::string, std::unique_ptr<Plan>>> g_single;
constexpr size_t SINGLE_PLANS = 32;

// plg_ft_apply keeps the device tables of its last momentum list (momenta,
// pairs, separable tables, phase index), keyed by the list, the field's
// shape and precision and the lattice: a caller that projects field after
// field with one list uploads them once.  Used under g_ws_mu only; every call
// synchronises its stream before it returns, so replacing them is safe.
struct FtTables {
  std::string key;
  int *d_ms = nullptr, *d_pair_ms = nullptr, *d_mom_pair = nullptr;
  uint16_t *d_pidx = nullptr;
  unsigned char *d_sep = nullptr;
  SepHost sh;
  int npairs = 0, zero_pair = 0;
  void release() {
    for (void *p : {(void *)d_ms, (void *)d_pair_ms, (void *)d_mom_pair, (void *)d_pidx, (void *)d_sep})
      if (p) hipFree(p);
    *this = FtTables();
  }
};
static FtTables g_ft;

static std::string single_key(int kind, int mask, const std::vector<plg_prop> &props, const int src[4], int totalT,
                              const int *moms, int nmom, int mode, int space) {
  std::ostringstream k;
  k << kind << ' ' << mask << ' ' << mode << ' ' << space << ' ' << totalT;
  for (int i = 0; i < 4; i++) k << ' ' << src[i];
  for (const plg_prop &p : props)
    k << " | " << p.data << ' ' << p.precision << ' ' << p.t_first << ' ' << p.nt << ' ' << p.comp_stride << ' '
      << (const void *)p.plane_exp << ' ' << p.rotate_sign << ' ' << p.bc_t0;
  k << " | " << nmom;
  if (moms && space == PLG_MOMENTUM_SPACE)
    for (int i = 0; i < 3 * nmom; i++) k << ' ' << moms[i];
  // a pinned projection tile (PLG_PROJ_CODE) is part of the plan
  if (const char *e = std::getenv("PLG_PROJ_CODE")) k << " code=" << e;
  if (const char *e = std::getenv("PLG_PROJ_SEP")) k << " sep=" << e;
  return k.str();
}

static Plan *single_lookup(const std::string &key) {
  for (auto it = g_single.begin(); it != g_single.end(); ++it)
    if (it->first == key) {
      g_single.splice(g_single.begin(), g_single, it);  // most recent first
      return g_single.front().second.get();
    }
  return nullptr;
}

static Plan *single_insert(const std::string &key, std::unique_ptr<Plan> P) {
  g_single.emplace_front(key, std::move(P));
  while (g_single.size() > SINGLE_PLANS) g_single.pop_back();
  return g_single.front().second.get();
}

// tiles of a single-call plan: the tuning cache's entry for the plan's
// geometry, else (PLG_TUNE_SINGLE unset or 1) a tuning run on the shared
// workspace whose result goes to the cache, as the first call of a
// geometry tunes in PLEGMA (lib/kernels/PLEGMA_kernel_tuner.cuh:240-292)
static int tune_single(Plan &P, hipStream_t st) {
  if (std::getenv("PLG_PROJ_CODE") || P.space != PLG_MOMENTUM_SPACE) return PLG_OK;
  int a, b, c;
  if (tune_lookup(tune_key(P), a, b, c)) {
    if (P.T == PLG_F64) P.mtile = a;
    else P.mbw = a;
    P.nsplit = b;
    P.tpb = c;
    return PLG_OK;
  }
  const char *e = std::getenv("PLG_TUNE_SINGLE");
  if (e && std::string(e) == "0") return PLG_OK;
  int rc;
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure(&g_ctx.ws_c, &g_ctx.ws_c_bytes, c_bytes(P)))) return rc;
    if ((rc = ensure(&g_ctx.ws_part, &g_ctx.ws_part_bytes, part_bytes_tune(P)))) return rc;
  }
  P.c = g_ctx.ws_c;
  P.part = g_ctx.ws_part;
  P.part_bytes = g_ctx.ws_part_bytes;
  rc = autotune(P, st);
  P.c = P.part = nullptr;
  P.part_bytes = 0;
  return rc;
}

// ------------------------------------------------------------------ no-allocation correlator
// Carves 256-byte aligned pieces out of the device's static scratch.
struct ScratchCarver {
  unsigned char *base;
  size_t used = 0;
  void *take(size_t bytes) {
    void *p = base + used;
    used += (std::max<size_t>(bytes, 1) + 255) / 256 * 256;
    return p;
  }
};

template <typename A>
static int put_async(ScratchCarver &sc, A **dst, const std::vector<A> &v, hipStream_t st) {
  *dst = (A *)sc.take(v.size() * sizeof(A));
  if (!v.empty()) HIPCHK(hipMemcpyAsync(*dst, v.data(), v.size() * sizeof(A), hipMemcpyHostToDevice, st));
  return PLG_OK;
}

// the device tables of a plan that lives in the static scratch: pointers
// into the array, cleared before the plan is destroyed (it owns none of them)
static void release_static(Plan &P) {
  P.d_props = nullptr;
  P.d_rows = nullptr;
  P.d_orows = nullptr;
  P.d_csel = nullptr;
  P.d_choff = nullptr;
  P.d_ms = P.d_pair_ms = P.d_mom_pair = nullptr;
  P.d_pidx = nullptr;
  P.d_sep = nullptr;
}

// device scratch a static call takes: its tables, the site correlator and
// the projection's partial sums (ScratchCarver's 256-byte pieces)
static size_t static_need(const Plan &P) {
  ScratchCarver probe{nullptr};
  probe.take(2 * (P.T == PLG_F64 ? sizeof(PropDev<double>) : sizeof(PropDev<float>)));
  probe.take(P.rows.size() * sizeof(RowDev));
  probe.take(P.orows.size() * sizeof(OutRowDev));
  probe.take(P.csel.size() * sizeof(int));
  probe.take(P.ch_off.size() * sizeof(int64_t));
  probe.take(P.ms.size() * sizeof(int));
  probe.take(P.pair_ms.size() * sizeof(int));
  probe.take(P.mom_pair.size() * sizeof(int));
  if (plan_sep(P)) probe.take(P.sep.blob.size());
  else if (P.npairs > 0 && P.T == PLG_F32) probe.take(phase_index_bytes(g_ctx.V3, P.npairs));
  probe.take(c_bytes(P));
  if (P.space == PLG_MOMENTUM_SPACE) probe.take(part_bytes(P));
  return probe.used;
}

static int correlator_static(Plan &P, ScratchCarver &sc, void *out, int out_precision, hipStream_t st) {
  int rc;
  // sizes first: nothing is copied or launched unless the whole call fits;
  // the separable projection's per-plane sums take more room than the GEMM's
  // partials, so a call that fits only with the GEMM uses the GEMM
  if (plan_sep(P) && static_need(P) > PLG_STATIC_SCRATCH_BYTES) {
    if (P.T == PLG_F64) P.mtile = P.gemm_code;
    else P.mbw = P.gemm_code;
  }
  const size_t need = static_need(P);
  if (need > PLG_STATIC_SCRATCH_BYTES)
    return fail(PLG_ERR_UNSUPPORTED, "plg_correlator_static: the correlator needs " + std::to_string(need) +
                                         " bytes of device scratch, above the static maximum of " +
                                         std::to_string((size_t)PLG_STATIC_SCRATCH_BYTES));
  if (P.T == PLG_F64) {
    std::vector<PropDev<double>> pd;
    fill_propdev(P.props, pd);
    PropDev<double> *d = nullptr;
    if ((rc = put_async(sc, &d, pd, st))) return rc;
    P.d_props = d;
  } else {
    std::vector<PropDev<float>> pd;
    fill_propdev(P.props, pd);
    PropDev<float> *d = nullptr;
    if ((rc = put_async(sc, &d, pd, st))) return rc;
    P.d_props = d;
  }
  if ((rc = put_async(sc, &P.d_rows, P.rows, st)) || (rc = put_async(sc, &P.d_orows, P.orows, st)) ||
      (rc = put_async(sc, &P.d_csel, P.csel, st)) || (rc = put_async(sc, &P.d_choff, P.ch_off, st)) ||
      (rc = put_async(sc, &P.d_ms, P.ms, st)) || (rc = put_async(sc, &P.d_pair_ms, P.pair_ms, st)) ||
      (rc = put_async(sc, &P.d_mom_pair, P.mom_pair, st)))
    return rc;
  if (plan_sep(P)) {
    unsigned char *d = nullptr;
    if ((rc = put_async(sc, &d, P.sep.blob, st))) return rc;
    P.d_sep = d;
  } else if (P.npairs > 0 && P.T == PLG_F32) {
    P.d_pidx = (uint16_t *)sc.take(phase_index_bytes(g_ctx.V3, P.npairs));
    ProjGeom g = plan_geom(P);
    if ((rc = launch_phase_index(g, P.d_pidx, st))) return fail(rc, "phase index launch failed");
  }
  void *c = sc.take(c_bytes(P));
  void *part = P.space == PLG_MOMENTUM_SPACE ? sc.take(part_bytes(P)) : nullptr;
  return enqueue(P, c, part, out, out_precision, st);
}

}  // namespace plg

using namespace plg;

struct plg_batch_s {
  Plan plan;
};

extern "C" {

int plg_version(void) { return 100; }
const char *plg_last_error(void) { return g_err.c_str(); }

static void finalize_locked();

int plg_init(const int L[4], int device) {
  std::lock_guard<std::mutex> sl(g_static_mu);
  std::lock_guard<std::mutex> wl(g_ws_mu);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!L) return fail(PLG_ERR_INVALID, "L is NULL");
  for (int i = 0; i < 4; i++)
    if (L[i] <= 0) return fail(PLG_ERR_INVALID, "lattice extents must be positive");
  if ((int64_t)L[0] * L[1] * L[2] * L[3] > (int64_t)1 << 32)
    return fail(PLG_ERR_INVALID, "lattice volume above 2^32 sites");
  int ndev = 0;
  HIPCHK(hipGetDeviceCount(&ndev));
  if (device < 0 || device >= ndev) return fail(PLG_ERR_INVALID, "no such device");
  HIPCHK(hipSetDevice(device));
  if (g_ctx.init) finalize_locked();
  // device address of the static scratch on this device (once per device)
  if (device < MAX_DEVICES && !g_scratch_dev[device]) {
    void *ptr = nullptr;
    const hipError_t se = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (se == hipSuccess && ptr) {
      g_scratch_dev[device] = ptr;
    } else {
      g_scratch_err[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(se);
      (void)hipGetLastError();
    }
  }
  std::memcpy(g_ctx.L, L, sizeof(int) * 4);
  g_ctx.V3 = (int64_t)L[0] * L[1] * L[2];
  g_ctx.device = device;
  int M = std::lcm(std::lcm(L[0], L[1]), L[2]);
  g_ctx.M = M;
  // exact phase table: cos/sin(2 pi k / M) for integer k, in FP64 on the host
  // and its bf16 pieces packed as (re, im) operand pairs of the two real rows
  // of a momentum: W[k][p] = cos_p | sin_p << 16, W[k][3+p] = -sin_p | cos_p << 16
  std::vector<double> c(M), s(M);
  std::vector<uint32_t> wt((size_t)M * 6);
  for (int k = 0; k < M; k++) {
    const double th = 2.0 * M_PI * (double)k / (double)M;
    c[k] = std::cos(th);
    s[k] = std::sin(th);
    uint16_t cp[3], sp[3];
    split3_host(c[k], cp);
    split3_host(s[k], sp);
    for (int p = 0; p < 3; p++) {
      wt[(size_t)6 * k + p] = (uint32_t)cp[p] | ((uint32_t)sp[p] << 16);
      wt[(size_t)6 * k + 3 + p] = (uint32_t)(sp[p] ^ 0x8000u) | ((uint32_t)cp[p] << 16);
    }
  }
  HIPCHK(hipMalloc((void **)&g_ctx.d_cos, sizeof(double) * M));
  HIPCHK(hipMalloc((void **)&g_ctx.d_sin, sizeof(double) * M));
  HIPCHK(hipMalloc((void **)&g_ctx.d_wtab, sizeof(uint32_t) * 6 * M));
  HIPCHK(hipMemcpy(g_ctx.d_cos, c.data(), sizeof(double) * M, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g_ctx.d_sin, s.data(), sizeof(double) * M, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(g_ctx.d_wtab, wt.data(), sizeof(uint32_t) * 6 * M, hipMemcpyHostToDevice));
  g_ctx.init = true;
  return PLG_OK;
}

int plg_finalize(void) {
  std::lock_guard<std::mutex> sl(g_static_mu);
  std::lock_guard<std::mutex> wl(g_ws_mu);
  std::lock_guard<std::mutex> lk(g_mu);
  finalize_locked();
  return PLG_OK;
}

// with g_ws_mu and g_mu held
static void finalize_locked() {
  if (!g_ctx.init) return;
  hipDeviceSynchronize();
  g_single.clear();
  g_ft.release();
  for (void *p : {(void *)g_ctx.d_cos, (void *)g_ctx.d_sin, (void *)g_ctx.d_wtab, g_ctx.ws_c, g_ctx.ws_part})
    if (p) hipFree(p);
  for (void *p : g_ctx.udsc_mem) hipFree(p);
  g_ctx = Context();
}

int plg_momlist(int q2max, int *nmom, int *xyz, int cap) {
  if (q2max < 0) return fail(PLG_ERR_INVALID, "The maximum number of Q2 cannot be negative");
  std::vector<int> v;
  momlist_host(q2max, v);
  const int n = (int)v.size() / 3;
  if (nmom) *nmom = n;
  if (xyz)
    for (int i = 0; i < std::min(n, cap) * 3; i++) xyz[i] = v[i];
  return PLG_OK;
}

int plg_window(const int source[4], int totalT, int *startT, int *endT, int *localT) {
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (!source || totalT <= 0 || totalT > g_ctx.L[3]) return fail(PLG_ERR_INVALID, "bad source or totalT");
  for (int i = 0; i < 4; i++)
    if (source[i] < 0 || source[i] >= g_ctx.L[i]) return fail(PLG_ERR_INVALID, "source outside the lattice");
  Window w = window_of(g_ctx.L, source, totalT);
  if (startT) *startT = w.startT;
  if (endT) *endT = w.endT;
  if (localT) *localT = w.localT;
  return PLG_OK;
}

int plg_row_time(const int source[4], int totalT, int j, int *t) {
  int lt;
  int rc = plg_window(source, totalT, nullptr, nullptr, &lt);
  if (rc) return rc;
  if (j < 0 || j >= lt) return fail(PLG_ERR_INVALID, "row outside the window");
  *t = row_time(g_ctx.L, source, window_of(g_ctx.L, source, totalT), j);
  return PLG_OK;
}

int plg_kind_shape(int kind, int group_mask, int *ngroups, int *ndatasets, int *nch) {
  int ng, nd, nc;
  int rc = kind_shape(kind, group_mask, ng, nd, nc);
  if (rc) return rc;
  if (ngroups) *ngroups = ng;
  if (ndatasets) *ndatasets = nd;
  if (nch) *nch = nc;
  return PLG_OK;
}

int plg_correlator_size(int kind, int group_mask, const int source[4], int totalT, int nmom, int space,
                        int64_t *elems) {
  int ng, nd, nc, lt;
  int rc = kind_shape(kind, group_mask, ng, nd, nc);
  if (rc) return rc;
  if ((rc = plg_window(source, totalT, nullptr, nullptr, &lt))) return rc;
  const int64_t inner = space == PLG_MOMENTUM_SPACE ? nmom : g_ctx.V3;
  *elems = (int64_t)ng * nd * lt * inner * nc;
  return PLG_OK;
}

int plg_correlator(int kind, int group_mask, const plg_prop *p1, const plg_prop *p2, const int source[4],
                   int totalT, const int *moms, int nmom, int mode, int space, int out_precision, void *out,
                   void *stream) {
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (!p1 || !out || !source) return fail(PLG_ERR_INVALID, "NULL argument");
  if (out_precision != PLG_F64 && out_precision != PLG_F32) return fail(PLG_ERR_INVALID, "bad out precision");
  int lt;
  int rc = plg_window(source, totalT, nullptr, nullptr, &lt);
  if (rc) return rc;
  hipStream_t st = (hipStream_t)stream;
  std::vector<plg_prop> props = {*p1, p2 ? *p2 : *p1};
  if (!same_view(props[0], props[1]) && (kind == PLG_TWOPT))
    return fail(PLG_ERR_INVALID, "twopt takes one propagator");
  std::lock_guard<std::mutex> wl(g_ws_mu);
  const std::string key = single_key(kind, group_mask, props, source, totalT, moms, nmom, mode, space);
  Plan *P = single_lookup(key);
  if (!P) {
    auto np = std::make_unique<Plan>();
    std::vector<JobIn> jobs = {{{source[0], source[1], source[2], source[3]}, totalT, 0, 1, 0}};
    if ((rc = build_plan(*np, kind, group_mask, props, jobs, moms, nmom, mode, PLG_SPLIT_BF16X6, space))) return rc;
    choose_tiles(*np);
    if ((rc = upload_plan(*np, st))) return rc;
    if ((rc = tune_single(*np, st))) return rc;
    P = single_insert(key, std::move(np));
  }
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure(&g_ctx.ws_c, &g_ctx.ws_c_bytes, c_bytes(*P)))) return rc;
    if ((rc = ensure(&g_ctx.ws_part, &g_ctx.ws_part_bytes, part_bytes(*P)))) return rc;
  }
  rc = enqueue(*P, g_ctx.ws_c, g_ctx.ws_part, out, out_precision, st);
  // the workspace is shared by single calls: finish this correlator before
  // the next call may use it, also after an error
  const hipError_t se = hipStreamSynchronize(st);
  if (rc) return rc;
  if (se != hipSuccess) return fail(PLG_ERR_HIP, std::string("correlator sync: ") + hipGetErrorString(se));
  return PLG_OK;
}

// ------------------------------------------------------------------ no-allocation correlator
int plg_correlator_static(int kind, int group_mask, const plg_prop *p1, const plg_prop *p2, const int source[4],
                          int totalT, const int *moms, int nmom, int mode, int space, int out_precision, void *out,
                          void *stream) {
  std::lock_guard<std::mutex> sl(g_static_mu);  // the whole call (plg_init waits for it)
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (!p1 || !out || !source) return fail(PLG_ERR_INVALID, "NULL argument");
  if (out_precision != PLG_F64 && out_precision != PLG_F32) return fail(PLG_ERR_INVALID, "bad out precision");
  int lt;
  int rc = plg_window(source, totalT, nullptr, nullptr, &lt);
  if (rc) return rc;
  const int dev = g_ctx.device;
  if (dev >= MAX_DEVICES) return fail(PLG_ERR_UNSUPPORTED, "device index above the static scratch table");
  if (!g_scratch_dev[dev])
    return fail(PLG_ERR_HIP, g_scratch_err[dev].empty() ? "static scratch unavailable" : g_scratch_err[dev]);
  hipStream_t st = (hipStream_t)stream;
  std::vector<plg_prop> props = {*p1, p2 ? *p2 : *p1};
  if (!same_view(props[0], props[1]) && (kind == PLG_TWOPT))
    return fail(PLG_ERR_INVALID, "twopt takes one propagator");
  Plan P;
  std::vector<JobIn> jobs = {{{source[0], source[1], source[2], source[3]}, totalT, 0, 1, 0}};
  if ((rc = build_plan(P, kind, group_mask, props, jobs, moms, nmom, mode, PLG_SPLIT_BF16X6, space))) return rc;
  choose_tiles(P);
  // tiles of an earlier tuning run of this geometry, if any (no tuning here:
  // it needs buffers of its own)
  if (!std::getenv("PLG_PROJ_CODE") && space == PLG_MOMENTUM_SPACE) {
    int a, b, c;
    if (tune_lookup(tune_key(P), a, b, c)) {
      if (P.T == PLG_F64) P.mtile = a;
      else P.mbw = a;
      P.nsplit = b;
      P.tpb = c;
    }
  }
  ScratchCarver sc{(unsigned char *)g_scratch_dev[dev]};
  rc = correlator_static(P, sc, out, out_precision, st);
  // the host tables of the copies and the scratch are this call's until its
  // work has finished, also after an error
  const hipError_t se = hipStreamSynchronize(st);
  release_static(P);
  if (rc) return rc;
  if (se != hipSuccess) return fail(PLG_ERR_HIP, std::string("correlator sync: ") + hipGetErrorString(se));
  return PLG_OK;
}

// ------------------------------------------------------------------ udsc baryons
// lib/PLEGMA_Correlator.cu:136-188 (product selection, datasets) and
// lib/kernels/PLEGMA_baryons_udsc.cu:267-300 (one product after the other)
static const char UDSC_LETTERS[4] = {'u', 'd', 's', 'c'};

static void udsc_select(const plg_prop *const fl[4], int only_st, int only_ch, std::vector<int> &prods, int &nd) {
  prods.clear();
  nd = 0;
  for (int i = 0; i < udsc_products(); i++) {
    const std::string f = udsc_flavours(i);
    bool ok = true;
    for (int k = 0; k < 4; k++)
      if (!fl[k] && f.find(UDSC_LETTERS[k]) != std::string::npos) ok = false;
    if (only_st && f.find('s') == std::string::npos) ok = false;
    if (only_ch && f.find('c') == std::string::npos) ok = false;
    if (ok) {
      prods.push_back(i);
      nd += udsc_nops(i);
    }
  }
}

int plg_udsc_count(int *nprod) {
  if (!nprod) return fail(PLG_ERR_INVALID, "NULL argument");
  *nprod = udsc_products();
  return PLG_OK;
}

int plg_udsc_product(int prod, char *flavours, int *nops) {
  if (prod < 0 || prod >= udsc_products()) return fail(PLG_ERR_INVALID, "no such flavour product");
  if (flavours) std::memcpy(flavours, udsc_flavours(prod), 4);
  if (nops) *nops = udsc_nops(prod);
  return PLG_OK;
}

int plg_udsc_op_name(int prod, int k, char *buf, int cap) {
  if (prod < 0 || prod >= udsc_products() || k < 0 || k >= udsc_nops(prod) || !buf || cap <= 0)
    return fail(PLG_ERR_INVALID, "bad product, operator or buffer");
  const char *n = udsc_op_name(prod, k);
  if ((int)std::strlen(n) >= cap) return fail(PLG_ERR_RANGE, "name buffer too small");
  std::strcpy(buf, n);
  return PLG_OK;
}

int plg_baryons_udsc_select(const plg_prop *up, const plg_prop *dn, const plg_prop *st, const plg_prop *ch,
                            int only_st, int only_ch, int *prods, int cap, int *nprods, int *ndatasets) {
  const plg_prop *fl[4] = {up, dn, st, ch};
  std::vector<int> v;
  int nd;
  udsc_select(fl, only_st, only_ch, v, nd);
  if (prods && (int)v.size() > cap) return fail(PLG_ERR_RANGE, "product buffer too small");
  if (prods) std::copy(v.begin(), v.end(), prods);
  if (nprods) *nprods = (int)v.size();
  if (ndatasets) *ndatasets = nd;
  return PLG_OK;
}

int plg_baryons_udsc_size(const plg_prop *up, const plg_prop *dn, const plg_prop *st, const plg_prop *ch,
                          int only_st, int only_ch, const int source[4], int totalT, int nmom, int space,
                          int64_t *elems) {
  int lt, nd, rc;
  if (!elems) return fail(PLG_ERR_INVALID, "NULL argument");
  if ((rc = plg_window(source, totalT, nullptr, nullptr, &lt))) return rc;
  const plg_prop *fl[4] = {up, dn, st, ch};
  std::vector<int> v;
  udsc_select(fl, only_st, only_ch, v, nd);
  *elems = (int64_t)nd * lt * (space == PLG_MOMENTUM_SPACE ? nmom : g_ctx.V3);
  return PLG_OK;
}

// the device schedule of a flavour product, uploaded once per context
static int udsc_device_schedule(int prod, UdscSched &out) {
  std::lock_guard<std::mutex> lk(g_mu);
  auto it = g_ctx.udsc.find(prod);
  if (it != g_ctx.udsc.end()) {
    out = it->second;
    return PLG_OK;
  }
  UdscSchedHost h;
  int rc = udsc_schedule(prod, h);
  if (rc) return fail(rc, "udsc schedule");
  short *q = nullptr;
  signed char *t = nullptr;
  uint16_t *e = nullptr;
  float *c = nullptr;
  if ((rc = upload(&q, h.quads, nullptr))) return rc;
  g_ctx.udsc_mem.push_back(q);
  if ((rc = upload(&t, h.tp, nullptr))) return rc;
  g_ctx.udsc_mem.push_back(t);
  if ((rc = upload(&e, h.ents, nullptr))) return rc;
  g_ctx.udsc_mem.push_back(e);
  if ((rc = upload(&c, h.coef, nullptr))) return rc;
  g_ctx.udsc_mem.push_back(c);
  h.s.quads = q;
  h.s.tp = t;
  h.s.ents = e;
  h.s.coef = c;
  g_ctx.udsc[prod] = h.s;
  out = h.s;
  return PLG_OK;
}

// stage_ms (optional): summed per-stage milliseconds over the products
static int udsc_check_args(const void *out, const int *source, int out_precision, int mode, int space,
                           const int *moms, int nmom) {
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (!out || !source) return fail(PLG_ERR_INVALID, "NULL argument");
  if (out_precision != PLG_F64 && out_precision != PLG_F32) return fail(PLG_ERR_INVALID, "bad out precision");
  if (mode != PLG_MODE_F64 && mode != PLG_MODE_MIXED) return fail(PLG_ERR_INVALID, "bad mode");
  if (space != PLG_MOMENTUM_SPACE && space != PLG_POSITION_SPACE) return fail(PLG_ERR_INVALID, "bad space");
  if (space == PLG_MOMENTUM_SPACE && (nmom <= 0 || !moms)) return fail(PLG_ERR_INVALID, "empty momentum list");
  return PLG_OK;
}

// one three-propagator table contraction (views S1, S2, S3 of the table),
// its operators written as datasets d0 .. d0 + nops - 1 of the output
static int udsc_one(const plg_prop *const views[3], const UdscSched &S, int nops, int d0, int nd_total,
                    const int source[4], int totalT, const int *moms, int nmom, int mode, int space, int out_precision,
                    void *out, hipStream_t hs, std::map<std::string, double> *stage_ms) {
  int lt, rc;
  if ((rc = plg_window(source, totalT, nullptr, nullptr, &lt))) return rc;
  const Window w = window_of(g_ctx.L, source, totalT);
  const int T = g_ctx.L[3];
  const int64_t inner = space == PLG_MOMENTUM_SPACE ? (int64_t)nmom : g_ctx.V3;
  Plan P;
  P.kind = -1;
  P.mode = mode;
  P.terms = PLG_SPLIT_BF16X6;
  P.space = space;
  P.T = mode == PLG_MODE_F64 ? PLG_F64 : PLG_F32;
  for (int i = 0; i < 3; i++) {
    if (!views[i]) return fail(PLG_ERR_INVALID, "NULL propagator view");
    P.props.push_back(*views[i]);
  }
  P.nmom = space == PLG_MOMENTUM_SPACE ? nmom : 0;
  if (P.nmom) P.moms.assign(moms, moms + 3 * P.nmom);
  P.ms = scaled_moms(P.moms.data(), P.nmom, 1);
  if (P.nmom > 0) P.npairs = make_pairs(P.ms, g_ctx.M, P.pair_ms, P.mom_pair, &P.zero_pair);
  P.nch_c = nops;
  for (int j = 0; j < lt; j++) {
    const int t = row_time(g_ctx.L, source, w, j);
    for (const plg_prop &v : P.props)
      if ((rc = check_prop(v, P.T, t))) return rc;
    P.rows.push_back({t, 0, 1, 0, 2});
  }
  for (int k = 0; k < nops; k++)
    for (int j = 0; j < lt; j++) {
      const int t = P.rows[j].t;
      OutRowDev o{};
      o.crow = j;
      o.ch_base = k;
      o.nch_row = 1;
      o.sx = source[0]; o.sy = source[1]; o.sz = source[2];
      o.base = ((int64_t)(d0 + k) * lt + j) * inner;
      // trilinear in the three views: one plane exponent of each
      auto plane = [&](const plg_prop &v) { return ((t - v.t_first) % T + T) % T; };
      o.ea = P.props[0].plane_exp; o.pla = plane(P.props[0]); o.ka = 1;
      o.eb = P.props[1].plane_exp; o.plb = plane(P.props[1]); o.kb = 1;
      o.ec = P.props[2].plane_exp; o.plc = plane(P.props[2]); o.kc = 1;
      P.orows.push_back(o);
    }
  P.csel = {0};
  P.ch_off = {0};
  P.m_stride = 1;
  P.out_elems = (int64_t)nd_total * lt * inner;
  P.nsites = (int64_t)lt * g_ctx.V3;
  P.launches.push_back({Launch::UDSC, 0, false, false, 0, 0});
  P.udsc = S;
  choose_tiles(P);
  P.tpb = 64;  // sites per workgroup of the udsc kernel
  if ((rc = upload_plan(P, hs))) return rc;
  std::lock_guard<std::mutex> wl(g_ws_mu);  // until the stream is synchronised below
  {
    std::lock_guard<std::mutex> lk(g_mu);
    if ((rc = ensure(&g_ctx.ws_c, &g_ctx.ws_c_bytes, c_bytes(P)))) return rc;
    if ((rc = ensure(&g_ctx.ws_part, &g_ctx.ws_part_bytes, part_bytes(P)))) return rc;
  }
  StageLog log;
  rc = enqueue(P, g_ctx.ws_c, g_ctx.ws_part, out, out_precision, hs, stage_ms ? &log : nullptr);
  // the plan's tables are freed on return: wait for every launch, also after an error
  const hipError_t se = hipStreamSynchronize(hs);
  if (!rc && se != hipSuccess) rc = fail(PLG_ERR_HIP, "udsc sync failed");
  for (size_t i = 0; i + 1 < log.ev.size(); i++) {
    float t = 0;
    if (!rc) hipEventElapsedTime(&t, log.ev[i], log.ev[i + 1]);
    (*stage_ms)[log.name[i + 1]] += t;
  }
  for (hipEvent_t e : log.ev) hipEventDestroy(e);
  return rc;
}

static int udsc_run(const plg_prop *up, const plg_prop *dn, const plg_prop *st, const plg_prop *ch, int only_st,
                    int only_ch, const int source[4], int totalT, const int *moms, int nmom, int mode, int space,
                    int out_precision, void *out, void *stream, std::map<std::string, double> *stage_ms) {
  int rc;
  if ((rc = udsc_check_args(out, source, out_precision, mode, space, moms, nmom))) return rc;
  const plg_prop *fl[4] = {up, dn, st, ch};
  std::vector<int> prods;
  int nd_total;
  udsc_select(fl, only_st, only_ch, prods, nd_total);
  int d0 = 0;
  for (int prod : prods) {
    const char *f = udsc_flavours(prod);
    const plg_prop *views[3];
    for (int i = 0; i < 3; i++)
      views[i] = fl[std::find(UDSC_LETTERS, UDSC_LETTERS + 4, f[i]) - UDSC_LETTERS];
    UdscSched S;
    if ((rc = udsc_device_schedule(prod, S))) return rc;
    const int nops = udsc_nops(prod);
    if ((rc = udsc_one(views, S, nops, d0, nd_total, source, totalT, moms, nmom, mode, space, out_precision, out,
                       (hipStream_t)stream, stage_ms)))
      return rc;
    d0 += nops;
  }
  return PLG_OK;
}

// Packed tables by content (the term order is a search over the table, a few
// milliseconds per thousand terms): the last few tables packed, keyed by a
// 64-bit hash of the counts, spins, colour slots and coefficients.
static std::mutex g_pack_mu;
static std::vector<std::pair<uint64_t, std::shared_ptr<const TetraPacked>>> g_pack_cache;
constexpr size_t PACK_CACHE = 8;

static uint64_t hash_bytes(const void *data, size_t n, uint64_t h) {
  const unsigned char *b = (const unsigned char *)data;
  size_t i = 0;
  for (; i + 8 <= n; i += 8) {
    uint64_t w;
    std::memcpy(&w, b + i, 8);
    h = (h ^ w) * 0x100000001b3ull;
    h ^= h >> 29;
  }
  for (; i < n; i++) h = (h ^ b[i]) * 0x100000001b3ull;
  return h;
}

static int tetra_pack_cached(const int16_t *idxs, const int16_t *cols, const float *vals, const int16_t *counts,
                             int nops, TetraPacked &out) {
  if (nops <= 0 || nops > TETRA_MAXOPS) return PLG_ERR_UNSUPPORTED;
  size_t nt = 0;
  for (int k = 0; k < nops; k++) {
    if (counts[k] < 0) return PLG_ERR_INVALID;
    nt += (size_t)counts[k];
  }
  uint64_t h = 0xcbf29ce484222325ull ^ (uint64_t)nops;
  h = hash_bytes(counts, nops * sizeof(int16_t), h);
  h = hash_bytes(idxs, nt * 8 * sizeof(int16_t), h);
  h = hash_bytes(cols, nt * 8 * sizeof(int16_t), h);
  h = hash_bytes(vals, nt * 2 * sizeof(float), h);
  {
    std::lock_guard<std::mutex> lk(g_pack_mu);
    for (auto &e : g_pack_cache)
      if (e.first == h) {
        out = *e.second;
        return PLG_OK;
      }
  }
  auto pk = std::make_shared<TetraPacked>();
  const int rc = tetra_pack(idxs, cols, vals, counts, nops, *pk);
  if (rc) return rc;
  out = *pk;
  std::lock_guard<std::mutex> lk(g_pack_mu);
  if (g_pack_cache.size() >= PACK_CACHE) g_pack_cache.erase(g_pack_cache.begin());
  g_pack_cache.emplace_back(h, std::move(pk));
  return PLG_OK;
}

// one four-propagator table (views P1..P4), its operators written as datasets
// d0 .. d0 + nops - 1 of the output
static int tetra_run(const plg_prop *const views[4], const int16_t *idxs, const int16_t *col_contr, const float *vals,
                     const int16_t *counts, int nops, int d0, const int source[4], int totalT, const int *moms,
                     int nmom, int mode, int space, int out_precision, void *out, void *stream) {
  int rc;
  if ((rc = udsc_check_args(out, source, out_precision, mode, space, moms, nmom))) return rc;
  if (!idxs || !col_contr || !vals || !counts || nops <= 0) return fail(PLG_ERR_INVALID, "empty table");
  for (int i = 0; i < 4; i++)
    if (!views[i]) return fail(PLG_ERR_INVALID, "NULL propagator view");
  TetraPacked pk;
  rc = tetra_pack_cached(idxs, col_contr, vals, counts, nops, pk);
  if (rc == PLG_ERR_INVALID) return fail(rc, "spin or colour index outside 0..3");
  if (rc)
    return fail(rc, "table outside the supported form (64 operators; colour slots a permutation of the four "
                    "sums, closing as <P1,P3><P2,P4>, Tr[(P3^dag P1)(P4^dag P2)], <P1,P4><P2,P3> or "
                    "Tr[(P4^dag P1)(P3^dag P2)]; 65536 distinct coefficients)");
  const bool two_sets = !pk.terms[1].empty();
  int lt;
  if ((rc = plg_window(source, totalT, nullptr, nullptr, &lt))) return rc;
  const Window w = window_of(g_ctx.L, source, totalT);
  const int T = g_ctx.L[3];
  const int64_t inner = space == PLG_MOMENTUM_SPACE ? (int64_t)nmom : g_ctx.V3;
  hipStream_t hs = (hipStream_t)stream;
  Plan P;
  P.kind = -1;
  P.mode = mode;
  P.terms = PLG_SPLIT_BF16X6;
  P.space = space;
  P.T = mode == PLG_MODE_F64 ? PLG_F64 : PLG_F32;
  for (int i = 0; i < 4; i++) P.props.push_back(*views[i]);
  P.nmom = space == PLG_MOMENTUM_SPACE ? nmom : 0;
  if (P.nmom) P.moms.assign(moms, moms + 3 * P.nmom);
  P.ms = scaled_moms(P.moms.data(), P.nmom, 1);
  if (P.nmom > 0) P.npairs = make_pairs(P.ms, g_ctx.M, P.pair_ms, P.mom_pair, &P.zero_pair);
  P.nch_c = nops;
  for (int j = 0; j < lt; j++) {
    const int t = row_time(g_ctx.L, source, w, j);
    for (const plg_prop &v : P.props)
      if ((rc = check_prop(v, P.T, t))) return rc;
    P.rows.push_back({t, 0, 1, 0, 2});
    P.rows4.push_back({t, {0, 1, 2, 3}});
  }
  for (int j = 0; j < lt; j++) P.rows4.push_back({P.rows[j].t, {0, 1, 3, 2}});  // set 1's views
  for (int k = 0; k < nops; k++)
    for (int j = 0; j < lt; j++) {
      const int t = P.rows[j].t;
      OutRowDev o{};
      o.crow = j;
      o.ch_base = k;
      o.nch_row = 1;
      o.sx = source[0]; o.sy = source[1]; o.sz = source[2];
      o.base = ((int64_t)(d0 + k) * lt + j) * inner;
      // multilinear in the four views: one plane exponent of each
      auto plane = [&](const plg_prop &v) { return ((t - v.t_first) % T + T) % T; };
      o.ea = P.props[0].plane_exp; o.pla = plane(P.props[0]); o.ka = 1;
      o.eb = P.props[1].plane_exp; o.plb = plane(P.props[1]); o.kb = 1;
      o.ec = P.props[2].plane_exp; o.plc = plane(P.props[2]); o.kc = 1;
      o.ed = P.props[3].plane_exp; o.pld = plane(P.props[3]); o.kd = 1;
      P.orows.push_back(o);
    }
  P.csel = {0};
  P.ch_off = {0};
  P.m_stride = 1;
  P.out_elems = (int64_t)(d0 + nops) * lt * inner;
  P.nsites = (int64_t)lt * g_ctx.V3;
  // set 0 writes every operator (zero where it has no terms), set 1 adds
  P.launches.push_back({Launch::TETRA, 0, false, false, 0, 0});
  if (two_sets) P.launches.push_back({Launch::TETRA, 1, false, false, 0, 0});
  TetraTerm *d_terms[2] = {nullptr, nullptr};
  float *d_coef = nullptr;
  auto free_tables = [&] {
    for (TetraTerm *p : d_terms) hipFree(p);
    hipFree(d_coef);
  };
  for (int h = 0; h < (two_sets ? 2 : 1) && !rc; h++) rc = upload(&d_terms[h], pk.terms[h], hs);
  if (!rc) rc = upload(&d_coef, pk.coef, hs);
  if (rc) {
    hipStreamSynchronize(hs);
    free_tables();
    return rc;
  }
  for (int h = 0; h < 2; h++) {
    TetraSched &S = h ? P.tetra2 : P.tetra;
    S.ncoef = (int)pk.coef.size() / 2;
    S.nops = nops;
    for (int k = 0; k < nops; k++) {
      S.first[k] = pk.first[h][k];
      S.count[k] = pk.count[h][k];
      S.ndirect[k] = pk.ndirect[h][k];
    }
    S.terms = d_terms[h];
    S.coef = d_coef;
    S.accumulate = h;
  }
  choose_tiles(P);
  P.tpb = 64;  // sites per workgroup of the kernel
  if (!(rc = upload_plan(P, hs)) && !(rc = upload(&P.d_rows4, P.rows4, hs))) {
    std::lock_guard<std::mutex> wl(g_ws_mu);  // until the stream is synchronised below
    {
      std::lock_guard<std::mutex> lk(g_mu);
      if (!(rc = ensure(&g_ctx.ws_c, &g_ctx.ws_c_bytes, c_bytes(P))))
        rc = ensure(&g_ctx.ws_part, &g_ctx.ws_part_bytes, part_bytes(P));
    }
    if (!rc) rc = enqueue(P, g_ctx.ws_c, g_ctx.ws_part, out, out_precision, hs);
    const hipError_t se = hipStreamSynchronize(hs);  // also after an error
    if (!rc && se != hipSuccess) rc = fail(PLG_ERR_HIP, "tetraquark table sync failed");
  }
  hipStreamSynchronize(hs);  // no launch may still read the terms
  free_tables();
  return rc;
}

int plg_tetraquark_table(const plg_prop *p1, const plg_prop *p2, const plg_prop *p3, const plg_prop *p4,
                         const int16_t *idxs, const int16_t *col_contr, const float *vals, const int16_t *counts,
                         int nops, const int source[4], int totalT, const int *moms, int nmom, int mode, int space,
                         int out_precision, void *out, void *stream) {
  const plg_prop *views[4] = {p1, p2, p3, p4};
  return tetra_run(views, idxs, col_contr, vals, counts, nops, 0, source, totalT, moms, nmom, mode, space,
                   out_precision, out, stream);
}

// ------------------------------------------------------------------ built-in tetraquark tables
// heavy-light: lib/PLEGMA_Correlator.cu:192-250 (product selection, datasets)
// and lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:163-212; bcud:
// lib/PLEGMA_Correlator.cu:252-310 and lib/kernels/PLEGMA_bcud_tetraquarks.cu:
// 159-212.  One product after the other, propagator per flavour letter:
// u, d -> LT, s -> ST, c -> CH, b -> BT.
static const char TT_LETTERS[5] = {'u', 'd', 's', 'c', 'b'};
static const int TT_VIEW_OF[5] = {0, 0, 1, 2, 3};  // letter -> LT, ST, CH, BT argument

struct TetraSet {
  int nprod;
  const char *const *prod;
  const int *nops;
  const char *const *const *names;
  const int16_t *const *count, *const *idx, *const *col;
  const float *const *val;
};
static const TetraSet TETRA_HL = {TT_NPROD, TT_PROD, TT_NOPS, TT_NAMES, TT_COUNT, TT_IDX, TT_COL, TT_VAL};
static const TetraSet TETRA_BCUD = {TB_NPROD, TB_PROD, TB_NOPS, TB_NAMES, TB_COUNT, TB_IDX, TB_COL, TB_VAL};

static void tetra_select(const TetraSet &ts, const plg_prop *const fl[4], int only_st, int only_ch,
                         std::vector<int> &prods, int &nd) {
  prods.clear();
  nd = 0;
  for (int i = 0; i < ts.nprod; i++) {
    const std::string f = ts.prod[i];
    bool ok = true;
    for (int k = 0; k < 5; k++)
      if (!fl[TT_VIEW_OF[k]] && f.find(TT_LETTERS[k]) != std::string::npos) ok = false;
    if (only_st && f.find('s') == std::string::npos) ok = false;
    if (only_ch && f.find('c') == std::string::npos) ok = false;
    if (ok) {
      prods.push_back(i);
      nd += ts.nops[i];
    }
  }
}

static int tetra_count(const TetraSet &ts, int *nprod) {
  if (!nprod) return fail(PLG_ERR_INVALID, "NULL argument");
  *nprod = ts.nprod;
  return PLG_OK;
}

static int tetra_product(const TetraSet &ts, int prod, char *flavours, int *nops) {
  if (prod < 0 || prod >= ts.nprod) return fail(PLG_ERR_INVALID, "no such flavour product");
  if (flavours) std::memcpy(flavours, ts.prod[prod], 5);
  if (nops) *nops = ts.nops[prod];
  return PLG_OK;
}

static int tetra_op_name(const TetraSet &ts, int prod, int k, char *buf, int cap) {
  if (prod < 0 || prod >= ts.nprod || k < 0 || k >= ts.nops[prod] || !buf || cap <= 0)
    return fail(PLG_ERR_INVALID, "bad product, operator or buffer");
  const char *n = ts.names[prod][k];
  if ((int)std::strlen(n) >= cap) return fail(PLG_ERR_RANGE, "name buffer too small");
  std::strcpy(buf, n);
  return PLG_OK;
}

static int tetra_select_api(const TetraSet &ts, const plg_prop *lt, const plg_prop *st, const plg_prop *ch,
                            const plg_prop *bt, int only_st, int only_ch, int *prods, int cap, int *nprods,
                            int *ndatasets) {
  const plg_prop *fl[4] = {lt, st, ch, bt};
  std::vector<int> v;
  int nd;
  tetra_select(ts, fl, only_st, only_ch, v, nd);
  if (prods && (int)v.size() > cap) return fail(PLG_ERR_RANGE, "product buffer too small");
  if (prods) std::copy(v.begin(), v.end(), prods);
  if (nprods) *nprods = (int)v.size();
  if (ndatasets) *ndatasets = nd;
  return PLG_OK;
}

static int tetra_size(const TetraSet &ts, const plg_prop *lt, const plg_prop *st, const plg_prop *ch,
                      const plg_prop *bt, int only_st, int only_ch, const int source[4], int totalT, int nmom,
                      int space, int64_t *elems) {
  int ltT, nd, rc;
  if (!elems) return fail(PLG_ERR_INVALID, "NULL argument");
  if ((rc = plg_window(source, totalT, nullptr, nullptr, &ltT))) return rc;
  const plg_prop *fl[4] = {lt, st, ch, bt};
  std::vector<int> v;
  tetra_select(ts, fl, only_st, only_ch, v, nd);
  *elems = (int64_t)nd * ltT * (space == PLG_MOMENTUM_SPACE ? nmom : g_ctx.V3);
  return PLG_OK;
}

static int tetra_products(const TetraSet &ts, const plg_prop *lt, const plg_prop *st, const plg_prop *ch,
                          const plg_prop *bt, int only_st, int only_ch, const int source[4], int totalT,
                          const int *moms, int nmom, int mode, int space, int out_precision, void *out,
                          void *stream) {
  int rc;
  if ((rc = udsc_check_args(out, source, out_precision, mode, space, moms, nmom))) return rc;
  const plg_prop *fl[4] = {lt, st, ch, bt};
  std::vector<int> prods;
  int nd_total;
  tetra_select(ts, fl, only_st, only_ch, prods, nd_total);
  int d0 = 0;
  for (int prod : prods) {
    const char *f = ts.prod[prod];
    const plg_prop *views[4];
    for (int i = 0; i < 4; i++) views[i] = fl[TT_VIEW_OF[std::find(TT_LETTERS, TT_LETTERS + 5, f[i]) - TT_LETTERS]];
    if ((rc = tetra_run(views, ts.idx[prod], ts.col[prod], ts.val[prod], ts.count[prod], ts.nops[prod], d0, source,
                        totalT, moms, nmom, mode, space, out_precision, out, stream)))
      return rc;
    d0 += ts.nops[prod];
  }
  return PLG_OK;
}

int plg_tetra_count(int *nprod) { return tetra_count(TETRA_HL, nprod); }
int plg_tetra_product(int prod, char *flavours, int *nops) { return tetra_product(TETRA_HL, prod, flavours, nops); }
int plg_tetra_op_name(int prod, int k, char *buf, int cap) { return tetra_op_name(TETRA_HL, prod, k, buf, cap); }

int plg_tetraquarks_select(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt,
                           int only_st, int only_ch, int *prods, int cap, int *nprods, int *ndatasets) {
  return tetra_select_api(TETRA_HL, lt, st, ch, bt, only_st, only_ch, prods, cap, nprods, ndatasets);
}

int plg_tetraquarks_size(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt,
                         int only_st, int only_ch, const int source[4], int totalT, int nmom, int space,
                         int64_t *elems) {
  return tetra_size(TETRA_HL, lt, st, ch, bt, only_st, only_ch, source, totalT, nmom, space, elems);
}

int plg_tetraquarks(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt, int only_st,
                    int only_ch, const int source[4], int totalT, const int *moms, int nmom, int mode, int space,
                    int out_precision, void *out, void *stream) {
  return tetra_products(TETRA_HL, lt, st, ch, bt, only_st, only_ch, source, totalT, moms, nmom, mode, space,
                        out_precision, out, stream);
}

int plg_tetra_bcud_count(int *nprod) { return tetra_count(TETRA_BCUD, nprod); }
int plg_tetra_bcud_product(int prod, char *flavours, int *nops) {
  return tetra_product(TETRA_BCUD, prod, flavours, nops);
}
int plg_tetra_bcud_op_name(int prod, int k, char *buf, int cap) {
  return tetra_op_name(TETRA_BCUD, prod, k, buf, cap);
}

int plg_tetraquarks_bcud_select(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt,
                                int only_st, int only_ch, int *prods, int cap, int *nprods, int *ndatasets) {
  return tetra_select_api(TETRA_BCUD, lt, st, ch, bt, only_st, only_ch, prods, cap, nprods, ndatasets);
}

int plg_tetraquarks_bcud_size(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt,
                              int only_st, int only_ch, const int source[4], int totalT, int nmom, int space,
                              int64_t *elems) {
  return tetra_size(TETRA_BCUD, lt, st, ch, bt, only_st, only_ch, source, totalT, nmom, space, elems);
}

int plg_tetraquarks_bcud(const plg_prop *lt, const plg_prop *st, const plg_prop *ch, const plg_prop *bt,
                         int only_st, int only_ch, const int source[4], int totalT, const int *moms, int nmom,
                         int mode, int space, int out_precision, void *out, void *stream) {
  return tetra_products(TETRA_BCUD, lt, st, ch, bt, only_st, only_ch, source, totalT, moms, nmom, mode, space,
                        out_precision, out, stream);
}

int plg_baryons_table(const plg_prop *p1, const plg_prop *p2, const plg_prop *p3, const int16_t *idxs,
                      const float *vals, const int16_t *counts, int nops, const int source[4], int totalT,
                      const int *moms, int nmom, int mode, int space, int out_precision, void *out, void *stream) {
  int rc;
  if ((rc = udsc_check_args(out, source, out_precision, mode, space, moms, nmom))) return rc;
  if (!idxs || !vals || !counts || nops <= 0) return fail(PLG_ERR_INVALID, "empty table");
  std::vector<UdscTerm> terms;
  std::vector<int> first(nops), count(nops);
  for (int k = 0; k < nops; k++) {
    if (counts[k] < 0) return fail(PLG_ERR_INVALID, "negative term count");
    first[k] = (int)terms.size();
    count[k] = counts[k];
    for (int i = 0; i < counts[k]; i++) {
      const size_t t = terms.size();
      UdscTerm u;
      for (int j = 0; j < 6; j++) u.m[j] = (signed char)idxs[6 * t + j];
      u.re = vals[2 * t];
      u.im = vals[2 * t + 1];
      terms.push_back(u);
    }
  }
  UdscSchedHost h;
  rc = udsc_schedule_terms(terms.data(), first.data(), count.data(), nops, h);
  if (rc == PLG_ERR_INVALID) return fail(rc, "spin index outside 0..3");
  if (rc)
    return fail(rc, "table exceeds the kernel's limits (16 operators, 256 spin quads, 4 terms per quad, "
                    "64 distinct coefficients)");
  hipStream_t hs = (hipStream_t)stream;
  std::vector<void *> mem;
  auto put = [&](auto &v, auto **d) {
    const int r = upload(d, v, hs);
    if (!r) mem.push_back(*d);
    return r;
  };
  short *q = nullptr;
  signed char *tp = nullptr;
  uint16_t *e = nullptr;
  float *c = nullptr;
  if (!(rc = put(h.quads, &q)) && !(rc = put(h.tp, &tp)) && !(rc = put(h.ents, &e)) && !(rc = put(h.coef, &c))) {
    h.s.quads = q;
    h.s.tp = tp;
    h.s.ents = e;
    h.s.coef = c;
    const plg_prop *views[3] = {p1, p2, p3};
    rc = udsc_one(views, h.s, nops, 0, nops, source, totalT, moms, nmom, mode, space, out_precision, out, hs,
                  nullptr);
  }
  hipStreamSynchronize(hs);  // no launch may still read the schedule
  for (void *p : mem) hipFree(p);
  return rc;
}

int plg_baryons_udsc(const plg_prop *up, const plg_prop *dn, const plg_prop *st, const plg_prop *ch, int only_st,
                     int only_ch, const int source[4], int totalT, const int *moms, int nmom, int mode, int space,
                     int out_precision, void *out, void *stream) {
  return udsc_run(up, dn, st, ch, only_st, only_ch, source, totalT, moms, nmom, mode, space, out_precision, out,
                  stream, nullptr);
}

int plg_baryons_udsc_profile(const plg_prop *up, const plg_prop *dn, const plg_prop *st, const plg_prop *ch,
                             int only_st, int only_ch, const int source[4], int totalT, const int *moms, int nmom,
                             int mode, int space, int out_precision, void *out, void *stream, int cap, int *nstages,
                             double *ms, char *names) {
  if (!nstages) return fail(PLG_ERR_INVALID, "NULL argument");
  std::map<std::string, double> sm;
  const int rc = udsc_run(up, dn, st, ch, only_st, only_ch, source, totalT, moms, nmom, mode, space, out_precision,
                          out, stream, &sm);
  if (rc) return rc;
  *nstages = (int)sm.size();
  int i = 0;
  for (auto &kv : sm) {
    if (i < cap) {
      if (ms) ms[i] = kv.second;
      if (names) {
        std::memset(names + 32 * i, 0, 32);
        std::strncpy(names + 32 * i, kv.first.c_str(), 31);
      }
    }
    i++;
  }
  return PLG_OK;
}

int plg_ft_apply(const void *field, int precision, int dof, int nt, const int *moms, int nmom, int sign,
                 int accumulate, void *out, int out_precision, void *stream) {
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (!field || !out || !moms || nmom <= 0 || dof <= 0) return fail(PLG_ERR_INVALID, "NULL or empty argument");
  if (precision != PLG_F64 && precision != PLG_F32) return fail(PLG_ERR_INVALID, "bad field precision");
  if (out_precision != PLG_F64 && out_precision != PLG_F32) return fail(PLG_ERR_INVALID, "bad out precision");
  if (nt != 1 && nt != g_ctx.L[3]) return fail(PLG_ERR_INVALID, "nt must be 1 (3D field) or T (4D field)");
  if (sign != 1 && sign != -1) return fail(PLG_ERR_INVALID, "sign must be +1 or -1");
  hipStream_t st = (hipStream_t)stream;
  // exp(sign i 2pi n.x/L): the GEMM applies exp(-i ...), so sign = +1 negates the momenta
  const std::vector<int> ms = scaled_moms(moms, nmom, -sign);
  const char *sep_env = std::getenv("PLG_PROJ_SEP");
  // the separable sums read the field with 16-byte loads
  const bool sep_off = (sep_env && std::string(sep_env) == "0") || ((uintptr_t)field & 15);
  std::ostringstream key;
  key << precision << ' ' << dof << ' ' << nt << ' ' << sep_off;
  for (int i = 0; i < 4; i++) key << ' ' << g_ctx.L[i];
  for (int v : ms) key << ' ' << v;
  // the tables and the shared workspace are this call's until the stream is
  // synchronised below
  std::unique_lock<std::mutex> wl(g_ws_mu);
  FtTables &F = g_ft;
  ProjGeom g;
  g.V3 = g_ctx.V3;
  for (int i = 0; i < 3; i++) g.L[i] = g_ctx.L[i];
  g.M = g_ctx.M;
  g.nmom = nmom;
  g.ncols = dof * nt;
  g.cos_tab = g_ctx.d_cos;
  g.sin_tab = g_ctx.d_sin;
  int rc = PLG_OK;
  if (F.key != key.str()) {
    F.release();
    std::vector<int> pair_ms, mom_pair;
    F.npairs = make_pairs(ms, g_ctx.M, pair_ms, mom_pair, &F.zero_pair);
    // the separable sums where the momenta suit them (PLG_PROJ_SEP=0: the GEMM)
    if (!sep_off) sep_build(g_ctx.L, g_ctx.M, ms, (int64_t)dof * nt, precision == PLG_F64, F.sh);
    rc = upload(&F.d_ms, ms, st);
    if (!rc) rc = upload(&F.d_pair_ms, pair_ms, st);
    if (!rc) rc = upload(&F.d_mom_pair, mom_pair, st);
    if (!rc && F.sh.ok) rc = upload(&F.d_sep, F.sh.blob, st);
    F.key = key.str();
  }
  const int npairs = F.npairs, zero_pair = F.zero_pair;
  const SepHost &sh = F.sh;
  g.mom_scaled = F.d_ms;
  g.npairs = npairs;
  g.pair_ms = F.d_pair_ms;
  g.mom_pair = F.d_mom_pair;
  g.zero_pair = zero_pair;
  g.ctile = 0;
  int code;
  int64_t tiles;
  if (precision == PLG_F64) {
    const int R = 2 * npairs;
    code = R <= 96 ? 32 : 64;
    tiles = ((g.ncols + 63) / 64) * ((R + code - 1) / code);
  } else {
    code = pick_proj(std::max(npairs - zero_pair, 1));
    tiles = proj_workgroups(code, g.ncols, std::max(npairs - zero_pair, 1));
  }
  g.nsplit = (int)std::max<int64_t>(1, std::min<int64_t>((1024 + tiles - 1) / tiles, g_ctx.V3 / 512));
  if (!rc && !sh.ok && npairs > 0 && precision == PLG_F32) {
    if (!F.d_pidx) {
      if (hipMalloc((void **)&F.d_pidx, phase_index_bytes(g_ctx.V3, npairs)) != hipSuccess)
        rc = fail(PLG_ERR_HIP, "phase index allocation failed");
      else if ((rc = launch_phase_index(g, F.d_pidx, st)))
        rc = fail(rc, "phase index launch failed");
    }
    g.pidx = F.d_pidx;
  }
  if (!rc) {
    std::lock_guard<std::mutex> lk(g_mu);
    rc = ensure(&g_ctx.ws_part, &g_ctx.ws_part_bytes,
                sh.ok ? sep_part_bytes(sh.g) : proj_part_bytes(g.nsplit, g.ncols, nmom, npairs));
  }
  if (!rc && sh.ok) {
    // the separable sums read the field where it lies: 16-site groups of a
    // column are contiguous in its [column][site] layout too
    SepGeom sg = sep_geom(sh, F.d_sep);
    sg.gstride = 16;
    sg.cstride = g_ctx.V3;
    rc = launch_project_sep(sg, field, precision, (double *)g_ctx.ws_part, st);
    if (!rc) {
      ProjGeom gs = g;
      gs.sep = 1;
      gs.npairs = 0;
      gs.nsplit = 1;
      gs.ctile = 1;
      rc = launch_ft_out(gs, (const double *)g_ctx.ws_part, dof, nt, accumulate, out, out_precision, st);
    }
    if (rc) rc = fail(rc, "FT launch failed");
  } else if (!rc && precision == PLG_F32) {
    // the mixed projection reads the tile-major layout of the site correlator
    std::lock_guard<std::mutex> lk(g_mu);
    rc = ensure(&g_ctx.ws_c, &g_ctx.ws_c_bytes, (size_t)g.ncols * ctile_sites(g_ctx.V3) * sizeof(cx<float>));
    if (!rc) rc = launch_to_tiles((const cx<float> *)field, g.ncols, g_ctx.V3, (cx<float> *)g_ctx.ws_c, st);
    g.ctile = 1;
    g.shifted = 1;
  }
  if (!rc && !sh.ok) {
    if (precision == PLG_F64)
      rc = launch_project_f64(g, (const cx<double> *)field, (cx<double> *)g_ctx.ws_part, code, st);
    else
      rc = launch_project_bf16(g, (const cx<float> *)g_ctx.ws_c, (double *)g_ctx.ws_part, PLG_SPLIT_BF16X6,
                               g_ctx.d_wtab, code, st);
    if (!rc) rc = launch_ft_out(g, (const double *)g_ctx.ws_part, dof, nt, accumulate, out, out_precision, st);
    if (rc) rc = fail(rc, "FT launch failed");
  }
  // the workspace is shared: wait for every launch, also after an error
  const hipError_t se = hipStreamSynchronize(st);
  if (!rc && se != hipSuccess) rc = fail(PLG_ERR_HIP, "FT synchronisation failed");
  if (rc) F.release();  // a table set that failed to build is not kept
  return rc;
}

int plg_batch_create(int kind, int group_mask, const plg_prop *props, int nprops, const plg_job *jobs, int njobs,
                     const int *moms, int nmom, int mode, int split, int flags, plg_batch *out) {
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (!props || nprops <= 0 || !jobs || njobs <= 0 || !out) return fail(PLG_ERR_INVALID, "empty batch");
  std::vector<plg_prop> pv(props, props + nprops);
  std::vector<JobIn> jv;
  for (int i = 0; i < njobs; i++) {
    const plg_job &j = jobs[i];
    if (j.prop < 0 || j.prop >= nprops) return fail(PLG_ERR_INVALID, "job propagator index out of range");
    if (j.prop2 < -1 || j.prop2 >= nprops) return fail(PLG_ERR_INVALID, "job prop2 index out of range");
    int lt;
    int rc = plg_window(j.source, j.totalT, nullptr, nullptr, &lt);
    if (rc) return rc;
    jv.push_back({{j.source[0], j.source[1], j.source[2], j.source[3]}, j.totalT, j.prop,
                  j.prop2 < 0 ? j.prop : j.prop2, j.out_offset});
  }
  plg_batch b = new plg_batch_s();
  Plan &P = b->plan;
  int rc = build_plan(P, kind, group_mask, pv, jv, moms, nmom, mode, split, PLG_MOMENTUM_SPACE);
  if (!rc) {
    choose_tiles(P);
    rc = upload_plan(P, nullptr);
  }
  if (!rc) {
    P.own_ws = true;
    rc = ensure(&P.c, &P.c_bytes, c_bytes(P));
    if (!rc && (flags & PLG_BATCH_TUNE)) {
      // room for every candidate (the largest K split, the separable form)
      rc = ensure(&P.part, &P.part_bytes, part_bytes_tune(P));
      if (!rc) rc = autotune(P, nullptr);
    }
    if (!rc) rc = ensure(&P.part, &P.part_bytes, part_bytes(P));
  }
  if (!rc && (flags & PLG_BATCH_GRAPH)) {
    if (hipStreamCreateWithFlags(&P.cap_stream, hipStreamNonBlocking) != hipSuccess)
      rc = fail(PLG_ERR_HIP, "stream creation failed");
  }
  if (rc) {
    delete b;
    return rc;
  }
  hipDeviceSynchronize();
  *out = b;
  return PLG_OK;
}

int plg_batch_run(plg_batch b, void *out, void *stream) {
  if (!b || !out) return fail(PLG_ERR_INVALID, "NULL argument");
  Plan &P = b->plan;
  hipStream_t st = (hipStream_t)stream;
  if (!P.cap_stream) return enqueue(P, P.c, P.part, out, PLG_F64, st);
  if (!P.exec || P.graph_out != out) {
    if (P.exec) hipGraphExecDestroy(P.exec);
    if (P.graph) hipGraphDestroy(P.graph);
    P.exec = nullptr;
    P.graph = nullptr;
    HIPCHK(hipStreamBeginCapture(P.cap_stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue(P, P.c, P.part, out, PLG_F64, P.cap_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(P.cap_stream, &g);
    if (rc) {
      if (g) hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) return fail(PLG_ERR_HIP, std::string("graph capture: ") + hipGetErrorString(e));
    P.graph = g;
    HIPCHK(hipGraphInstantiate(&P.exec, P.graph, nullptr, nullptr, 0));
    P.graph_out = out;
  }
  HIPCHK(hipGraphLaunch(P.exec, st));
  return PLG_OK;
}

int plg_batch_profile(plg_batch b, void *out, void *stream, int cap, int *nstages, double *ms, char *names) {
  if (!b || !out || !nstages) return fail(PLG_ERR_INVALID, "NULL argument");
  Plan &P = b->plan;
  hipStream_t st = (hipStream_t)stream;
  StageLog log;
  int rc = enqueue(P, P.c, P.part, out, PLG_F64, st, &log);
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = fail(PLG_ERR_HIP, "profile sync failed");
  const int n = (int)log.ev.size() - 1;
  if (!rc) {
    *nstages = n;
    for (int i = 0; i < n && i < cap; i++) {
      float t = 0;
      hipEventElapsedTime(&t, log.ev[i], log.ev[i + 1]);
      if (ms) ms[i] = t;
      if (names) {
        std::memset(names + 32 * i, 0, 32);
        std::strncpy(names + 32 * i, log.name[i + 1].c_str(), 31);
      }
    }
  }
  for (hipEvent_t e : log.ev) hipEventDestroy(e);
  return rc;
}

int plg_batch_info(plg_batch b, int64_t *out_elems, int64_t *nsites, int *nrows) {
  if (!b) return fail(PLG_ERR_INVALID, "NULL batch");
  if (out_elems) *out_elems = b->plan.out_elems;
  if (nsites) *nsites = b->plan.nsites;
  if (nrows) *nrows = (int)b->plan.rows.size();
  return PLG_OK;
}

int plg_batch_destroy(plg_batch b) {
  if (b) {
    hipDeviceSynchronize();
    delete b;
  }
  return PLG_OK;
}

int plg_synth_prop(const plg_prop *dst, uint64_t seed, uint32_t stream_id, int t0, double lambda, void *stream) {
  return plg_synth_prop_perturbed(dst, seed, stream_id, 0, 0.0, t0, lambda, stream);
}

int plg_synth_prop_perturbed(const plg_prop *dst, uint64_t seed, uint32_t base_stream, uint32_t pert_stream,
                             double eps, int t0, double lambda, void *stream) {
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (!dst || !dst->data) return fail(PLG_ERR_INVALID, "NULL propagator");
  if (dst->nt <= 0 || dst->nt > g_ctx.L[3] || dst->comp_stride < (int64_t)dst->nt * g_ctx.V3)
    return fail(PLG_ERR_INVALID, "bad propagator view");
  int rc = launch_synth(*dst, g_ctx.L, seed, base_stream, pert_stream, eps, t0, lambda, (hipStream_t)stream);
  return rc ? fail(rc, "synth launch failed") : PLG_OK;
}

int plg_cast_prop(const plg_prop *dst, const plg_prop *src, void *stream) {
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (!dst || !src || dst->precision != PLG_F32 || src->precision != PLG_F64)
    return fail(PLG_ERR_INVALID, "cast copies an F64 view into an F32 view");
  for (int pl = 0; pl < dst->nt; pl++) {
    const int t = (dst->t_first + pl) % g_ctx.L[3];
    int rc = check_prop(*src, PLG_F64, t);
    if (rc) return rc;
  }
  int rc = launch_cast(*dst, *src, g_ctx.V3, g_ctx.L[3], (hipStream_t)stream);
  return rc ? fail(rc, "cast launch failed") : PLG_OK;
}

int plg_prop_rotate_physical(const plg_prop *p, int sign, void *stream) {
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (!p || !p->data) return fail(PLG_ERR_INVALID, "NULL propagator");
  if (sign != 1 && sign != -1) return fail(PLG_ERR_INVALID, "the sign can be only +-1");
  if (p->precision != PLG_F64 && p->precision != PLG_F32) return fail(PLG_ERR_INVALID, "bad precision");
  if (p->nt <= 0 || p->nt > g_ctx.L[3] || p->comp_stride < (int64_t)p->nt * g_ctx.V3)
    return fail(PLG_ERR_INVALID, "propagator view: bad nt/comp_stride");
  const int rc = launch_prop_rotate(*p, g_ctx.V3, sign, (hipStream_t)stream);
  return rc ? fail(rc, "rotation launch failed") : PLG_OK;
}

int plg_prop_apply_boundaries(const plg_prop *p, int t0, void *stream) {
  if (!g_ctx.init) return fail(PLG_ERR_NOT_INIT, "plg_init not called");
  if (!p || !p->data) return fail(PLG_ERR_INVALID, "NULL propagator");
  if (p->precision != PLG_F64 && p->precision != PLG_F32) return fail(PLG_ERR_INVALID, "bad precision");
  if (p->nt <= 0 || p->nt > g_ctx.L[3] || p->comp_stride < (int64_t)p->nt * g_ctx.V3)
    return fail(PLG_ERR_INVALID, "propagator view: bad nt/comp_stride");
  // planes whose global time is below t0
  std::vector<int> planes;
  for (int pl = 0; pl < p->nt; pl++)
    if ((p->t_first + pl) % g_ctx.L[3] < t0) planes.push_back(pl);
  if (planes.empty()) return PLG_OK;
  hipStream_t st = (hipStream_t)stream;
  int *d_planes = nullptr;
  int rc = upload(&d_planes, planes, st);
  if (!rc) rc = launch_prop_negate(*p, g_ctx.V3, d_planes, (int)planes.size(), st);
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = PLG_ERR_HIP;
  if (d_planes) hipFree(d_planes);
  return rc ? fail(rc, "boundary launch failed") : PLG_OK;
}

}  // extern "C"
