// PLEGMA_Correlator.h -- the two-point correlator with the interface of
// include/PLEGMA_Correlator.h:18-331 for one process on one GPU: window
// accessors, dataset/group bookkeeping, H_elem in the reference's layout,
// contract* overwriting the buffer, the ASCII writer and the HDF5 writer
// (objects of the reference's HDF5 file in the record file of
// PLEGMA_h5twin.h, as this build has no HDF5 library).  Every contraction runs
// in libplegma_amd.so through plg_correlator: Float = double is the FP64
// path (PLG_MODE_F64), Float = float the mixed path (FP32 contraction,
// split-bf16 MFMA projection, FP64 sums).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>

#include "PLEGMA_FT.h"
#include "PLEGMA_Propagator.h"
#include "PLEGMA_h5twin.h"

namespace plegma {

template <typename Float> class PLEGMA_Correlator {
protected:
  const site source;
  const int totalT;
  std::vector<int> shape;
  std::vector<std::string> datasets;
  std::vector<std::string> groups;
  std::string description;
  const CORR_SPACE corr_space;
  std::shared_ptr<Float> corr_pos_space;  // host buffer of position-space results
  size_t pos_size = 0;
  std::shared_ptr<PLEGMA_FT<Float>> corr_mom_space;
  void *d_out = nullptr;  // device result buffer, reused across contractions
  size_t d_bytes = 0;

  int L(int d) const { return Lattice::get().L[d]; }

  // lib/PLEGMA_Correlator.cu:25-41: (re)allocate and zero the host buffer
  void initialize() {
    if (corr_space == MOMENTUM_SPACE) {
      corr_mom_space->checkAllocation((int)getSiteSize());
    } else {
      pos_size = getTotalSize() * 2;
      corr_pos_space.reset(new Float[pos_size](), std::default_delete<Float[]>());
    }
  }

  void run(int kind, int mask, const PLEGMA_Propagator<Float> &p1, const PLEGMA_Propagator<Float> &p2) {
    if (p1.getAllocation() == NONE || p2.getAllocation() == NONE) PLEGMA_error("propagator is not allocated");
    int64_t n = 0;
    std::vector<int> moms = corr_space == MOMENTUM_SPACE ? corr_mom_space->intMomenta() : std::vector<int>{};
    const int nmom = (int)moms.size() / 3;
    const int space = corr_space == MOMENTUM_SPACE ? PLG_MOMENTUM_SPACE : PLG_POSITION_SPACE;
    plg_check(plg_correlator_size(kind, mask, source.data(), totalT, nmom, space, &n), "plg_correlator_size");
    if ((size_t)n != getTotalSize())  // PLEGMA_mesons.cuh:140-141
      PLEGMA_error("Site size of the correlator does not match the contraction (%lld vs %zu)", (long long)n,
                   getTotalSize());
    const size_t bytes = (size_t)n * 2 * sizeof(Float);
    if (bytes > d_bytes) {
      if (d_out) (void)hipFree(d_out);
      hip_check(hipMalloc(&d_out, bytes), "correlator result buffer");
      d_bytes = bytes;
    }
    plg_prop v1 = p1.view(), v2 = p2.view();
    const int mode = sizeof(Float) == 8 ? PLG_MODE_F64 : PLG_MODE_MIXED;
    plg_check(plg_correlator(kind, mask, &v1, &v2, source.data(), totalT, moms.empty() ? nullptr : moms.data(), nmom,
                             mode, space, plg_precision_of<Float>(), d_out, nullptr),
              "plg_correlator");
    hip_check(hipMemcpy(H_elem(), d_out, bytes, hipMemcpyDeviceToHost), "result copy");
  }

public:
  // include/PLEGMA_Correlator.h:41-79 with one process (procPosition = 0)
  bool hasSource(int dir) const { return dir >= 0 && dir < 4 && source[dir] >= 0 && source[dir] < L(dir); }
  bool hasSource() const {
    bool r = true;
    for (int d = 0; r && d < 4; d++) r &= hasSource(d);
    return r;
  }
  int startT() const {
    int s = 0, e = 0, lt = 0;
    plg_check(plg_window(source.data(), totalT, &s, &e, &lt), "plg_window");
    return s;
  }
  int endT() const {
    int s = 0, e = 0, lt = 0;
    plg_check(plg_window(source.data(), totalT, &s, &e, &lt), "plg_window");
    return e;
  }
  int localT() const {
    int s = 0, e = 0, lt = 0;
    plg_check(plg_window(source.data(), totalT, &s, &e, &lt), "plg_window");
    return lt;
  }

  // include/PLEGMA_Correlator.h:81-85
  PLEGMA_Correlator(CORR_SPACE corr_space, site source, int Q2_max = 0, int totalT = Lattice::get().L[3])
      : source(source), totalT(totalT), corr_space(corr_space) {
    if (!Lattice::get().init) PLEGMA_error("PLEGMA_amd_init must be called before creating correlators");
    if (corr_space == MOMENTUM_SPACE) corr_mom_space.reset(new PLEGMA_FT<Float>(Q2_max, 3, false, localT()));
  }
  PLEGMA_Correlator(const PLEGMA_Correlator &) = delete;
  PLEGMA_Correlator &operator=(const PLEGMA_Correlator &) = delete;
  ~PLEGMA_Correlator() {
    if (d_out) (void)hipFree(d_out);
  }

  CORR_SPACE getCorrSpace() const { return corr_space; }
  Float *H_elem() const { return corr_space == MOMENTUM_SPACE ? corr_mom_space->H_elem() : corr_pos_space.get(); }
  site getSource() { return source; }
  int getTotalT() { return totalT; }
  size_t nDatasets() const { return std::max(datasets.size(), (size_t)1); }
  size_t nGroups() const { return std::max(groups.size(), (size_t)1); }
  // include/PLEGMA_Correlator.h:113-127
  size_t getSiteSize() const {
    size_t s = nDatasets() * nGroups();
    for (int n : shape) s *= n;
    return s;
  }
  size_t getVolSize() const {
    if (corr_space == MOMENTUM_SPACE) return (size_t)corr_mom_space->Nmoms() * corr_mom_space->DimT();
    return (size_t)Lattice::V3() * localT();
  }
  size_t getTotalSize() const { return getSiteSize() * getVolSize(); }
  void setQ2max(int Q2_max) { corr_mom_space.reset(new PLEGMA_FT<Float>(Q2_max, 3, false, localT())); }
  void setFixMomVec(std::vector<int> &v) { corr_mom_space.reset(new PLEGMA_FT<Float>(v, 3, false, localT())); }
  void setFixMomList(std::vector<std::vector<int>> &l) {
    corr_mom_space.reset(new PLEGMA_FT<Float>(l, 3, false, localT()));
  }
  std::vector<std::vector<int>> getMomList() {
    std::vector<std::vector<int>> out;
    if (corr_space != MOMENTUM_SPACE) return out;
    for (auto &m : corr_mom_space->MomList()) {
      out.push_back({});
      for (auto x : m) out.back().push_back((int)std::lround(x));
    }
    return out;
  }
  std::vector<std::string> getDatasets() const { return datasets; }
  std::vector<std::string> getGroups() const { return groups; }
  void setDatasets(std::vector<std::string> d) {
    if (datasets.size() == d.size() || (datasets.empty() && d.size() == 1)) datasets = d;
    else PLEGMA_error("Given vector size do not match. This would change the correlator size.");
  }
  void setDatasets(std::string s) { setDatasets(std::vector<std::string>{s}); }
  void setGroups(std::vector<std::string> d) {
    if (groups.size() == d.size() || (groups.empty() && d.size() == 1)) groups = d;
    else PLEGMA_error("Given vector size do not match. This would change the correlator size.");
  }
  void setGroups(std::string s) { setGroups(std::vector<std::string>{s}); }
  std::string getDescription() const { return description; }

  // ---- contractions (lib/PLEGMA_Correlator.cu:43-135)
  void contractMesons(PLEGMA_Propagator<Float> &prop1, PLEGMA_Propagator<Float> &prop2) {
    shape = {10};
    datasets = {"twop_meson_1", "twop_meson_2"};
    groups = {"mesons"};
    description = "pseudoscalar, scalar, g5g1, g5g2, g5g3, g5g4, g1, g2, g3, g4";
    initialize();
    run(PLG_MESONS, 0, prop1, prop2);
  }
  void contractMesonsNew(PLEGMA_Propagator<Float> &prop1, PLEGMA_Propagator<Float> &prop2) {
    shape = {10};
    datasets = {"twop_meson"};
    groups = {"mesons"};
    description = "pseudoscalar, scalar, g5g1, g5g2, g5g3, g5g4, g1, g2, g3, g4";
    initialize();
    run(PLG_MESONS_NEW, 0, prop1, prop2);
  }
  void contractMesonsAll(PLEGMA_Propagator<Float> &prop1, PLEGMA_Propagator<Float> &prop2) {
    shape = {34};
    datasets = {"twop_meson"};
    groups = {"mesons"};
    description = "pseudoscalar, scalar";
    for (int p = 0; p < 2; p++) {
      const std::string va = p ? "" : "g5";
      for (int i = 1; i < 5; i++)
        for (int j = 1; j < 5; j++)
          description += ", (" + va + "g" + std::to_string(i) + "," + va + "g" + std::to_string(j) + ")";
    }
    initialize();
    run(PLG_MESONS_ALL, 0, prop1, prop2);
  }
  // groups: nucl_nucl; with group_mask also deltapp_deltamm_{11,22,33}
  // (PLG_G_* bits; the reference's PLEGMA_LIGHT_BARYONS build adds them)
  void contractBaryons(PLEGMA_Propagator<Float> &prop1, PLEGMA_Propagator<Float> &prop2,
                       int group_mask = PLG_G_NUCL) {
    shape = {16};
    datasets = {"twop_baryon_1", "twop_baryon_2"};
    groups.clear();
    const char *names[10] = {"baryons/nucl_nucl", "baryons/nucl_nucl2", "baryons/nucl2_nucl",
                             "baryons/nucl2_nucl2", "baryons/deltap_deltaz_11", "baryons/deltap_deltaz_22",
                             "baryons/deltap_deltaz_33", "baryons/deltapp_deltamm_11",
                             "baryons/deltapp_deltamm_22", "baryons/deltapp_deltamm_33"};
    for (int b = 0; b < 10; b++)
      if (group_mask & (1 << b)) groups.push_back(names[b]);
    description = "1,g1,g2,g3,g4,g5,g5g1,g5g2,g5g3,g5g4,s12,s13,s23,s41,s42,s43";
    initialize();
    run(PLG_BARYONS, group_mask, prop1, prop2);
  }
  // ---- kinds without a single reference function
  // 16 mesons + proton/Delta++ groups on one propagator (the fused kernel of
  // the mixed path); datasets: one
  void contractTwoPoint(PLEGMA_Propagator<Float> &prop, int group_mask = PLG_G_NUCL | PLG_G_DELTAPP_ALL) {
    shape = {16};
    datasets = {"twop"};
    groups = {"mesons16"};
    const char *names[4] = {"baryons/nucl_nucl", "baryons/deltapp_deltamm_11", "baryons/deltapp_deltamm_22",
                            "baryons/deltapp_deltamm_33"};
    const int bits[4] = {0, 7, 8, 9};
    for (int k = 0; k < 4; k++)
      if (group_mask & (1 << bits[k])) groups.push_back(names[k]);
    description = "mesons: GAMMAS order; baryons: mu*4 + mu'";
    initialize();
    run(PLG_TWOPT, group_mask, prop, prop);
  }
  // heavy-light mesons and the four-propagator meson-meson block, the
  // counterpart of contractTetraquarks (lib/kernels/PLEGMA_heavy_light_tetraquarks.cu:163-212)
  void contractMesonMeson(PLEGMA_Propagator<Float> &propLight, PLEGMA_Propagator<Float> &propHeavy) {
    shape = {20};
    datasets = {"meson_meson"};
    groups = {"heavy_light"};
    description = "Tr[G SL g5 G SH^dag g5] for the 16 GAMMAS, then (Tr A)^2 - Tr[A A] for G = g5, g1, g2, g3";
    initialize();
    run(PLG_MESON_MESON, 0, propLight, propHeavy);
  }

  // lib/PLEGMA_Correlator.cu:136-188: every flavour product ("ucc" ... "sss")
  // whose propagators are allocated (NONE = flavour absent), optionally only
  // those holding s / c; datasets = the products' operator names in order
  void contractBaryonsUDSC(PLEGMA_Propagator<Float> &propUP, PLEGMA_Propagator<Float> &propDN,
                           PLEGMA_Propagator<Float> &propST, PLEGMA_Propagator<Float> &propCH,
                           bool only_st = false, bool only_ch = false) {
    shape = {};
    description = "";
    groups = {};
    datasets.clear();
    PLEGMA_Propagator<Float> *fl[4] = {&propUP, &propDN, &propST, &propCH};
    plg_prop v[4];
    const plg_prop *vp[4];
    for (int i = 0; i < 4; i++) {
      vp[i] = nullptr;
      if (fl[i]->getAllocation() != NONE) {
        v[i] = fl[i]->view();
        vp[i] = &v[i];
      }
    }
    int prods[64], np = 0, nd = 0;
    plg_check(plg_baryons_udsc_select(vp[0], vp[1], vp[2], vp[3], only_st, only_ch, prods, 64, &np, &nd),
              "plg_baryons_udsc_select");
    for (int i = 0; i < np; i++) {
      int nops = 0;
      plg_check(plg_udsc_product(prods[i], nullptr, &nops), "plg_udsc_product");
      for (int k = 0; k < nops; k++) {
        char name[64];
        plg_check(plg_udsc_op_name(prods[i], k, name, sizeof(name)), "plg_udsc_op_name");
        datasets.push_back(name);
      }
    }
    if (datasets.empty()) return;
    initialize();
    std::vector<int> moms = corr_space == MOMENTUM_SPACE ? corr_mom_space->intMomenta() : std::vector<int>{};
    const int nmom = (int)moms.size() / 3;
    const int space = corr_space == MOMENTUM_SPACE ? PLG_MOMENTUM_SPACE : PLG_POSITION_SPACE;
    int64_t n = 0;
    plg_check(plg_baryons_udsc_size(vp[0], vp[1], vp[2], vp[3], only_st, only_ch, source.data(), totalT, nmom,
                                    space, &n),
              "plg_baryons_udsc_size");
    if ((size_t)n != getTotalSize()) PLEGMA_error("udsc correlator size mismatch (%lld vs %zu)", (long long)n,
                                                  getTotalSize());
    const size_t bytes = (size_t)n * 2 * sizeof(Float);
    if (bytes > d_bytes) {
      if (d_out) (void)hipFree(d_out);
      hip_check(hipMalloc(&d_out, bytes), "correlator result buffer");
      d_bytes = bytes;
    }
    const int mode = sizeof(Float) == 8 ? PLG_MODE_F64 : PLG_MODE_MIXED;
    plg_check(plg_baryons_udsc(vp[0], vp[1], vp[2], vp[3], only_st, only_ch, source.data(), totalT,
                               moms.empty() ? nullptr : moms.data(), nmom, mode, space, plg_precision_of<Float>(),
                               d_out, nullptr),
              "plg_baryons_udsc");
    hip_check(hipMemcpy(H_elem(), d_out, bytes, hipMemcpyDeviceToHost), "result copy");
  }

  // lib/PLEGMA_Correlator.cu:192-250: every heavy-light tetraquark product
  // ("udbb", "usbb") whose propagators are allocated (NONE = flavour absent:
  // u, d read propLT, s propST, c propCH, b propBT), optionally only those
  // holding s / c; datasets = the products' operator names in order.  A
  // double propagator runs the FP64 path, a float one the mixed path; the
  // result is stored in Float.
  template <typename Float2>
  void contractTetraquarks(PLEGMA_Propagator<Float2> &propLT, PLEGMA_Propagator<Float2> &propST,
                           PLEGMA_Propagator<Float2> &propCH, PLEGMA_Propagator<Float2> &propBT,
                           bool only_st = false, bool only_ch = false) {
    static const TetraFamily hl = {plg_tetraquarks_select, plg_tetra_product, plg_tetra_op_name,
                                   plg_tetraquarks_size, plg_tetraquarks, "plg_tetraquarks"};
    tetraquarks(hl, propLT, propST, propCH, propBT, only_st, only_ch);
  }

  // lib/PLEGMA_Correlator.cu:252-310: the bcud tetraquark product ("udcb",
  // J = 1 and J = 0 operators), same selection rules and storage
  template <typename Float2>
  void contractTetraquarksBCUD(PLEGMA_Propagator<Float2> &propLT, PLEGMA_Propagator<Float2> &propST,
                               PLEGMA_Propagator<Float2> &propCH, PLEGMA_Propagator<Float2> &propBT,
                               bool only_st = false, bool only_ch = false) {
    static const TetraFamily bcud = {plg_tetraquarks_bcud_select, plg_tetra_bcud_product, plg_tetra_bcud_op_name,
                                     plg_tetraquarks_bcud_size, plg_tetraquarks_bcud, "plg_tetraquarks_bcud"};
    tetraquarks(bcud, propLT, propST, propCH, propBT, only_st, only_ch);
  }

 private:
  // the C entry points of one family of built-in tetraquark tables
  struct TetraFamily {
    int (*select)(const plg_prop *, const plg_prop *, const plg_prop *, const plg_prop *, int, int, int *, int,
                  int *, int *);
    int (*product)(int, char *, int *);
    int (*op_name)(int, int, char *, int);
    int (*size)(const plg_prop *, const plg_prop *, const plg_prop *, const plg_prop *, int, int, const int *,
                int, int, int, int64_t *);
    int (*run)(const plg_prop *, const plg_prop *, const plg_prop *, const plg_prop *, int, int, const int *, int,
               const int *, int, int, int, int, void *, void *);
    const char *what;
  };

  template <typename Float2>
  void tetraquarks(const TetraFamily &f, PLEGMA_Propagator<Float2> &propLT, PLEGMA_Propagator<Float2> &propST,
                   PLEGMA_Propagator<Float2> &propCH, PLEGMA_Propagator<Float2> &propBT, bool only_st,
                   bool only_ch) {
    shape = {};
    description = "";
    groups = {};
    datasets.clear();
    PLEGMA_Propagator<Float2> *fl[4] = {&propLT, &propST, &propCH, &propBT};
    plg_prop v[4];
    const plg_prop *vp[4];
    for (int i = 0; i < 4; i++) {
      vp[i] = nullptr;
      if (fl[i]->getAllocation() != NONE) {
        v[i] = fl[i]->view();
        vp[i] = &v[i];
      }
    }
    int prods[32], np = 0, nd = 0;
    plg_check(f.select(vp[0], vp[1], vp[2], vp[3], only_st, only_ch, prods, 32, &np, &nd), f.what);
    for (int i = 0; i < np; i++) {
      int nops = 0;
      plg_check(f.product(prods[i], nullptr, &nops), f.what);
      for (int k = 0; k < nops; k++) {
        char name[64];
        plg_check(f.op_name(prods[i], k, name, sizeof(name)), f.what);
        datasets.push_back(name);
      }
    }
    if (datasets.empty()) return;
    initialize();
    std::vector<int> moms = corr_space == MOMENTUM_SPACE ? corr_mom_space->intMomenta() : std::vector<int>{};
    const int nmom = (int)moms.size() / 3;
    const int space = corr_space == MOMENTUM_SPACE ? PLG_MOMENTUM_SPACE : PLG_POSITION_SPACE;
    int64_t n = 0;
    plg_check(f.size(vp[0], vp[1], vp[2], vp[3], only_st, only_ch, source.data(), totalT, nmom, space, &n), f.what);
    if ((size_t)n != getTotalSize()) PLEGMA_error("tetraquark correlator size mismatch (%lld vs %zu)", (long long)n,
                                                  getTotalSize());
    const size_t bytes = (size_t)n * 2 * sizeof(Float);
    if (bytes > d_bytes) {
      if (d_out) (void)hipFree(d_out);
      hip_check(hipMalloc(&d_out, bytes), "correlator result buffer");
      d_bytes = bytes;
    }
    const int mode = sizeof(Float2) == 8 ? PLG_MODE_F64 : PLG_MODE_MIXED;
    plg_check(f.run(vp[0], vp[1], vp[2], vp[3], only_st, only_ch, source.data(), totalT,
                    moms.empty() ? nullptr : moms.data(), nmom, mode, space, plg_precision_of<Float>(), d_out,
                    nullptr),
              f.what);
    hip_check(hipMemcpy(H_elem(), d_out, bytes, hipMemcpyDeviceToHost), "result copy");
  }

 public:
  // ---- output (lib/PLEGMA_Correlator.cu:721-806, momentum space)
  void writeASCII(std::string filename) const {
    if (corr_space != MOMENTUM_SPACE) PLEGMA_error("WriteASCII do not support writing in position space.");
    const int T = L(3), Nm = corr_mom_space->Nmoms(), lt = localT();
    const size_t ss = getSiteSize(), sr = ss / (nDatasets() * nGroups());
    char tail[64];
    std::snprintf(tail, sizeof(tail), "_sx%02dsy%02dsz%02dst%02d.dat", source[0], source[1], source[2], source[3]);
    FILE *f = std::fopen((filename + tail).c_str(), "w");
    if (!f) PLEGMA_error("Error opening file for writing");
    const auto moms = corr_mom_space->MomList();
    const Float *H = H_elem();
    for (int it = 0; it < T; it++) {
      const int t = (it + source[3]) % T;  // lattice time of line it
      int j = -1;                         // buffer row holding t
      for (int r = 0; r < lt && j < 0; r++) {
        int tr = 0;
        plg_check(plg_row_time(source.data(), totalT, r, &tr), "plg_row_time");
        if (tr == t) j = r;
      }
      if (j < 0) continue;
      for (int m = 0; m < Nm; m++) {
        std::fprintf(f, "%d  %+d  %+d  %+d ", it, (int)std::round(moms[m][0]), (int)std::round(moms[m][1]),
                     (int)std::round(moms[m][2]));
        for (size_t id = 0; id < nDatasets(); id++)
          for (size_t ig = 0; ig < nGroups(); ig++)
            for (size_t is = 0; is < sr; is++) {
              const size_t k = (((ig * nDatasets() + id) * lt + j) * Nm + m) * sr + is;
              std::fprintf(f, "%+e %+eI ", (double)H[2 * k], (double)H[2 * k + 1]);
            }
        std::fprintf(f, "\n");
      }
    }
    std::fclose(f);
  }
  // lib/PLEGMA_Correlator.cu:809-850 (fill_H5_shapes) and 893-946
  // (writeHDF5) into the record file of PLEGMA_h5twin.h.  Per group
  // /sxXXsyXXszXXstXX/<group>: "mvec" (Nmom x 3 int, momentum space) and one
  // dataset per name of shape (totalT, Nmom | Lz, Ly, Lx, shape..., 2) with
  // its "description" attribute.  The time axis counts from the source time
  // (tau = t - t0 mod T) and the spatial axes from the source position,
  // where the reference's start offsets with periodic wrap put them
  // (PLEGMA_hdf5.h:332-379).
  void writeHDF5(std::string filename) const {
    const bool mom = corr_space == MOMENTUM_SPACE;
    const int T = L(3), lt = localT();
    const int Lx = L(0), Ly = L(1), Lz = L(2);
    const size_t npt = mom ? (size_t)corr_mom_space->Nmoms() : (size_t)Lattice::V3();
    const size_t sr = getSiteSize() / (nDatasets() * nGroups());
    std::vector<uint64_t> dims = {(uint64_t)totalT};
    std::string descr = "shape: ";
    if (mom) {
      descr += "/time/moms";
      dims.push_back(npt);
    } else {
      descr += "/t/z/y/x";
      dims.insert(dims.end(), {(uint64_t)Lz, (uint64_t)Ly, (uint64_t)Lx});
    }
    if (!shape.empty()) {
      descr += "/" + description;
      for (int n : shape) dims.push_back((uint64_t)n);
    }
    descr += "/re-im";
    dims.push_back(2);
    // destination of buffer row j and point x in the dataset
    std::vector<size_t> row_dst(lt), pt_dst(npt);
    for (int j = 0; j < lt; j++) {
      int t = 0;
      plg_check(plg_row_time(source.data(), totalT, j, &t), "plg_row_time");
      const int tau = ((t - source[3]) % T + T) % T;
      if (tau >= totalT) PLEGMA_error("row %d (t = %d) lies outside the correlator's time extent", j, t);
      row_dst[j] = (size_t)tau;
    }
    for (size_t x = 0; x < npt; x++) {
      if (mom) {
        pt_dst[x] = x;
      } else {
        const int ix = (int)(x % Lx), iy = (int)(x / Lx % Ly), iz = (int)(x / ((size_t)Lx * Ly));
        const int sx = (ix - source[0] + Lx) % Lx, sy = (iy - source[1] + Ly) % Ly, sz = (iz - source[2] + Lz) % Lz;
        pt_dst[x] = ((size_t)sz * Ly + sy) * Lx + sx;
      }
    }
    const size_t block = (size_t)totalT * npt * sr * 2;  // one dataset, in Floats
    std::vector<Float> buf(block);
    char top[64];
    std::snprintf(top, sizeof(top), "/sx%02dsy%02dsz%02dst%02d/", source[0], source[1], source[2], source[3]);
    std::vector<int32_t> mvec;
    std::vector<uint64_t> mdims;
    if (mom) {
      for (const auto &m : corr_mom_space->MomList())
        for (auto v : m) mvec.push_back((int32_t)std::lround(v));
      // a list of one momentum stays a flat vector of 3 (PLEGMA_hdf5.h:683-691)
      mdims = npt == 1 ? std::vector<uint64_t>{3} : std::vector<uint64_t>{(uint64_t)npt, 3};
    }
    // the HDF5 file name as include/io/PLEGMA_hdf5.h:541-560 forms it ("x.h5",
    // "x.h5/<group path>" -> x.h5, anything else -> "<name>.h5"; the group path
    // is unused as the correlator's paths are absolute); the record file is
    // that name + ".twin"
    std::string h5name = filename;
    const size_t ext = filename.rfind(".h5");
    if (ext == std::string::npos || ext + 3 < filename.size())
      h5name = ext != std::string::npos && filename[ext + 3] == '/' ? filename.substr(0, ext + 3) : filename + ".h5";
    try {
      H5Twin w(h5name + ".twin");
      const Float *H = H_elem();
      for (size_t g = 0; g < nGroups(); g++) {
        const std::string gpath = std::string(top) + (groups.empty() ? "" : groups[g]) + "/";
        if (mom) w.write_dataset(gpath + "mvec", mvec.data(), mdims);
        for (size_t d = 0; d < nDatasets(); d++) {
          const Float *src = H + (g * nDatasets() + d) * (size_t)lt * npt * sr * 2;
          if ((size_t)lt != (size_t)totalT) std::fill(buf.begin(), buf.end(), Float(0));
          for (int j = 0; j < lt; j++)
            for (size_t x = 0; x < npt; x++)
              std::memcpy(&buf[((row_dst[j] * npt) + pt_dst[x]) * sr * 2], src + ((size_t)j * npt + x) * sr * 2,
                          sr * 2 * sizeof(Float));
          const std::string name = datasets.empty() ? "arr" : datasets[d];
          w.write_dataset(gpath + name, buf.data(), dims, {{"description", descr}});
        }
      }
    } catch (const std::exception &e) {
      PLEGMA_error("writeHDF5: %s", e.what());
    }
  }
  // include/PLEGMA_io.h:8-34 and 53-73: the format from the file name's
  // extension (.h5/.hdf5 also before a '/'; .txt/.dat/.ascii/.raw; .lime/.bin)
  static FILE_FORMAT deduceFormat(const std::string &filename) {
    for (std::string ext : {".h5", ".hdf5"}) {
      const size_t pos = filename.rfind(ext);
      if (pos == std::string::npos) continue;
      if (pos + ext.size() == filename.size() || filename[pos + ext.size()] == '/') return HDF5_FORMAT;
    }
    auto ends = [&](const std::string &ext) {
      return filename.size() >= ext.size() && filename.compare(filename.size() - ext.size(), ext.size(), ext) == 0;
    };
    for (std::string ext : {".txt", ".dat", ".ascii", ".raw"})
      if (ends(ext)) return ASCII_FORMAT;
    for (std::string ext : {".lime", ".bin"})
      if (ends(ext)) return LIME_FORMAT;
    std::fprintf(stderr, "PLEGMA WARNING: Deduce type of %s failed.\n", filename.c_str());
    return DEFAULT_FORMAT;
  }
  void writeFile(std::string filename, FILE_FORMAT format) const {
    switch (format) {
    case ASCII_FORMAT: writeASCII(filename); break;
    case HDF5_FORMAT: writeHDF5(filename); break;
    case LIME_FORMAT: PLEGMA_error("Writing a correlator in LIME format is not supported");
    default: PLEGMA_error("No default writing defined");
    }
  }
  void writeFile(std::string filename) const { writeFile(filename, deduceFormat(filename)); }
};

}  // namespace plegma
