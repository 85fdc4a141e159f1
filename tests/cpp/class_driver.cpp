// class_driver.cpp -- drives the C++ class layer (include/plegma/) the way a
// PLEGMA driver does (plegma/Calc_2pt.cpp:121-132, 160-266): one lattice,
// seeded propagators, a correlator per contraction, H_elem, the ASCII file and
// the HDF5 writer's record file.  Writes each H_elem as raw little-endian Float
// pairs to <outdir>/<name>.bin for the GPU tests to compare with the CPU oracle.
//
//   class_driver <outdir> <f64|f32>
#include <cstdio>
#include <string>

#include "plegma/PLEGMA_Correlator.h"

using namespace plegma;

template <typename Float> static void dump(const PLEGMA_Correlator<Float> &c, const std::string &path) {
  FILE *f = std::fopen(path.c_str(), "wb");
  if (!f) PLEGMA_error("cannot write %s", path.c_str());
  std::fwrite(c.H_elem(), sizeof(Float), 2 * c.getTotalSize(), f);
  std::fclose(f);
  std::printf("%s %zu\n", path.c_str(), c.getTotalSize());
}

template <typename Float> static int run(const std::string &out) {
  const std::string tag = sizeof(Float) == 8 ? "f64" : "f32";
  PLEGMA_Propagator<Float> u, d;
  if (sizeof(Float) == 8) {
    u.synthesize(1234, 1, 0, 0.5);
    d.synthesize(1234, 2, 0, 0.5);
  } else {
    PLEGMA_Propagator<double> u64, d64;
    u64.synthesize(1234, 1, 0, 0.5);
    d64.synthesize(1234, 2, 0, 0.5);
    u.castFrom(u64);
    d.castFrom(d64);
  }
  site src(1, 2, 3, 5);
  {
    PLEGMA_Correlator<Float> c(MOMENTUM_SPACE, src, 4, 6);
    c.contractMesons(u, d);
    dump(c, out + "/mesons_" + tag + ".bin");
    c.writeASCII(out + "/mesons_" + tag);
    c.writeHDF5(out + "/corr_" + tag);  // -> corr_<tag>.h5.twin
  }
  {
    PLEGMA_Correlator<Float> c(MOMENTUM_SPACE, src, 4, 6);
    c.contractBaryons(u, d, PLG_G_NUCL | PLG_G_DELTAPP_ALL);
    dump(c, out + "/baryons_" + tag + ".bin");
    c.writeFile(out + "/corr_" + tag + ".h5");  // format from the extension
  }
  {
    PLEGMA_Correlator<Float> c(MOMENTUM_SPACE, src, 1);  // full time extent
    c.contractMesonsAll(u, d);
    dump(c, out + "/mesons_all_" + tag + ".bin");
    c.writeASCII(out + "/mesons_all_" + tag);
    // same group as contractMesons: its mvec stays, twop_meson is added
    c.writeFile(out + "/corr_" + tag + ".h5/ignored/group", HDF5_FORMAT);
  }
  {
    // a second write of a dataset already in the file is skipped
    PLEGMA_Correlator<Float> c(MOMENTUM_SPACE, src, 4, 6);
    c.contractMesons(d, u);
    c.writeHDF5(out + "/corr_" + tag + ".h5");
  }
  {
    // position space over a window that wraps the time boundary
    PLEGMA_Correlator<Float> c(POSITION_SPACE, src, 0, 4);
    c.contractMesonsNew(u, d);
    dump(c, out + "/pos_mesons_" + tag + ".bin");
    c.writeHDF5(out + "/pos_" + tag + ".h5");
  }
  {
    PLEGMA_Correlator<Float> c(MOMENTUM_SPACE, src, 4, 6);
    c.contractMesonMeson(u, d);
    dump(c, out + "/mesmes_" + tag + ".bin");
  }
  {
    // udsc baryons with the charm propagator absent: the products of u, d, s
    PLEGMA_Propagator<Float> s, none(NONE);
    if (sizeof(Float) == 8) {
      s.synthesize(1234, 3, 0, 0.8);
    } else {
      PLEGMA_Propagator<double> s64;
      s64.synthesize(1234, 3, 0, 0.8);
      s.castFrom(s64);
    }
    PLEGMA_Correlator<Float> c(MOMENTUM_SPACE, src, 2, 6);
    c.contractBaryonsUDSC(u, d, s, none);
    dump(c, out + "/udsc_" + tag + ".bin");
    c.writeHDF5(out + "/udsc_" + tag + ".h5");
    FILE *f = std::fopen((out + "/udsc_" + tag + ".names").c_str(), "w");
    for (auto &n : c.getDatasets()) std::fprintf(f, "%s\n", n.c_str());
    std::fclose(f);
  }
  {
    // heavy-light tetraquarks with the charm propagator absent and the light
    // one as LT: products "udbb" and "usbb"; u (seed stream 1) as LT, the
    // strange stream 3 as ST, a heavy bottom (stream 4, decay 1.2) as BT
    PLEGMA_Propagator<Float> s, b, none(NONE);
    if (sizeof(Float) == 8) {
      s.synthesize(1234, 3, 0, 0.8);
      b.synthesize(1234, 4, 0, 1.2);
    } else {
      PLEGMA_Propagator<double> s64, b64;
      s64.synthesize(1234, 3, 0, 0.8);
      b64.synthesize(1234, 4, 0, 1.2);
      s.castFrom(s64);
      b.castFrom(b64);
    }
    PLEGMA_Correlator<double> c(MOMENTUM_SPACE, src, 2, 6);
    c.contractTetraquarks(u, s, none, b);
    dump(c, out + "/tetra_" + tag + ".bin");
    FILE *f = std::fopen((out + "/tetra_" + tag + ".names").c_str(), "w");
    for (auto &n : c.getDatasets()) std::fprintf(f, "%s\n", n.c_str());
    std::fclose(f);
    // bcud tetraquarks: u as LT, a charm propagator (stream 5, decay 1.0)
    // as CH, the bottom as BT, the strange one absent
    PLEGMA_Propagator<Float> ch;
    if (sizeof(Float) == 8) {
      ch.synthesize(1234, 5, 0, 1.0);
    } else {
      PLEGMA_Propagator<double> c64;
      c64.synthesize(1234, 5, 0, 1.0);
      ch.castFrom(c64);
    }
    PLEGMA_Correlator<double> cb(MOMENTUM_SPACE, src, 2, 6);
    cb.contractTetraquarksBCUD(u, none, ch, b);
    dump(cb, out + "/bcud_" + tag + ".bin");
    f = std::fopen((out + "/bcud_" + tag + ".names").c_str(), "w");
    for (auto &n : cb.getDatasets()) std::fprintf(f, "%s\n", n.c_str());
    std::fclose(f);
  }
  {
    // a propagator as a 144-component field through PLEGMA_FT::apply
    PLEGMA_FT<Float> ft(2);
    ft.apply(u, FT_GEMV, -1);
    const std::string path = out + "/ft_" + tag + ".bin";
    FILE *f = std::fopen(path.c_str(), "wb");
    std::fwrite(ft.H_elem(), sizeof(Float), (size_t)2 * ft.Nmoms() * ft.DimT() * ft.Dof(), f);
    std::fclose(f);
  }
  return 0;
}

int main(int argc, char **argv) {
  if (argc != 3) {
    std::fprintf(stderr, "usage: %s <outdir> <f64|f32>\n", argv[0]);
    return 2;
  }
  PLEGMA_amd_init({4, 4, 4, 8}, 0);
  const int rc = std::string(argv[2]) == "f32" ? run<float>(argv[1]) : run<double>(argv[1]);
  PLEGMA_amd_finalize();
  return rc;
}
