// PLEGMA_amd_core.h -- shared pieces of the C++ class layer over the C ABI
// (include/plegma_amd.h): lattice set-up, error handling, enums and the
// source-site type of the reference.
//
// The classes of this directory (PLEGMA_Propagator, PLEGMA_FT,
// PLEGMA_Correlator) keep the names, constructor arguments, accessors and
// H_elem layouts of the reference's classes (include/PLEGMA_Correlator.h,
// include/PLEGMA_FT.h, include/PLEGMA_Field.h) for the two-point path, for one
// process on one GPU.  Host C++17; device memory through the HIP runtime API.
#pragma once

#ifndef __HIP_PLATFORM_AMD__
#define __HIP_PLATFORM_AMD__ 1
#endif
#include <hip/hip_runtime_api.h>

#include <array>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "../plegma_amd.h"

namespace plegma {

// include/global/PLEGMA_enums.h:7-8
enum CORR_SPACE { POSITION_SPACE, MOMENTUM_SPACE };
enum FILE_FORMAT { ASCII_FORMAT, HDF5_FORMAT, LIME_FORMAT, DEFAULT_FORMAT };
// include/global/PLEGMA_enums.h:10
enum ALLOCATION_FLAG { EVERY = -1, NONE = 0, HOST, DEVICE, BOTH };

// include/global/PLEGMA_structs.h:4-12
struct site : std::array<int, 4> {
  site() = default;
  site(const std::array<int, 4> &v) : std::array<int, 4>(v) {}
  site(int x, int y, int z, int t) : std::array<int, 4>{{x, y, z, t}} {}
};

// include/global/PLEGMA_prints.hpp:49-81: print and abort (one process here)
[[noreturn]] inline void PLEGMA_error(const char *fmt, ...) {
  std::va_list ap;
  va_start(ap, fmt);
  std::fprintf(stderr, "PLEGMA ERROR: ");
  std::vfprintf(stderr, fmt, ap);
  std::fprintf(stderr, "\n");
  va_end(ap);
  std::fflush(stderr);
  std::abort();
}

// a failing C-ABI call is a PLEGMA_error with the library's message
inline void plg_check(int rc, const char *what) {
  if (rc != PLG_OK) PLEGMA_error("%s failed (%d): %s", what, rc, plg_last_error());
}
inline void hip_check(hipError_t e, const char *what) {
  if (e != hipSuccess) PLEGMA_error("%s failed: %s", what, hipGetErrorString(e));
}

// lattice of the process (replaces the HGC_totalL globals that PLEGMA_init sets)
struct Lattice {
  int L[4] = {0, 0, 0, 0};
  bool init = false;
  static Lattice &get() {
    static Lattice l;
    return l;
  }
  static long long V3() { return (long long)get().L[0] * get().L[1] * get().L[2]; }
  static long long V() { return V3() * get().L[3]; }
};

// one process, one GPU: L = {Lx, Ly, Lz, Lt}
inline void PLEGMA_amd_init(const std::array<int, 4> &L, int device = 0) {
  plg_check(plg_init(L.data(), device), "plg_init");
  for (int i = 0; i < 4; i++) Lattice::get().L[i] = L[i];
  Lattice::get().init = true;
}
inline void PLEGMA_amd_finalize() {
  plg_check(plg_finalize(), "plg_finalize");
  Lattice::get().init = false;
}

template <typename Float> constexpr int plg_precision_of() { return sizeof(Float) == 8 ? PLG_F64 : PLG_F32; }

}  // namespace plegma
