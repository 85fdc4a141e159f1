// PLEGMA_Propagator.h -- device propagator of the two-point path: 144
// complex components (mu, nu, a, b) per lattice site in the reference's SoA
// layout, element (i*V + site) complex, i = ((mu*4 + nu)*3 + a)*3 + b
// (lib/PLEGMA_Propagator.cu:39-44, memory of lib/PLEGMA_Field.cu:225-264).
// Keeps the accessors the Correlator path reads (include/PLEGMA_Field.h:124-159).
#pragma once
#include "PLEGMA_amd_core.h"

namespace plegma {

template <typename Float> class PLEGMA_Propagator {
  Float *d_elem = nullptr;
  int *d_plane_exp = nullptr;  // FP32 only: per-time-slice power-of-two exponents
  ALLOCATION_FLAG allocation = NONE;
  size_t total_length = 0;

public:
  // alloc_flag NONE: "flavour absent" (PLEGMA_Correlator.cu:151-154); otherwise
  // device memory (a host copy is not kept: the path never reads one)
  explicit PLEGMA_Propagator(ALLOCATION_FLAG alloc_flag = BOTH) {
    if (!Lattice::get().init) PLEGMA_error("PLEGMA_amd_init must be called before creating fields");
    total_length = (size_t)Lattice::V();
    if (alloc_flag == NONE) return;
    hip_check(hipMalloc((void **)&d_elem, Bytes_total()), "propagator allocation");
    allocation = DEVICE;
  }
  PLEGMA_Propagator(const PLEGMA_Propagator &) = delete;
  PLEGMA_Propagator &operator=(const PLEGMA_Propagator &) = delete;
  ~PLEGMA_Propagator() {
    if (d_elem) (void)hipFree(d_elem);
    if (d_plane_exp) (void)hipFree(d_plane_exp);
  }

  Float *D_elem() const { return d_elem; }
  ALLOCATION_FLAG getAllocation() const { return allocation; }
  int Field_length() const { return 144; }
  size_t Total_length() const { return total_length; }
  size_t Bytes_total() const { return Total_length() * Field_length() * 2 * sizeof(Float); }

  // view of the whole propagator for the C ABI
  plg_prop view() const {
    plg_prop v{};
    v.data = d_elem;
    v.precision = plg_precision_of<Float>();
    v.t_first = 0;
    v.nt = Lattice::get().L[3];
    v.comp_stride = (int64_t)total_length;
    v.plane_exp = d_plane_exp;
    return v;
  }

  // ---- preparation in place (include/PLEGMA_Propagator.h:57-59)
  // entries at global time t < t0 negated (lib/PLEGMA_Propagator.cu:98-100)
  void applyBoundaries_device(int t0) {
    plg_prop v = view();
    plg_check(plg_prop_apply_boundaries(&v, t0, nullptr), "plg_prop_apply_boundaries");
  }
  // twisted-mass rotation to the physical basis, sign +-1 (lib/PLEGMA_Propagator.cu:103-106)
  void rotateToPhysicalBase_device(int sign) {
    if (sign != +1 && sign != -1) PLEGMA_error("The sign can be only +-1\n");
    plg_prop v = view();
    plg_check(plg_prop_rotate_physical(&v, sign, nullptr), "plg_prop_rotate_physical");
    hip_check(hipDeviceSynchronize(), "rotateToPhysicalBase_device");
  }

  // ---- inputs of benchmarks and tests (no reference counterpart)
  // seeded N(0,1) entries times exp(-lambda |t - t0|); FP32 propagators carry
  // per-time-slice exponents so that fast decays stay in range
  void synthesize(uint64_t seed, uint32_t stream_id, int t0, double lambda) {
    ensure_plane_exp();
    plg_prop v = view();
    plg_check(plg_synth_prop(&v, seed, stream_id, t0, lambda, nullptr), "plg_synth_prop");
    hip_check(hipDeviceSynchronize(), "synthesize");
  }
  void synthesize_perturbed(uint64_t seed, uint32_t base_stream, uint32_t pert_stream, double eps, int t0,
                            double lambda) {
    ensure_plane_exp();
    plg_prop v = view();
    plg_check(plg_synth_prop_perturbed(&v, seed, base_stream, pert_stream, eps, t0, lambda, nullptr),
              "plg_synth_prop_perturbed");
    hip_check(hipDeviceSynchronize(), "synthesize_perturbed");
  }
  // round-to-nearest copy of an FP64 propagator (with per-slice exponents)
  template <typename F2> void castFrom(const PLEGMA_Propagator<F2> &src) {
    ensure_plane_exp();
    plg_prop d = view(), s = src.view();
    plg_check(plg_cast_prop(&d, &s, nullptr), "plg_cast_prop");
    hip_check(hipDeviceSynchronize(), "castFrom");
  }

private:
  void ensure_plane_exp() {
    if (allocation == NONE) PLEGMA_error("propagator is not allocated");
    if (sizeof(Float) == 4 && !d_plane_exp)
      hip_check(hipMalloc((void **)&d_plane_exp, Lattice::get().L[3] * sizeof(int)), "plane exponents");
  }
};

}  // namespace plegma
