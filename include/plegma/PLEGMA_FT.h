// PLEGMA_FT.h -- momentum list and host result buffer of the momentum
// projection, with the interface of include/PLEGMA_FT.h:11-133 (3D lists).
// H_elem holds Nmoms * DimT * dof complex Float; the correlator path writes its
// [group][dataset][row][momentum][channel] block into it.
#pragma once
#include <cmath>
#include <complex>
#include <cstring>
#include <memory>

#include "PLEGMA_amd_core.h"

namespace plegma {

// include/PLEGMA_FT.h:9
enum FT_TYPE { FT_NAIVE, FT_GEMV, FT_FFT };

template <typename Float> class PLEGMA_FT {
public:
  typedef std::vector<Float> VFloat;
  typedef std::vector<VFloat> VVFloat;

private:
  int Q2_max = 0;
  int dof = 0;
  std::shared_ptr<Float> h_elem;
  size_t sizeN = 0;
  int dims = 3;
  int dimT = 1;
  bool accum = false;
  VVFloat momList;

  void check_dims(int d) {
    if (d != 3 && d != 4) PLEGMA_error("This class transforms only 3 and 4 dimensions");
    if (d == 4) PLEGMA_error("4D momentum lists are not provided by the MI355X two-point path");
  }

public:
  // lib/PLEGMA_FT.cu:14-21, list of lib/PLEGMA_FT.cu:50-72
  PLEGMA_FT(int Q2_max, int D3D4 = 3, bool accum = false, int dimT = Lattice::get().L[3])
      : Q2_max(Q2_max), dims(D3D4), dimT(dimT), accum(accum) {
    check_dims(D3D4);
    if (Q2_max < 0) PLEGMA_error("The maximum number of Q2 cannot be negative");
    if (dimT < 0 || dimT > Lattice::get().L[3])
      PLEGMA_error("The time dimension cannot be negative or larger than local size");
    int n = 0;
    plg_check(plg_momlist(Q2_max, &n, nullptr, 0), "plg_momlist");
    std::vector<int> xyz(3 * (size_t)n);
    plg_check(plg_momlist(Q2_max, &n, xyz.data(), n), "plg_momlist");
    for (int i = 0; i < n; i++) momList.push_back({(Float)xyz[3 * i], (Float)xyz[3 * i + 1], (Float)xyz[3 * i + 2]});
  }
  // lib/PLEGMA_FT.cu:23-33
  template <typename T>
  PLEGMA_FT(std::vector<T> mom, int D3D4 = 3, bool accum = false, int dimT = Lattice::get().L[3])
      : dims(D3D4), dimT(dimT), accum(accum) {
    check_dims(D3D4);
    if ((int)mom.size() != dims) PLEGMA_error("The size of the momentum vector does not match the dimensionality of FT");
    momList.push_back(VFloat(mom.begin(), mom.end()));
  }
  // lib/PLEGMA_FT.cu:35-45
  template <typename T>
  PLEGMA_FT(std::vector<std::vector<T>> &moms, int D3D4 = 3, bool accum = false, int dimT = Lattice::get().L[3])
      : dims(D3D4), dimT(dimT), accum(accum) {
    check_dims(D3D4);
    for (auto &m : moms) {
      if (m.size() % dims != 0) PLEGMA_error("The size of the momentum vector does not match the dimensionality of FT");
      momList.push_back(VFloat(m.begin(), m.end()));
    }
  }

  int Nmoms() const { return (int)momList.size(); }
  VVFloat MomList() const { return momList; }
  int Dims() const { return dims; }
  int DimT() const { return dimT; }
  bool IsAccum() const { return accum; }
  Float *H_elem() const { return h_elem.get(); }
  int Dof() const { return dof; }

  // integer momenta of the list (the C ABI takes integer triples)
  std::vector<int> intMomenta() const {
    std::vector<int> v;
    for (auto &m : momList)
      for (int i = 0; i < 3; i++) v.push_back((int)std::lround(m[i]));
    return v;
  }

  // lib/PLEGMA_FT.cu:74-80
  void checkAllocation(int newDof) {
    dof = newDof;
    sizeN = (size_t)Nmoms() * dimT * dof * 2;
    h_elem.reset(new Float[sizeN], std::default_delete<Float[]>());
    zero();
  }
  void zero() {
    if (h_elem) std::memset(h_elem.get(), 0, sizeN * sizeof(Float));
  }
  void scale(Float a) {
    for (size_t i = 0; i < sizeN; i++) h_elem.get()[i] *= a;
  }
  // lib/PLEGMA_FT.cu:144-162: H_elem[t][dof][mom] = sum_x f(x, t) exp(sign i 2pi n.x/L)
  // (FT_NAIVE and FT_GEMV are the same sum; the buffer is re-allocated and
  // overwritten, as checkAllocation does in the reference).  Field: any type
  // with D_elem(), Field_length(), Total_length() in the PLEGMA_Field layout.
  template <class Field> void apply(const Field &f, FT_TYPE type = FT_GEMV, int sign = -1) {
    if (type == FT_FFT) PLEGMA_error("Not implemented yet");
    if (sign != +1 && sign != -1) PLEGMA_error("Sign should be either +1 or -1");
    const bool is3d = f.Total_length() != (size_t)Lattice::V();
    if (is3d) dimT = 1;
    else if (dimT != Lattice::get().L[3]) PLEGMA_error("a 4D field is transformed over all T time slices");
    checkAllocation(f.Field_length());
    const size_t bytes = sizeN * sizeof(Float);
    void *d = nullptr;
    hip_check(hipMalloc(&d, bytes), "FT result buffer");
    std::vector<int> moms = intMomenta();
    const int prec = plg_precision_of<Float>();
    const int rc = plg_ft_apply(f.D_elem(), prec, f.Field_length(), dimT, moms.data(), Nmoms(), sign, 0, d, prec,
                                nullptr);
    if (rc == PLG_OK) hip_check(hipMemcpy(h_elem.get(), d, bytes, hipMemcpyDeviceToHost), "FT result copy");
    (void)hipFree(d);
    plg_check(rc, "plg_ft_apply");
  }

  // lib/PLEGMA_FT.cu:164-182 (the buffer read as [t][dof][mom])
  void mulConstMomentumPhases(std::vector<int> src, int sign) {
    if (dims == 3 && src.size() != 3) PLEGMA_error("src size is incompatible with the dimensionality of the FT");
    if (sign != +1 && sign != -1) PLEGMA_error("Sign should be either +1 or -1");
    if (!h_elem) PLEGMA_error("Apply first FT and then the const phases");
    std::complex<Float> *h2 = (std::complex<Float> *)h_elem.get();
    const Float pi = (Float)3.14159265358979323846;
    for (int imom = 0; imom < Nmoms(); imom++) {
      Float phase = 0.;
      for (int i = 0; i < dims; i++) phase += ((Float)src[i] * momList[imom][i]) / ((Float)Lattice::get().L[i]);
      phase *= (Float)2. * pi;
      const std::complex<Float> e(std::cos(phase), sign * std::sin(phase));
      for (int idf = 0; idf < dof; idf++)
        for (int it = 0; it < dimT; it++) h2[(size_t)it * dof * Nmoms() + (size_t)idf * Nmoms() + imom] *= e;
    }
  }
};

}  // namespace plegma
