// plg_common.h -- shared device/host definitions of the 2pt path.
//
// Spin algebra: every matrix of the 16-element Clifford basis used by PLEGMA
// (lib/kernels/PLEGMA_gammas.cuh:6-41, GAMMAS order of
// include/global/PLEGMA_enums.h:22) is a signed permutation: row i has one
// non-zero, at column i ^ c (c in {0,1,2,3}), with value i^ph[i].  The
// tables below are derived at compile time from gamma_1..gamma_4.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace plg {

// ------------------------------------------------------------------ spin algebra
struct SPerm {
  int col[4];  // column of the non-zero in row i
  int ph[4];   // value = i^ph (0:1, 1:i, 2:-1, 3:-i)
};

constexpr SPerm sp_mul(const SPerm &a, const SPerm &b) {
  SPerm r{};
  for (int i = 0; i < 4; i++) {
    r.col[i] = b.col[a.col[i]];
    r.ph[i] = (a.ph[i] + b.ph[a.col[i]]) & 3;
  }
  return r;
}
constexpr SPerm sp_phase(const SPerm &a, int ph) {
  SPerm r = a;
  for (int i = 0; i < 4; i++) r.ph[i] = (a.ph[i] + ph) & 3;
  return r;
}
// adjoint: (A^dag)[j][i] = conj(A[i][j])
constexpr SPerm sp_dag(const SPerm &a) {
  SPerm r{};
  for (int i = 0; i < 4; i++) {
    r.col[a.col[i]] = i;
    r.ph[a.col[i]] = (4 - a.ph[i]) & 3;
  }
  return r;
}

constexpr SPerm G_ONE{{0, 1, 2, 3}, {0, 0, 0, 0}};
constexpr SPerm G_1{{3, 2, 1, 0}, {1, 1, 3, 3}};
constexpr SPerm G_2{{3, 2, 1, 0}, {0, 2, 2, 0}};
constexpr SPerm G_3{{2, 3, 0, 1}, {1, 3, 3, 1}};
constexpr SPerm G_4{{0, 1, 2, 3}, {0, 0, 2, 2}};
constexpr SPerm G_5 = sp_mul(sp_mul(sp_mul(G_1, G_2), G_3), G_4);
// sigma_{mu nu} = -(i/2)[g_mu, g_nu] = -i g_mu g_nu for mu != nu
constexpr SPerm sp_sigma(const SPerm &a, const SPerm &b) { return sp_phase(sp_mul(a, b), 3); }

constexpr SPerm gamma_op(int op) {
  switch (op) {
  case 0: return G_ONE;
  case 1: return G_1;
  case 2: return G_2;
  case 3: return G_3;
  case 4: return G_4;
  case 5: return G_5;
  case 6: return sp_mul(G_5, G_1);
  case 7: return sp_mul(G_5, G_2);
  case 8: return sp_mul(G_5, G_3);
  case 9: return sp_mul(G_5, G_4);
  case 10: return sp_sigma(G_1, G_2);
  case 11: return sp_sigma(G_1, G_3);
  case 12: return sp_sigma(G_2, G_3);
  case 13: return sp_sigma(G_4, G_1);
  case 14: return sp_sigma(G_4, G_2);
  default: return sp_sigma(G_4, G_3);
  }
}
// charge conjugation C = i g2 g4
constexpr SPerm G_C = sp_phase(sp_mul(G_2, G_4), 1);
// diquark matrix A = C Gamma (0: C g5, k=1..3: C g_k) and Abar = g4 A^dag g4
constexpr SPerm diquark(int which) { return sp_mul(G_C, gamma_op(which == 0 ? 5 : which)); }
constexpr SPerm diquark_bar(int which) { return sp_mul(sp_mul(G_4, sp_dag(diquark(which))), G_4); }

// xor class of a signed permutation
constexpr int sp_class(const SPerm &a) { return a.col[0]; }

// channel order of contractMesons (lib/PLEGMA_Correlator.cu:47-50)
constexpr int MESON10_OPS[10] = {5, 0, 6, 7, 8, 9, 1, 2, 3, 4};

// ------------------------------------------------------------------ complex
template <typename T> struct cx {
  T re, im;
};
template <typename T> __host__ __device__ inline cx<T> cmake(T r, T i) { return cx<T>{r, i}; }

// Twisted-mass rotation to the physical basis of the four spin entries
// a = S[m][n], b = S[m][n+2], c = S[m+2][n], d = S[m+2][n+2] of one colour
// pair (spin indices mod 4), in place:
//   S'[mu][nu] = (S[mu][nu] - S[mu+2][nu+2] + i s (S[mu][nu+2] + S[mu+2][nu])) / 2
// (lib/kernels/PLEGMA_propagator_utils.cuh:103-118, same evaluation order).
template <typename T>
__host__ __device__ __forceinline__ void rot_physical(cx<T> &a, cx<T> &b, cx<T> &c, cx<T> &d, int s) {
  const T sg = (T)s, h = (T)0.5;
  const cx<T> A = a, B = b, C = c, D = d;
  auto term = [&](cx<T> p, cx<T> q, cx<T> r) {  // (p - q + i s r) / 2
    return cmake<T>(h * ((p.re - q.re) + (-sg * r.im)), h * ((p.im - q.im) + sg * r.re));
  };
  a = term(A, D, B + C);
  d = term(D, A, C + B);
  b = term(B, C, A + D);
  c = term(C, B, D + A);
}

// rotation to the physical basis of the 16 spin entries s[mu * 4 + nu] of one
// colour pair, in registers
template <typename T> __device__ __forceinline__ void rot_spins(cx<T> (&s)[16], int sign) {
#pragma unroll
  for (int m = 0; m < 2; m++)
#pragma unroll
    for (int n = 0; n < 2; n++) rot_physical(s[m * 4 + n], s[m * 4 + n + 2], s[(m + 2) * 4 + n], s[(m + 2) * 4 + n + 2], sign);
}

// Loads and stores of HBM data through global (address space 1) pointers.
// Through a generic pointer the compiler emits flat instructions, which also
// count in the LDS wait counter: every LDS wait would then wait for the
// prefetch of the next tile as well.
template <typename T> __device__ __forceinline__ cx<T> gload(const cx<T> *p) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  const v2 v = *(const __attribute__((address_space(1))) v2 *)p;
  return cmake<T>(v[0], v[1]);
}
template <typename T> __device__ __forceinline__ void gstore(cx<T> *p, const cx<T> &v) {
  typedef T v2 __attribute__((ext_vector_type(2)));
  *(__attribute__((address_space(1))) v2 *)p = v2{v.re, v.im};
}
template <typename T> __host__ __device__ inline cx<T> operator+(cx<T> a, cx<T> b) {
  return {a.re + b.re, a.im + b.im};
}
template <typename T> __host__ __device__ inline cx<T> operator-(cx<T> a, cx<T> b) {
  return {a.re - b.re, a.im - b.im};
}
template <typename T> __host__ __device__ inline cx<T> operator*(cx<T> a, cx<T> b) {
  return {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
template <typename T> __host__ __device__ inline cx<T> operator*(T s, cx<T> a) {
  return {s * a.re, s * a.im};
}
template <typename T> __host__ __device__ inline cx<T> cconj(cx<T> a) { return {a.re, -a.im}; }
// a += b * c
template <typename T> __host__ __device__ inline void cfma(cx<T> &a, cx<T> b, cx<T> c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(-b.im, c.im, a.re);
  a.im = fma(b.re, c.im, a.im);
  a.im = fma(b.im, c.re, a.im);
}
// a += b * conj(c)
template <typename T> __host__ __device__ inline void cfma_conj(cx<T> &a, cx<T> b, cx<T> c) {
  a.re = fma(b.re, c.re, a.re);
  a.re = fma(b.im, c.im, a.re);
  a.im = fma(b.im, c.re, a.im);
  a.im = fma(-b.re, c.im, a.im);
}
// bf16 bit pattern of a finite float, rounded to nearest even, and back
__device__ __forceinline__ uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf16_val(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// a += i^PH * b * c as four FMAs (the phase folds into operand signs)
template <int PH, typename T> __host__ __device__ inline void cfma_iph(cx<T> &a, cx<T> b, cx<T> c) {
  constexpr int p = PH & 3;
  if constexpr (p == 0) {
    a.re = fma(b.re, c.re, a.re);
    a.re = fma(-b.im, c.im, a.re);
    a.im = fma(b.re, c.im, a.im);
    a.im = fma(b.im, c.re, a.im);
  } else if constexpr (p == 1) {  // i (x + iy) = -y + ix
    a.re = fma(-b.re, c.im, a.re);
    a.re = fma(-b.im, c.re, a.re);
    a.im = fma(b.re, c.re, a.im);
    a.im = fma(-b.im, c.im, a.im);
  } else if constexpr (p == 2) {
    a.re = fma(-b.re, c.re, a.re);
    a.re = fma(b.im, c.im, a.re);
    a.im = fma(-b.re, c.im, a.im);
    a.im = fma(-b.im, c.re, a.im);
  } else {  // -i (x + iy) = y - ix
    a.re = fma(b.re, c.im, a.re);
    a.re = fma(b.im, c.re, a.re);
    a.im = fma(-b.re, c.re, a.im);
    a.im = fma(b.im, c.im, a.im);
  }
}
// a += i^PH * b * conj(c) as four FMAs
template <int PH, typename T> __host__ __device__ inline void cfma_conj_iph(cx<T> &a, cx<T> b, cx<T> c) {
  // b conj(c) = (br cr + bi ci) + i (bi cr - br ci)
  constexpr int p = PH & 3;
  if constexpr (p == 0) {
    a.re = fma(b.re, c.re, a.re);
    a.re = fma(b.im, c.im, a.re);
    a.im = fma(b.im, c.re, a.im);
    a.im = fma(-b.re, c.im, a.im);
  } else if constexpr (p == 1) {
    a.re = fma(b.re, c.im, a.re);
    a.re = fma(-b.im, c.re, a.re);
    a.im = fma(b.re, c.re, a.im);
    a.im = fma(b.im, c.im, a.im);
  } else if constexpr (p == 2) {
    a.re = fma(-b.re, c.re, a.re);
    a.re = fma(-b.im, c.im, a.re);
    a.im = fma(-b.im, c.re, a.im);
    a.im = fma(b.re, c.im, a.im);
  } else {
    a.re = fma(b.im, c.re, a.re);
    a.re = fma(-b.re, c.im, a.re);
    a.im = fma(-b.re, c.re, a.im);
    a.im = fma(-b.im, c.im, a.im);
  }
}
// multiply by i^ph (compile-time ph folds to swaps/negations)
template <typename T> __host__ __device__ inline cx<T> cmul_iph(cx<T> a, int ph) {
  switch (ph & 3) {
  case 0: return a;
  case 1: return {-a.im, a.re};
  case 2: return {-a.re, -a.im};
  default: return {a.im, -a.re};
  }
}

// ------------------------------------------------------------------ epsilon
// eps permutations in the order of lib/kernels/PLEGMA_kernel_utils.cuh:49-56
constexpr int EPS[6][3] = {{0, 1, 2}, {2, 0, 1}, {1, 2, 0}, {2, 1, 0}, {0, 2, 1}, {1, 0, 2}};
constexpr int SGN_EPS[6] = {+1, +1, +1, -1, -1, -1};

// propagator component index: ((mu*4 + nu)*3 + a)*3 + b
__host__ __device__ constexpr int pidx(int mu, int nu, int a, int b) {
  return ((mu * 4 + nu) * 3 + a) * 3 + b;
}

// compile-time loop: f(std::integral_constant<int, I>) for I = 0..N-1, so that
// indices into register arrays are constants in every instance
template <int I> struct ic {
  static constexpr int value = I;
  constexpr operator int() const { return I; }
};
template <int B, int E, typename F> __host__ __device__ __forceinline__ void static_for(F &&f) {
  if constexpr (B < E) {
    f(ic<B>{});
    static_for<B + 1, E>(f);
  }
}

}  // namespace plg
