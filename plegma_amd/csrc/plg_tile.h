// plg_tile.h -- the site tile shared by the per-site contraction kernels
// (baryons, fused two-point, meson-meson): 16 lattice sites per 256-thread
// workgroup, 16 lanes (one DPP row of a wavefront) per site, propagator
// components staged in LDS as [component][site].
#pragma once
#include "plg_internal.h"

namespace plg {

constexpr int BT_SITES = 16;  // sites per workgroup tile
constexpr int BT_THREADS = 256;
// LDS strides chosen for the access patterns of a wavefront (4 sites x 16
// lanes, 8-byte complex = 2 banks): a component row of 17 sites makes the 16
// lanes of a site that read components 9 l + i fall on 16 distinct bank
// pairs; a per-site exchange stride of 292 (= 8 banks mod 32) keeps the 4
// sites x 4 columns of transposed reads on distinct banks.
constexpr int BT_SP = BT_SITES + 1;
constexpr int BT_WS = 2 * 16 * 9 + 4;

// per meson channel op (GAMMAS order): xor class and phases of X = g5 G
struct MesonX {
  int cls[16];
  int ph[16][4];
};
constexpr MesonX make_mesonx() {
  MesonX m{};
  for (int op = 0; op < 16; op++) {
    SPerm X = sp_mul(G_5, gamma_op(op));
    m.cls[op] = sp_class(X);
    for (int i = 0; i < 4; i++) m.ph[op][i] = X.ph[i];
  }
  return m;
}

// ---- DPP row reductions (16-lane rows = the lanes of one site)
template <int CTRL> __device__ __forceinline__ float dpp(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
template <int CTRL> __device__ __forceinline__ double dpp(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, 0xF, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
template <typename T> __device__ __forceinline__ T rowsum16(T v) {
  v += dpp<0xB1>(v);   // quad_perm [1,0,3,2]
  v += dpp<0x4E>(v);   // quad_perm [2,3,0,1]
  v += dpp<0x141>(v);  // row_half_mirror: the other quad of the half row
  v += dpp<0x140>(v);  // row_mirror: the other half row
  return v;
}

// The 16 lanes of a site are one DPP row of one wavefront, so data a site's
// lanes exchange through LDS needs ordering at wavefront scope only (LDS
// operations of a wavefront complete in issue order).
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// runtime phase i^ph (lane-dependent)
template <typename T> __device__ __forceinline__ cx<T> cmul_iph_rt(cx<T> a, int ph) {
  ph &= 3;
  const T re = (ph & 1) ? -a.im : a.re;
  const T im = (ph & 1) ? a.re : a.im;
  return (ph & 2) ? cmake<T>(-re, -im) : cmake<T>(re, im);
}

// rotation to the physical basis of a staged tile S[component][site] in
// place (all threads of the workgroup; the caller synchronises after it)
template <typename T> __device__ __forceinline__ void rot_tile(cx<T> (*S)[BT_SP], int sign) {
  for (int it = threadIdx.x; it < BT_SITES * 36; it += BT_THREADS) {
    const int s = it % BT_SITES, cq = it / BT_SITES, cp = cq % 9, q = cq / 9, m = q >> 1, n = q & 1;
    cx<T> a = S[(m * 4 + n) * 9 + cp][s], b = S[(m * 4 + n + 2) * 9 + cp][s];
    cx<T> c = S[((m + 2) * 4 + n) * 9 + cp][s], d = S[((m + 2) * 4 + n + 2) * 9 + cp][s];
    rot_physical(a, b, c, d, sign);
    S[(m * 4 + n) * 9 + cp][s] = a;
    S[(m * 4 + n + 2) * 9 + cp][s] = b;
    S[((m + 2) * 4 + n) * 9 + cp][s] = c;
    S[((m + 2) * 4 + n + 2) * 9 + cp][s] = d;
  }
}

// entry i of a compile-time 4-vector at a runtime index (selects, no scratch)
__device__ __forceinline__ int pick4(const int (&v)[4], int i) {
  int r = 0;
#pragma unroll
  for (int k = 0; k < 4; k++) r += i == k ? v[k] : 0;
  return r;
}

}  // namespace plg
