// plg_mlp.hip -- the learned part: a small multilayer perceptron with BF16
// weights and activations and FP32 accumulation, trained and evaluated on
// gfx950 by the kernels of this file (no library GEMM, no autograd).
//
// The network (no counterpart in the reference) maps cheap per-source inputs
// (zero-momentum correlators) to expensive outputs (high-momentum nucleon
// correlators):  x -> [Linear -> SiLU] x (L-1) -> Linear -> y.
//
// Rounding points (a CPU restatement mirrors them exactly):
//   forward   X_0 = bf16(x);  Z_l = X_l W_l^T + b_l  (bf16 x bf16 products,
//             FP32 sums);  X_{l+1} = bf16(silu(Z_l));  y = Z_{L-1}
//   loss      mean squared error over batch x outputs, dY = 2 (y - t) / (B N)
//   backward  dW_l = bf16(dZ_l)^T X_l,  db_l = sum_b dZ_l,
//             dX_l = bf16(dZ_l) W_l,  dZ_{l-1} = dX_l * silu'(Z_{l-1})
//   update    Adam on FP32 master weights, bf16 copies rewritten each step
// Every product is a v_mfma_f32_16x16x32_bf16 tile.
#include "plg_internal.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace plg {

typedef __bf16 bf8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16x8 __attribute__((ext_vector_type(8)));

constexpr int MG_THREADS = 256;  // 4 waves, a 32 x 32 tile of C, 16 x 16 per wave

// C[m][n] (+ bias[n]) = sum_k A(m, k) B(k, n) with A(m, k) = A[m sam + k sak],
// B(k, n) = B[k sbk + n sbn] (bf16 bit patterns), FP32 accumulation.
// Optional epilogue: C (FP32, row stride ldc) and Cb = bf16(act(C)).
struct GemmArgs {
  int M, N, K;
  const uint16_t *A;
  int64_t sam, sak;
  const uint16_t *B;
  int64_t sbk, sbn;
  float *C;
  int64_t ldc;
  uint16_t *Cb;
  const float *bias;
  int silu;  // Cb = bf16(silu(C)) when set, else bf16(C)
  // split K: ks > 1 workgroup layers (blockIdx.z) sum K ranges of kper into
  // Cs[z][m][n]; k_gemm_reduce adds them and applies the epilogue
  int ks, kper;
  float *Cs;
};

__device__ __forceinline__ float silu(float z) { return z / (1.0f + expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + expf(-z));
  return s * (1.0f + z * (1.0f - s));
}

__global__ __launch_bounds__(MG_THREADS) void k_gemm_bf16(GemmArgs g) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int m0 = blockIdx.y * 32 + (wave >> 1) * 16;
  const int n0 = blockIdx.x * 32 + (wave & 1) * 16;
  // operand fragments of 16x16x32: lane l holds A[m0 + l%16][k0 + 8(l/16) + i]
  // and B[k0 + 8(l/16) + i][n0 + l%16], i < 8
  const int am = m0 + (lane & 15), bn = n0 + (lane & 15);
  const int kq = 8 * (lane >> 4);
  const int kb = blockIdx.z * g.kper, ke = min(g.K, kb + g.kper);
  f4 acc = {0, 0, 0, 0};
  for (int k0 = kb; k0 < ke; k0 += 32) {
    u16x8 a, b;
#pragma unroll
    for (int i = 0; i < 8; i++) {
      const int k = k0 + kq + i;
      a[i] = (am < g.M && k < ke) ? g.A[am * g.sam + k * g.sak] : (uint16_t)0;
      b[i] = (bn < g.N && k < ke) ? g.B[k * g.sbk + bn * g.sbn] : (uint16_t)0;
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf8, a), __builtin_bit_cast(bf8, b), acc,
                                                 0, 0, 0);
  }
  // D of 16x16x32 (f32): lane l, value v -> row 4 (l/16) + v, column l % 16
#pragma unroll
  for (int v = 0; v < 4; v++) {
    const int m = m0 + 4 * (lane >> 4) + v, n = n0 + (lane & 15);
    if (g.ks > 1) {
      if (m < g.M && n < g.N) g.Cs[((int64_t)blockIdx.z * g.M + m) * g.N + n] = acc[v];
      continue;
    }
    if (m < g.M && n < g.N) {
      float z = acc[v] + (g.bias ? g.bias[n] : 0.0f);
      if (g.C) g.C[m * g.ldc + n] = z;
      if (g.Cb) g.Cb[m * g.ldc + n] = bf16_rne(g.silu ? silu(z) : z);
    }
  }
}

// the split-K partials summed in order of z, then the GEMM's epilogue
__global__ void k_gemm_reduce(GemmArgs g) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)g.M * g.N) return;
  const int m = (int)(e / g.N), n = (int)(e % g.N);
  float z = 0;
  for (int k = 0; k < g.ks; k++) z += g.Cs[(int64_t)k * g.M * g.N + e];
  z += g.bias ? g.bias[n] : 0.0f;
  if (g.C) g.C[m * g.ldc + n] = z;
  if (g.Cb) g.Cb[m * g.ldc + n] = bf16_rne(g.silu ? silu(z) : z);
}

__global__ void k_to_bf16(const float *__restrict__ x, uint16_t *__restrict__ y, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = bf16_rne(x[i]);
}

// dY = 2 (y - t) / (B N) (FP32 and bf16), per-block partial sums of (y - t)^2
__global__ void k_mse(const float *__restrict__ y, const float *__restrict__ t, float *__restrict__ dy,
                      uint16_t *__restrict__ dyb, float *__restrict__ part, int64_t n, float scale) {
  __shared__ float red[256];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float e2 = 0;
  if (i < n) {
    const float e = y[i] - t[i];
    const float d = 2.0f * e * scale;
    dy[i] = d;
    dyb[i] = bf16_rne(d);
    e2 = e * e;
  }
  red[threadIdx.x] = e2;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// dZ = dX * silu'(Z), FP32 and bf16
__global__ void k_silu_back(const float *__restrict__ dx, const float *__restrict__ z, float *__restrict__ dz,
                            uint16_t *__restrict__ dzb, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const float d = dx[i] * silu_grad(z[i]);
    dz[i] = d;
    dzb[i] = bf16_rne(d);
  }
}

// db[n] = sum_b dZ[b][n] in two deterministic stages.  k_bias_part: block
// (column group cx, row block rb) sums the rows [rb*rpb, (rb+1)*rpb) of 16
// columns; thread (row lane r, column j) takes the rows r, r + 16, ... with
// four independent accumulators (loads in flight together), and the 16 row
// lanes are added in LDS in a fixed order into part[rb][n].  k_bias_sum adds
// the row blocks' partials in order.  (One block per 16 columns walking all
// rows took about 35 us per call at B = 2048, most of a training step.)
constexpr int MG_BROWS = 16;  // row blocks of the bias gradient
__global__ __launch_bounds__(256) void k_bias_part(const float *__restrict__ dz, float *__restrict__ part, int B,
                                                   int N, int rpb) {
  __shared__ float red[16][17];
  const int j = threadIdx.x & 15, r = threadIdx.x >> 4;
  const int n = blockIdx.x * 16 + j;
  const int b0 = blockIdx.y * rpb, b1 = min(B, b0 + rpb);
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  if (n < N) {
    int b = b0 + r;
    for (; b + 48 < b1; b += 64)
#pragma unroll
      for (int u = 0; u < 4; u++) s[u] += dz[(int64_t)(b + 16 * u) * N + n];
    for (int u = 0; b < b1; b += 16, u++) s[u & 3] += dz[(int64_t)b * N + n];
  }
  red[r][j] = (s[0] + s[1]) + (s[2] + s[3]);
  __syncthreads();
  if (r == 0 && n < N) {
    float t = 0;
    for (int k = 0; k < 16; k++) t += red[k][j];
    part[(int64_t)blockIdx.y * N + n] = t;
  }
}
__global__ void k_bias_sum(const float *__restrict__ part, float *__restrict__ db, int nrb, int N) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float t = 0;
  for (int k = 0; k < nrb; k++) t += part[(int64_t)k * N + n];
  db[n] = t;
}

// the Adam step counter and its bias corrections 1 / (1 - b^step), on the
// device so that a captured training step replays unchanged
__global__ void k_adam_step(int *__restrict__ step, float *__restrict__ coef, float b1, float b2) {
  const int s = ++step[0];
  coef[0] = 1.0f / (1.0f - powf(b1, (float)s));
  coef[1] = 1.0f / (1.0f - powf(b2, (float)s));
}
// Adam (Kingma & Ba 2015) on FP32 master weights; wb: bf16 copy (NULL for biases)
__global__ void k_adam(float *__restrict__ w, float *__restrict__ m, float *__restrict__ v,
                       const float *__restrict__ g, uint16_t *__restrict__ wb, int64_t n, float lr, float b1,
                       float b2, float eps, const float *__restrict__ coef) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float c1 = coef[0], c2 = coef[1];
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.0f - b1) * gi;
  const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  const float wi = w[i] - lr * (mi * c1) / (sqrtf(vi * c2) + eps);
  w[i] = wi;
  if (wb) wb[i] = bf16_rne(wi);
}

static int blocks(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

}  // namespace plg

using namespace plg;

// ------------------------------------------------------------------ the network object
struct plg_mlp_s {
  std::vector<int> dims;  // dims[0] = inputs, dims[L] = outputs
  int L = 0, max_batch = 0;
  int64_t nparams = 0;
  std::vector<int64_t> w_off, b_off;  // offsets of W_l [out][in] and b_l [out] in the parameter vector
  float *p = nullptr, *g = nullptr, *m = nullptr, *v = nullptr;
  uint16_t *pb = nullptr;  // bf16 copy, same offsets (biases unused)
  // activations, per layer l: X_l (bf16, [B][dims[l]]), Z_l (FP32, [B][dims[l+1]])
  std::vector<uint16_t *> X;
  std::vector<float *> Z;
  float *dz = nullptr, *dx = nullptr, *part = nullptr;
  float *bpart = nullptr;  // bias-gradient partials [MG_BROWS][widest]
  uint16_t *dzb = nullptr;
  int64_t step = 0;
  float *ksw = nullptr;     // split-K partials of the products
  int *d_step = nullptr;    // Adam step on the device
  float *d_coef = nullptr;  // its bias corrections
  // one training step captured as a graph, for the arguments it was captured with
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const float *gx = nullptr, *gt = nullptr;
  int gB = 0;
  float glr = 0;
  std::vector<void *> allocs;
  ~plg_mlp_s() {
    if (exec) hipGraphExecDestroy(exec);
    if (graph) hipGraphDestroy(graph);
    if (cap) hipStreamDestroy(cap);
    for (void *q : allocs) hipFree(q);
  }
};

static int mlp_alloc(plg_mlp_s *net, void **dst, size_t bytes) {
  if (hipMalloc(dst, bytes) != hipSuccess) return fail(PLG_ERR_ALLOC, "MLP device allocation failed");
  net->allocs.push_back(*dst);
  return PLG_OK;
}

// K split for long, narrow products (the weight gradients sum over the
// batch): up to MG_KSPLIT ranges of at least 128, when the tile grid is small
constexpr int MG_KSPLIT = 16;
static int gemm_ksplit(int M, int N, int K) {
  const int64_t tiles = (int64_t)((N + 31) / 32) * ((M + 31) / 32);
  if (K < 512 || tiles >= 512) return 1;
  return std::max(1, std::min(MG_KSPLIT, K / 128));
}
static void gemm(GemmArgs a, float *ws, hipStream_t st) {
  a.ks = ws ? gemm_ksplit(a.M, a.N, a.K) : 1;
  a.kper = a.ks > 1 ? ((a.K + a.ks - 1) / a.ks + 31) / 32 * 32 : a.K;
  a.ks = a.ks > 1 ? (a.K + a.kper - 1) / a.kper : 1;
  a.Cs = ws;
  dim3 grid((unsigned)((a.N + 31) / 32), (unsigned)((a.M + 31) / 32), (unsigned)a.ks);
  k_gemm_bf16<<<grid, MG_THREADS, 0, st>>>(a);
  if (a.ks > 1) k_gemm_reduce<<<blocks((int64_t)a.M * a.N), 256, 0, st>>>(a);
}

static int mlp_forward(plg_mlp_s *net, const float *x, int B, hipStream_t st) {
  const int64_t n0 = (int64_t)B * net->dims[0];
  k_to_bf16<<<blocks(n0), 256, 0, st>>>(x, net->X[0], n0);
  for (int l = 0; l < net->L; l++) {
    const int Ni = net->dims[l], No = net->dims[l + 1];
    const bool last = l == net->L - 1;
    GemmArgs a{};
    a.M = B; a.N = No; a.K = Ni;
    a.A = net->X[l]; a.sam = Ni; a.sak = 1;              // X_l [b][i]
    a.B = net->pb + net->w_off[l]; a.sbk = 1; a.sbn = Ni;  // W_l^T: (i, o) -> W[o][i]
    a.C = net->Z[l]; a.ldc = No;
    a.Cb = last ? nullptr : net->X[l + 1];
    a.bias = net->p + net->b_off[l];
    a.silu = 1;
    gemm(a, net->ksw, st);
  }
  return hipGetLastError() == hipSuccess ? PLG_OK : fail(PLG_ERR_HIP, "MLP forward launch failed");
}

// loss partials + full backward into net->g
static int mlp_backward(plg_mlp_s *net, const float *t, int B, hipStream_t st) {
  const int Lx = net->L;
  const int64_t ny = (int64_t)B * net->dims[Lx];
  k_mse<<<blocks(ny), 256, 0, st>>>(net->Z[Lx - 1], t, net->dz, net->dzb, net->part, ny,
                                   1.0f / (float)ny);
  for (int l = Lx - 1; l >= 0; l--) {
    const int Ni = net->dims[l], No = net->dims[l + 1];
    // dW_l[o][i] = sum_b dZ[b][o] X_l[b][i]
    GemmArgs a{};
    a.M = No; a.N = Ni; a.K = B;
    a.A = net->dzb; a.sam = 1; a.sak = No;
    a.B = net->X[l]; a.sbk = Ni; a.sbn = 1;
    a.C = net->g + net->w_off[l]; a.ldc = Ni;
    gemm(a, net->ksw, st);
    {
      const int rpb = (B + MG_BROWS - 1) / MG_BROWS;
      k_bias_part<<<dim3(blocks(No, 16), MG_BROWS), 256, 0, st>>>(net->dz, net->bpart, B, No, rpb);
      k_bias_sum<<<blocks(No), 256, 0, st>>>(net->bpart, net->g + net->b_off[l], MG_BROWS, No);
    }
    if (l > 0) {
      // dX_l[b][i] = sum_o dZ[b][o] W_l[o][i], then dZ_{l-1} = dX_l silu'(Z_{l-1})
      GemmArgs d{};
      d.M = B; d.N = Ni; d.K = No;
      d.A = net->dzb; d.sam = No; d.sak = 1;
      d.B = net->pb + net->w_off[l]; d.sbk = Ni; d.sbn = 1;
      d.C = net->dx; d.ldc = Ni;
      gemm(d, net->ksw, st);
      const int64_t n = (int64_t)B * Ni;
      k_silu_back<<<blocks(n), 256, 0, st>>>(net->dx, net->Z[l - 1], net->dz, net->dzb, n);
    }
  }
  return hipGetLastError() == hipSuccess ? PLG_OK : fail(PLG_ERR_HIP, "MLP backward launch failed");
}

static int mlp_loss(plg_mlp_s *net, int B, hipStream_t st, double *loss) {
  const int64_t ny = (int64_t)B * net->dims[net->L];
  const int nb = blocks(ny);
  std::vector<float> h(nb);
  if (hipMemcpyAsync(h.data(), net->part, nb * sizeof(float), hipMemcpyDeviceToHost, st) != hipSuccess ||
      hipStreamSynchronize(st) != hipSuccess)
    return fail(PLG_ERR_HIP, "MLP loss read-back failed");
  double s = 0;
  for (float v : h) s += v;
  *loss = s / (double)ny;
  return PLG_OK;
}

// one Adam step: forward, backward, step counter, updates (no allocation,
// no memset: capturable)
static int train_step(plg_mlp_s *net, const float *x, const float *t, int B, float lr, hipStream_t st) {
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  int rc;
  if ((rc = mlp_forward(net, x, B, st)) || (rc = mlp_backward(net, t, B, st))) return rc;
  k_adam_step<<<1, 1, 0, st>>>(net->d_step, net->d_coef, b1, b2);
  // the parameters (W_0, b_0, W_1, ...) are one contiguous array: one update
  // launch for all of them (the biases' bf16 copies are written but unused)
  k_adam<<<blocks(net->nparams), 256, 0, st>>>(net->p, net->m, net->v, net->g, net->pb, net->nparams, lr, b1, b2,
                                               eps, net->d_coef);
  return hipGetLastError() == hipSuccess ? PLG_OK : fail(PLG_ERR_HIP, "MLP update launch failed");
}

static int check_batch(plg_mlp_s *net, int B) {
  if (!net) return fail(PLG_ERR_INVALID, "null MLP handle");
  if (B <= 0 || B > net->max_batch)
    return fail(PLG_ERR_RANGE, "batch " + std::to_string(B) + " outside 1.." + std::to_string(net->max_batch));
  return PLG_OK;
}

extern "C" {

int plg_mlp_create(const int *dims, int nlayers, int max_batch, plg_mlp *out) {
  if (!dims || !out || nlayers < 1 || nlayers > 16 || max_batch <= 0)
    return fail(PLG_ERR_INVALID, "plg_mlp_create: bad arguments");
  for (int l = 0; l <= nlayers; l++)
    if (dims[l] <= 0 || dims[l] > 65536) return fail(PLG_ERR_INVALID, "plg_mlp_create: bad layer width");
  auto *net = new plg_mlp_s();
  net->dims.assign(dims, dims + nlayers + 1);
  net->L = nlayers;
  net->max_batch = max_batch;
  int64_t off = 0, widest = 0;
  for (int l = 0; l < nlayers; l++) {
    net->w_off.push_back(off);
    off += (int64_t)dims[l] * dims[l + 1];
    net->b_off.push_back(off);
    off += dims[l + 1];
  }
  for (int l = 0; l <= nlayers; l++) widest = std::max<int64_t>(widest, dims[l]);
  net->nparams = off;
  int rc = PLG_OK;
  const size_t pbytes = (size_t)off * sizeof(float);
  rc = rc ? rc : mlp_alloc(net, (void **)&net->p, pbytes);
  rc = rc ? rc : mlp_alloc(net, (void **)&net->g, pbytes);
  rc = rc ? rc : mlp_alloc(net, (void **)&net->m, pbytes);
  rc = rc ? rc : mlp_alloc(net, (void **)&net->v, pbytes);
  rc = rc ? rc : mlp_alloc(net, (void **)&net->pb, (size_t)off * sizeof(uint16_t));
  for (int l = 0; l < nlayers && !rc; l++) {
    uint16_t *xl;
    float *zl;
    rc = mlp_alloc(net, (void **)&xl, (size_t)max_batch * dims[l] * sizeof(uint16_t));
    rc = rc ? rc : mlp_alloc(net, (void **)&zl, (size_t)max_batch * dims[l + 1] * sizeof(float));
    net->X.push_back(xl);
    net->Z.push_back(zl);
  }
  const size_t wb = (size_t)max_batch * widest;
  rc = rc ? rc : mlp_alloc(net, (void **)&net->dz, wb * sizeof(float));
  rc = rc ? rc : mlp_alloc(net, (void **)&net->dx, wb * sizeof(float));
  rc = rc ? rc : mlp_alloc(net, (void **)&net->dzb, wb * sizeof(uint16_t));
  rc = rc ? rc : mlp_alloc(net, (void **)&net->part, (size_t)blocks((int64_t)wb) * sizeof(float));
  rc = rc ? rc : mlp_alloc(net, (void **)&net->bpart, (size_t)MG_BROWS * widest * sizeof(float));
  {
    // split-K partials: a bound of ks x M x N over the network's products
    // at any batch up to max_batch (the split ignoring the tile count)
    size_t kw = 1;
    for (int l = 0; l < nlayers; l++) {
      const int Ni = dims[l], No = dims[l + 1];
      const int64_t shapes[3][3] = {{max_batch, No, Ni}, {No, Ni, max_batch}, {max_batch, Ni, No}};
      for (auto &sh : shapes)
        if (sh[2] >= 512) kw = std::max(kw, (size_t)std::min<int64_t>(MG_KSPLIT, sh[2] / 128) * sh[0] * sh[1]);
    }
    rc = rc ? rc : mlp_alloc(net, (void **)&net->ksw, kw * sizeof(float));
  }
  rc = rc ? rc : mlp_alloc(net, (void **)&net->d_step, sizeof(int));
  rc = rc ? rc : mlp_alloc(net, (void **)&net->d_coef, 2 * sizeof(float));
  if (!rc && hipMemset(net->d_step, 0, sizeof(int)) != hipSuccess) rc = fail(PLG_ERR_HIP, "MLP step clear failed");
  if (!rc && (hipMemset(net->p, 0, pbytes) != hipSuccess || hipMemset(net->pb, 0, off * sizeof(uint16_t)) != hipSuccess))
    rc = fail(PLG_ERR_HIP, "MLP parameter clear failed");
  if (rc) {
    delete net;
    return rc;
  }
  *out = net;
  return PLG_OK;
}

int plg_mlp_nparams(plg_mlp net, int64_t *n) {
  if (!net || !n) return fail(PLG_ERR_INVALID, "null argument");
  *n = net->nparams;
  return PLG_OK;
}

// parameters: for each layer, W [out][in] then b [out], FP32.  Resets Adam.
int plg_mlp_set_params(plg_mlp net, const float *host) {
  if (!net || !host) return fail(PLG_ERR_INVALID, "null argument");
  const size_t bytes = (size_t)net->nparams * sizeof(float);
  if (hipMemcpy(net->p, host, bytes, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemset(net->m, 0, bytes) != hipSuccess || hipMemset(net->v, 0, bytes) != hipSuccess)
    return fail(PLG_ERR_HIP, "MLP parameter upload failed");
  k_to_bf16<<<blocks(net->nparams), 256>>>(net->p, net->pb, net->nparams);
  net->step = 0;
  if (hipMemset(net->d_step, 0, sizeof(int)) != hipSuccess) return fail(PLG_ERR_HIP, "MLP step clear failed");
  if (hipDeviceSynchronize() != hipSuccess) return fail(PLG_ERR_HIP, "MLP parameter conversion failed");
  return PLG_OK;
}

int plg_mlp_get_params(plg_mlp net, float *host) {
  if (!net || !host) return fail(PLG_ERR_INVALID, "null argument");
  if (hipDeviceSynchronize() != hipSuccess ||
      hipMemcpy(host, net->p, (size_t)net->nparams * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
    return fail(PLG_ERR_HIP, "MLP parameter read-back failed");
  return PLG_OK;
}

int plg_mlp_forward(plg_mlp net, const float *x, int batch, float *y, void *stream) {
  int rc = check_batch(net, batch);
  if (rc) return rc;
  if (!x || !y) return fail(PLG_ERR_INVALID, "null argument");
  hipStream_t st = (hipStream_t)stream;
  if ((rc = mlp_forward(net, x, batch, st))) return rc;
  const size_t bytes = (size_t)batch * net->dims[net->L] * sizeof(float);
  if (hipMemcpyAsync(y, net->Z[net->L - 1], bytes, hipMemcpyDeviceToDevice, st) != hipSuccess)
    return fail(PLG_ERR_HIP, "MLP output copy failed");
  return PLG_OK;
}

int plg_mlp_gradient(plg_mlp net, const float *x, const float *t, int batch, double *loss, float *grad_host,
                     void *stream) {
  int rc = check_batch(net, batch);
  if (rc) return rc;
  if (!x || !t) return fail(PLG_ERR_INVALID, "null argument");
  hipStream_t st = (hipStream_t)stream;
  if ((rc = mlp_forward(net, x, batch, st)) || (rc = mlp_backward(net, t, batch, st))) return rc;
  if (loss && (rc = mlp_loss(net, batch, st, loss))) return rc;
  if (grad_host) {
    if (hipMemcpyAsync(grad_host, net->g, (size_t)net->nparams * sizeof(float), hipMemcpyDeviceToHost, st) !=
            hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return fail(PLG_ERR_HIP, "MLP gradient read-back failed");
  }
  return PLG_OK;
}

int plg_mlp_train(plg_mlp net, const float *x, const float *t, int batch, int steps, float lr, double *loss,
                  void *stream) {
  int rc = check_batch(net, batch);
  if (rc) return rc;
  if (!x || !t || steps < 0 || !(lr > 0)) return fail(PLG_ERR_INVALID, "plg_mlp_train: bad arguments");
  hipStream_t st = (hipStream_t)stream;
  // Steps are replayed from a graph of one step (captured again when the
  // inputs, batch or rate change) unless PLG_MLP_EAGER=1 or the call is
  // short; both forms launch the same kernels with the same arguments.
  const char *ev = std::getenv("PLG_MLP_EAGER");
  const bool eager = steps < 4 || (ev && std::atoi(ev) != 0);
  if (eager) {
    for (int s = 0; s < steps; s++)
      if ((rc = train_step(net, x, t, batch, lr, st))) return rc;
  } else {
    if (!net->exec || net->gx != x || net->gt != t || net->gB != batch || net->glr != lr) {
      if (net->exec) hipGraphExecDestroy(net->exec);
      if (net->graph) hipGraphDestroy(net->graph);
      net->exec = nullptr;
      net->graph = nullptr;
      if (!net->cap && hipStreamCreateWithFlags(&net->cap, hipStreamNonBlocking) != hipSuccess)
        return fail(PLG_ERR_HIP, "MLP capture stream creation failed");
      // the capture stream starts after the work queued on `stream`
      hipEvent_t ev0;
      if (hipEventCreateWithFlags(&ev0, hipEventDisableTiming) != hipSuccess) return fail(PLG_ERR_HIP, "event");
      hipEventRecord(ev0, st);
      hipStreamWaitEvent(net->cap, ev0, 0);
      hipEventDestroy(ev0);
      if (hipStreamBeginCapture(net->cap, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return fail(PLG_ERR_HIP, "MLP step capture failed");
      rc = train_step(net, x, t, batch, lr, net->cap);
      hipGraph_t g = nullptr;
      const hipError_t ce = hipStreamEndCapture(net->cap, &g);
      if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
      }
      if (ce != hipSuccess || !g || hipGraphInstantiate(&net->exec, g, nullptr, nullptr, 0) != hipSuccess) {
        if (g) hipGraphDestroy(g);
        return fail(PLG_ERR_HIP, "MLP step graph instantiation failed");
      }
      net->graph = g;
      net->gx = x;
      net->gt = t;
      net->gB = batch;
      net->glr = lr;
    }
    for (int s = 0; s < steps; s++)
      if (hipGraphLaunch(net->exec, st) != hipSuccess) return fail(PLG_ERR_HIP, "MLP step graph launch failed");
  }
  net->step += steps;
  if (loss) {
    // loss of the parameters after the last update
    if ((rc = mlp_forward(net, x, batch, st)) || (rc = mlp_backward(net, t, batch, st))) return rc;
    if ((rc = mlp_loss(net, batch, st, loss))) return rc;
  }
  return PLG_OK;
}

int plg_mlp_destroy(plg_mlp net) {
  if (!net) return PLG_OK;
  hipDeviceSynchronize();
  delete net;
  return PLG_OK;
}

}  // extern "C"
