// Small MLP that scores SPR moves (32 features -> 128 -> 128 -> 1, GELU), with
// hand-written training and inference kernels: BF16-input / FP32-accumulate
// MFMA GEMMs (v_mfma_f32_16x16x32_bf16), FP32 master weights and Adam.
// The reference has no learned component; this serves the learned SPR
// proposal that re-ranks TreeNeighborhoods moves (lib/treeNeighborhoods.ml).
//
// Parameters (flat FP32): W1[n_in][H] b1[H] W2[H][H] b2[H] w3[H] b3[1].
// Loss: 0.5 * mean((y - t)^2).
#include <algorithm>
#include <cmath>

#include "common.hpp"
#include "rng.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

struct phy_mlp {
  phy_ctx* ctx = nullptr;
  int n_in = 32, H = 128;
  int64_t n_params = 0;
  float *params = nullptr, *grads = nullptr, *m1 = nullptr, *m2 = nullptr;
  int64_t step = 0;
  // activations for up to cap rows: Z1 H1 Z2 H2 [cap][H], y [cap], dy, dH [cap][H] x2
  float* act = nullptr;
  int64_t cap = 0;
  // partial results of the split-K weight-gradient GEMMs and column sums
  float* ws = nullptr;
  int64_t ws_floats = 0;
};

namespace {

__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad(float x) {
  const float cdf = 0.5f * (1.0f + erff(x * 0.70710678118654752f));
  const float pdf = 0.3989422804014327f * expf(-0.5f * x * x);
  return cdf + x * pdf;
}

// C[M][N] (+)= op(A)[M][K] op(B)[K][N] (+ bias[n]), A at A[m*sam + k*sak],
// B at B[k*sbk + n*sbn], C row-major with ldc.  One wave per 16x16 tile;
// operands are rounded to BF16, accumulation is FP32.  Split K: workgroup z
// takes k in [z * kps, (z + 1) * kps) and writes its partial tile at
// Cm + z * c_split (the caller sums the splits).
__global__ __launch_bounds__(64) void k_gemm(const float* __restrict__ A, int64_t sam, int64_t sak,
                                             const float* __restrict__ B, int64_t sbk, int64_t sbn,
                                             float* __restrict__ Cm, int64_t ldc, int M, int N, int K,
                                             int accumulate, const float* __restrict__ bias,
                                             int kps = 1 << 30, int64_t c_split = 0) {
  const int l = threadIdx.x, q = l >> 4, r16 = l & 15;
  const int m0 = blockIdx.y * 16, n0 = blockIdx.x * 16;
  const int am = m0 + r16, bn = n0 + r16;
  const int kb = (int)blockIdx.z * kps, ke = min(K, kb + kps);
  Cm += (int64_t)blockIdx.z * c_split;
  f32x4 acc = {0.f, 0.f, 0.f, 0.f};
  for (int k0 = kb; k0 < ke; k0 += 32) {
    bf16x8 a, b;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const int k = k0 + 8 * q + v;
      const bool ka = k < ke;
      a[v] = (__bf16)((ka && am < M) ? A[(int64_t)am * sam + (int64_t)k * sak] : 0.0f);
      b[v] = (__bf16)((ka && bn < N) ? B[(int64_t)k * sbk + (int64_t)bn * sbn] : 0.0f);
    }
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, acc, 0, 0, 0);
  }
  // lane l holds C[m0 + 4q + v][n0 + l%16]
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int m = m0 + 4 * q + v, n = n0 + r16;
    if (m < M && n < N) {
      float x = acc[v] + (bias ? bias[n] : 0.0f);
      float* c = Cm + (int64_t)m * ldc + n;
      *c = accumulate ? *c + x : x;
    }
  }
}

__global__ void k_gelu_fwd(const float* __restrict__ Z, float* __restrict__ Hh, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) Hh[i] = gelu(Z[i]);
}

// y[m] = H2[m] . w3 + b3  (FP32 dot product: the 128 -> 1 layer is a GEMV)
__global__ void k_out(const float* __restrict__ H2, const float* __restrict__ w3,
                      const float* __restrict__ b3, float* __restrict__ y, int64_t M, int H) {
  const int64_t m = (int64_t)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  if (m >= M) return;
  float s = 0.f;
  for (int j = l; j < H; j += 64) s += H2[m * H + j] * w3[j];
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
  if (l == 0) y[m] = s + b3[0];
}

// dy = (y - t) / B; loss partials; dH2 = dy * w3; db3, dw3 accumulated later
__global__ void k_loss(const float* __restrict__ y, const float* __restrict__ t, float* __restrict__ dy,
                       double* __restrict__ loss, int64_t M) {
  const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double e = 0.0;
  if (m < M) {
    const float d = y[m] - t[m];
    dy[m] = d / (float)M;
    e = 0.5 * (double)d * (double)d / (double)M;
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) e += __shfl_xor(e, off, 64);
  if ((threadIdx.x & 63) == 0 && loss) atomicAdd(loss, e);
}

// dZ[m][j] = (dy[m] * w3[j] or dH[m][j]) * gelu'(Z[m][j])
__global__ void k_gelu_bwd(const float* __restrict__ dH, const float* __restrict__ dy,
                           const float* __restrict__ w3, const float* __restrict__ Z,
                           float* __restrict__ dZ, int64_t M, int H) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M * H) return;
  const int64_t m = i / H;
  const int j = (int)(i % H);
  const float g = dH ? dH[i] : dy[m] * w3[j];
  dZ[i] = g * gelu_grad(Z[i]);
}

// out[i] = sum over s < splits of part[s * n + i] (fixed order), written or
// added to out
__global__ void k_sum_splits(const float* __restrict__ part, int splits, int64_t n,
                             float* __restrict__ out, int accumulate) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += part[(int64_t)z * n + i];
  out[i] = accumulate ? out[i] + s : s;
}

// column sums by row chunks: part[c][j] = sum of X[m][j] over the rows of
// chunk c (threads over columns: coalesced rows)
__global__ void k_colsum_chunks(const float* __restrict__ X, int64_t M, int H, int64_t rows,
                                float* __restrict__ part) {
  const int64_t c = blockIdx.x;
  const int64_t m0 = c * rows, m1 = min(M, m0 + rows);
  for (int j = threadIdx.x; j < H; j += blockDim.x) {
    float s = 0.f;
    for (int64_t m = m0; m < m1; ++m) s += X[m * H + j];
    part[c * H + j] = s;
  }
}

__global__ void k_adam(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, int64_t n, float lr, float b1, float b2, float eps,
                       float c1, float c2) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.f - b1) * gi;
  const float b = b2 * m2[i] + (1.f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  p[i] -= lr * (a / c1) / (sqrtf(b / c2) + eps);
}

struct Views {
  float *W1, *b1, *W2, *b2, *w3, *b3;
};
Views views(float* base, int n_in, int H) {
  Views v;
  v.W1 = base;
  v.b1 = v.W1 + (int64_t)n_in * H;
  v.W2 = v.b1 + H;
  v.b2 = v.W2 + (int64_t)H * H;
  v.w3 = v.b2 + H;
  v.b3 = v.w3 + H;
  return v;
}

// split-K plan of a weight gradient (reduction over K = the batch): enough
// workgroups to fill the chip, each a multiple of 32 rows long
struct SplitK {
  int splits, kps;
};
SplitK split_k(int64_t K, int tiles) {
  SplitK s;
  const int64_t want = std::max<int64_t>(1, std::min<int64_t>((4096 + tiles - 1) / tiles, (K + 255) / 256));
  s.kps = (int)((((K + want - 1) / want) + 31) / 32 * 32);
  s.splits = (int)((K + s.kps - 1) / s.kps);
  return s;
}

// grad (M x N, ldc = N) = A^T B over K rows, split over K into m->ws, then summed
int grad_gemm(phy_ctx* ctx, phy_mlp* m, const float* A, int64_t lda, const float* Bm, int64_t ldb,
              float* grad, int M, int N, int64_t K) {
  const dim3 t((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16));
  const SplitK s = split_k(K, (int)(t.x * t.y));
  const int64_t n = (int64_t)M * N;
  hipLaunchKernelGGL(k_gemm, dim3(t.x, t.y, (unsigned)s.splits), dim3(64), 0, ctx->stream, A, (int64_t)1,
                     lda, Bm, ldb, (int64_t)1, m->ws, (int64_t)N, M, N, (int)K, 0, (const float*)nullptr,
                     s.kps, n);
  hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, m->ws,
                     s.splits, n, grad, 0);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

// out[j] = sum_m X[m][j], by row chunks into m->ws, then summed
int colsum(phy_ctx* ctx, phy_mlp* m, const float* X, int64_t M, int H, float* out) {
  const int64_t rows = std::max<int64_t>(64, (M + 1023) / 1024);
  const int64_t chunks = (M + rows - 1) / rows;
  hipLaunchKernelGGL(k_colsum_chunks, dim3((unsigned)chunks), dim3(std::min(H, 256)), 0, ctx->stream, X, M,
                     H, rows, m->ws);
  hipLaunchKernelGGL(k_sum_splits, dim3((unsigned)((H + 255) / 256)), dim3(256), 0, ctx->stream, m->ws,
                     (int)chunks, (int64_t)H, out, 0);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

int ensure_cap(phy_mlp* m, int64_t B) {
  if (B <= m->cap) return PHY_OK;
  if (m->act) PHY_HIP(m->ctx, hipFree(m->act));
  m->act = nullptr;
  const int64_t cap = std::max<int64_t>(B, 1024);
  PHY_HIP(m->ctx, hipMalloc((void**)&m->act, sizeof(float) * (size_t)cap * (6 * m->H + 2)));
  m->cap = cap;
  return PHY_OK;
}

inline dim3 tiles(int M, int N) { return dim3((unsigned)((N + 15) / 16), (unsigned)((M + 15) / 16)); }

// forward; activations kept for backward
int forward(phy_ctx* ctx, phy_mlp* m, const float* X, int64_t B, float* y) {
  const int H = m->H, n_in = m->n_in;
  Views p = views(m->params, n_in, H);
  float* Z1 = m->act;
  float* H1 = Z1 + m->cap * H;
  float* Z2 = H1 + m->cap * H;
  float* H2 = Z2 + m->cap * H;
  const int Mi = (int)B;
  KernelSpan span(ctx, "mlp_forward");
  hipLaunchKernelGGL(k_gemm, tiles(Mi, H), dim3(64), 0, ctx->stream, X, (int64_t)n_in, (int64_t)1, p.W1,
                     (int64_t)H, (int64_t)1, Z1, (int64_t)H, Mi, H, n_in, 0, p.b1, 1 << 30, (int64_t)0);
  const int64_t nh = B * H;
  hipLaunchKernelGGL(k_gelu_fwd, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, ctx->stream, Z1, H1, nh);
  hipLaunchKernelGGL(k_gemm, tiles(Mi, H), dim3(64), 0, ctx->stream, H1, (int64_t)H, (int64_t)1, p.W2,
                     (int64_t)H, (int64_t)1, Z2, (int64_t)H, Mi, H, H, 0, p.b2, 1 << 30, (int64_t)0);
  hipLaunchKernelGGL(k_gelu_fwd, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, ctx->stream, Z2, H2, nh);
  hipLaunchKernelGGL(k_out, dim3((unsigned)((B + 3) / 4)), dim3(256), 0, ctx->stream, H2, p.w3, p.b3, y, B, H);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

int backward(phy_ctx* ctx, phy_mlp* m, const float* X, const float* t, int64_t B, double* loss) {
  const int H = m->H, n_in = m->n_in;
  Views p = views(m->params, n_in, H);
  Views g = views(m->grads, n_in, H);
  float* Z1 = m->act;
  float* H1 = Z1 + m->cap * H;
  float* Z2 = H1 + m->cap * H;
  float* H2 = Z2 + m->cap * H;
  float* dZ2 = H2 + m->cap * H;
  float* dH1 = dZ2 + m->cap * H;  // reused as dZ1
  float* y = dH1 + m->cap * H;
  float* dy = y + m->cap;
  const int Mi = (int)B;
  int rc = forward(ctx, m, X, B, y);
  if (rc) return rc;
  KernelSpan span(ctx, "mlp_backward");
  hipLaunchKernelGGL(k_loss, dim3((unsigned)((B + 255) / 256)), dim3(256), 0, ctx->stream, y, t, dy, loss, B);
  // output layer: dw3 = H2^T dy (GEMM with N = 1), db3 = sum dy
  if ((rc = grad_gemm(ctx, m, H2, H, dy, 1, g.w3, H, 1, B))) return rc;
  if ((rc = colsum(ctx, m, dy, B, 1, g.b3))) return rc;
  const int64_t nh = B * H;
  hipLaunchKernelGGL(k_gelu_bwd, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, ctx->stream,
                     (const float*)nullptr, dy, p.w3, Z2, dZ2, B, H);
  // dW2 = H1^T dZ2 ; db2 = colsum dZ2 ; dH1 = dZ2 W2^T
  if ((rc = grad_gemm(ctx, m, H1, H, dZ2, H, g.W2, H, H, B))) return rc;
  if ((rc = colsum(ctx, m, dZ2, B, H, g.b2))) return rc;
  hipLaunchKernelGGL(k_gemm, tiles(Mi, H), dim3(64), 0, ctx->stream, dZ2, (int64_t)H, (int64_t)1, p.W2,
                     (int64_t)1, (int64_t)H, dH1, (int64_t)H, Mi, H, H, 0, (const float*)nullptr,
                     1 << 30, (int64_t)0);
  // dZ1 = dH1 * gelu'(Z1) (in place) ; dW1 = X^T dZ1 ; db1 = colsum dZ1
  hipLaunchKernelGGL(k_gelu_bwd, dim3((unsigned)((nh + 255) / 256)), dim3(256), 0, ctx->stream, dH1,
                     (const float*)nullptr, (const float*)nullptr, Z1, dH1, B, H);
  if ((rc = grad_gemm(ctx, m, X, n_in, dH1, H, g.W1, n_in, H, B))) return rc;
  if ((rc = colsum(ctx, m, dH1, B, H, g.b1))) return rc;
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

}  // namespace

extern "C" {

int phy_mlp_create(phy_ctx* ctx, int32_t n_in, int32_t n_hidden, uint64_t seed, phy_mlp** out) {
  if (!ctx || !out) return PHY_EINVAL;
  *out = nullptr;
  if (n_in < 1 || n_in > 4096 || n_hidden < 16 || n_hidden > 4096 || n_hidden % 16)
    return set_err(ctx, PHY_EINVAL, "mlp: n_in in 1..4096, n_hidden a multiple of 16");
  phy_mlp* m = new phy_mlp();
  m->ctx = ctx;
  m->n_in = n_in;
  m->H = n_hidden;
  m->n_params = (int64_t)n_in * n_hidden + n_hidden + (int64_t)n_hidden * n_hidden + n_hidden + n_hidden + 1;
  // He-style initialisation on the host, deterministic in seed
  std::vector<float> p((size_t)m->n_params, 0.0f);
  phy::Rng r(seed, 17);
  Views v = views(p.data(), n_in, n_hidden);
  for (int64_t i = 0; i < (int64_t)n_in * n_hidden; ++i) v.W1[i] = (float)(r.normal() * std::sqrt(2.0 / n_in));
  for (int64_t i = 0; i < (int64_t)n_hidden * n_hidden; ++i) v.W2[i] = (float)(r.normal() * std::sqrt(2.0 / n_hidden));
  for (int i = 0; i < n_hidden; ++i) v.w3[i] = (float)(r.normal() * std::sqrt(1.0 / n_hidden));
  const size_t bytes = sizeof(float) * (size_t)m->n_params;
  hipError_t e = hipMalloc((void**)&m->params, bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&m->grads, bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&m->m1, bytes);
  if (e == hipSuccess) e = hipMalloc((void**)&m->m2, bytes);
  if (e == hipSuccess) e = hipMemcpy(m->params, p.data(), bytes, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemset(m->m1, 0, bytes);
  if (e == hipSuccess) e = hipMemset(m->m2, 0, bytes);
  if (e == hipSuccess) e = hipMemset(m->grads, 0, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    phy_mlp_destroy(m);
    return set_err(ctx, PHY_ENOMEM, "mlp: device allocation failed");
  }
  *out = m;
  return PHY_OK;
}

void phy_mlp_destroy(phy_mlp* m) {
  if (!m) return;
  if (m->params) (void)hipFree(m->params);
  if (m->grads) (void)hipFree(m->grads);
  if (m->m1) (void)hipFree(m->m1);
  if (m->m2) (void)hipFree(m->m2);
  if (m->act) (void)hipFree(m->act);
  if (m->ws) (void)hipFree(m->ws);
  delete m;
}

int64_t phy_mlp_n_params(const phy_mlp* m) { return m ? m->n_params : 0; }

int phy_mlp_get_params(phy_ctx* ctx, const phy_mlp* m, float* h_out) {
  if (!ctx || !m || !h_out) return PHY_EINVAL;
  PHY_HIP(ctx, hipMemcpyAsync(h_out, m->params, sizeof(float) * m->n_params, hipMemcpyDeviceToHost, ctx->stream));
  PHY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PHY_OK;
}

int phy_mlp_set_params(phy_ctx* ctx, phy_mlp* m, const float* h_in) {
  if (!ctx || !m || !h_in) return PHY_EINVAL;
  PHY_HIP(ctx, hipMemcpyAsync(m->params, h_in, sizeof(float) * m->n_params, hipMemcpyHostToDevice, ctx->stream));
  PHY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PHY_OK;
}

int phy_mlp_forward(phy_ctx* ctx, phy_mlp* m, const float* d_x, int64_t B, float* d_y) {
  if (!ctx || !m || !d_x || !d_y || B < 0) return PHY_EINVAL;
  if (B == 0) return PHY_OK;
  if (B > (1 << 30)) return set_err(ctx, PHY_ERANGE, "mlp: batch too large");
  int rc = ensure_cap(m, B);
  if (rc) return rc;
  return forward(ctx, m, d_x, B, d_y);
}

int phy_mlp_backward(phy_ctx* ctx, phy_mlp* m, const float* d_x, const float* d_t, int64_t B,
                     float* d_grads, double* d_loss) {
  if (!ctx || !m || !d_x || !d_t || B < 1) return PHY_EINVAL;
  if (B > (1 << 30)) return set_err(ctx, PHY_ERANGE, "mlp: batch too large");
  int rc = ensure_cap(m, B);
  if (rc) return rc;
  {
    // split-K partials of the three weight gradients, and the column-sum chunks
    const int H = m->H;
    auto part = [&](int M, int N) {
      const SplitK s = split_k(B, ((N + 15) / 16) * ((M + 15) / 16));
      return (int64_t)s.splits * M * N;
    };
    const int64_t rows = std::max<int64_t>(64, (B + 1023) / 1024);
    const int64_t need = std::max({part(H, 1), part(H, H), part(m->n_in, H),
                                   ((B + rows - 1) / rows) * (int64_t)H});
    if (m->ws_floats < need) {
      if (m->ws) PHY_HIP(ctx, hipFree(m->ws));
      m->ws = nullptr;
      m->ws_floats = 0;
      PHY_HIP(ctx, hipMalloc((void**)&m->ws, sizeof(float) * (size_t)need));
      m->ws_floats = need;
    }
  }
  if (d_loss && (rc = phy_zero_u64(ctx, reinterpret_cast<uint64_t*>(d_loss), 1))) return rc;
  rc = backward(ctx, m, d_x, d_t, B, d_loss);
  if (rc) return rc;
  if (d_grads)
    PHY_HIP(ctx, hipMemcpyAsync(d_grads, m->grads, sizeof(float) * m->n_params, hipMemcpyDeviceToDevice, ctx->stream));
  return PHY_OK;
}

int phy_mlp_step(phy_ctx* ctx, phy_mlp* m, const float* d_x, const float* d_t, int64_t B, float lr,
                 double* d_loss) {
  int rc = phy_mlp_backward(ctx, m, d_x, d_t, B, nullptr, d_loss);
  if (rc) return rc;
  m->step += 1;
  const float b1 = 0.9f, b2 = 0.999f;
  const float c1 = 1.0f - std::pow(b1, (float)m->step), c2 = 1.0f - std::pow(b2, (float)m->step);
  hipLaunchKernelGGL(k_adam, dim3((unsigned)((m->n_params + 255) / 256)), dim3(256), 0, ctx->stream,
                     m->params, m->grads, m->m1, m->m2, m->n_params, lr, b1, b2, 1e-8f, c1, c2);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

}  // extern "C"
