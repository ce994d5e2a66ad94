// Felsenstein pruning for 4-state (DNA) models on gfx950.
//
// Likelihood_c is all TODO in the reference (lib/likelihood_c.ml:1-33); the
// algorithm is textbook Felsenstein (1981) pruning (cited at
// lib/mlModel.mli:30-31) over the reference's model layer (Q, pi, Gamma rates,
// P(t) = exp(Qt); lib/mlModel.ml:204-213,390-421,585-587, lib/mlmodel.c:325-342).
//
// Mapping: one thread owns one site pattern and all R rate classes; the tree
// is walked with the uniform post-order program of phy_topo_create, so every
// transition matrix P(rate_r t_b) is wave-uniform and is read through the
// scalar cache into SGPRs, and each 4x4 matrix-vector product is 16 FMAs
// with an SGPR operand.  Partials are R x 4 values in registers; intermediate
// results wait in the accumulator or, in the fused form, in a per-thread LDS
// stack (depth <= ceil(log2 N)); in the stored form every internal partial is
// written once to HBM (structure of arrays, pattern-contiguous) for the
// derivative pass and a stacked operand is read back from there.
//
// Underflow: after every node the partial is rescaled by an exact power of
// two when its maximum drops below 2^-24 (FP32) or 2^-128 (FP64); the
// per-pattern int32 exponent carries the factor, so FP32 partials cover the
// FP64 dynamic range without rounding from the scaling itself.
#include <cmath>

#include "common.hpp"
#include "model_host.hpp"

struct phy_model {
  phy_ctx* ctx = nullptr;
  int K = 4, R = 4;
  std::vector<double> pi, Q, U, lam, Ui, rates, probs;
  // device constants: U[K*K], lam[K], Ui[K*K], rates[R], pi[K], probs[R]
  double* d_consts = nullptr;
  // per-call scratch of the per-node API (grow-only)
  void* d_scratch = nullptr;
  size_t scratch_bytes = 0;
};

namespace {

constexpr int K4 = 4;

__device__ __forceinline__ int cst_U(int K) { return 0; }
__device__ __forceinline__ int cst_lam(int K) { return K * K; }
__device__ __forceinline__ int cst_Ui(int K) { return K * K + K; }
__device__ __forceinline__ int cst_rates(int K) { return 2 * K * K + K; }

template <typename real>
struct Thr;
// Rescale thresholds.  The derivative pass multiplies three rescaled
// factors (M, P_a D_a, P_b D_b), so a threshold T must keep T^3 well inside
// the format's normal range: 2^-72 (FP32, min normal 2^-126) and 2^-384
// (FP64, min normal 2^-1022).
template <>
struct Thr<float> {
  static constexpr float lo = 0x1p-24f;
};
template <>
struct Thr<double> {
  static constexpr double lo = 0x1p-128;
};

// P, P' and P'' of `nb` branches x R rates: one thread per (branch, rate, i, j).
// t_b = d_t[b] if d_t, else t0 (b == 0) / t1 (b == 1).
template <typename real>
__global__ void k_pmats(const double* __restrict__ cst, int K, int R, const double* __restrict__ d_t,
                        double t0, double t1, int nb, real* __restrict__ P, real* __restrict__ P1,
                        real* __restrict__ P2) {
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per = (int64_t)R * K * K;
  if (idx >= (int64_t)nb * per) return;
  const int b = (int)(idx / per);
  const int rem = (int)(idx % per);
  const int r = rem / (K * K), i = (rem / K) % K, j = rem % K;
  const double* U = cst + cst_U(K);
  const double* lam = cst + cst_lam(K);
  const double* Ui = cst + cst_Ui(K);
  const double rate = cst[cst_rates(K) + r];
  const double t = d_t ? d_t[b] : (b == 0 ? t0 : t1);
  const double rt = rate * t;
  double p = 0.0, p1 = 0.0, p2 = 0.0;
  for (int k = 0; k < K; ++k) {
    const double ex = exp(lam[k] * rt);
    const double uu = U[i * K + k] * Ui[k * K + j];
    p += uu * ex;
    const double rl = rate * lam[k];
    p1 += uu * rl * ex;
    p2 += uu * rl * rl * ex;
  }
  // compose_gtr rules (lib/mlmodel.c:332-341): t < 1e-10 gives the identity
  if (!(rt >= 1e-10)) p = (i == j) ? 1.0 : 0.0;
  P[idx] = (real)p;
  if (P1) {
    P1[idx] = (real)p1;
    P2[idx] = (real)p2;
  }
}

template <typename real>
__device__ __forceinline__ void matvec4(const real* __restrict__ P, const real* x, real* y) {
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    real s = P[i * 4 + 0] * x[0];
    s = fma(P[i * 4 + 1], x[1], s);
    s = fma(P[i * 4 + 2], x[2], s);
    s = fma(P[i * 4 + 3], x[3], s);
    y[i] = s;
  }
}

// x^T P (row vector times matrix)
template <typename real>
__device__ __forceinline__ void vecmat4(const real* __restrict__ P, const real* x, real* y) {
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    real s = x[0] * P[0 * 4 + j];
    s = fma(x[1], P[1 * 4 + j], s);
    s = fma(x[2], P[2 * 4 + j], s);
    s = fma(x[3], P[3 * 4 + j], s);
    y[j] = s;
  }
}

template <typename real>
__device__ __forceinline__ void tipvec(uint32_t code, real* x) {
#pragma unroll
  for (int j = 0; j < 4; ++j) x[j] = (real)((code >> j) & 1u);
}

__device__ __forceinline__ uint32_t tip_code(const uint64_t* __restrict__ tips, int64_t stride,
                                             int leaf, int64_t p) {
  const uint64_t w = tips[(int64_t)leaf * stride + (p >> 4)];
  return (uint32_t)(w >> (4 * (p & 15))) & 15u;
}

// exact power-of-two rescale of n values when their maximum is below Thr
template <typename real, int n>
__device__ __forceinline__ void rescale(real* v, int* e) {
  real mx = v[0];
#pragma unroll
  for (int k = 1; k < n; ++k) mx = fmax(mx, v[k]);
  if (mx < Thr<real>::lo && mx > (real)0) {
    int ex;
    (void)frexp(mx, &ex);
#pragma unroll
    for (int k = 0; k < n; ++k) v[k] = ldexp(v[k], -ex);
    *e += ex;
  }
}

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Small ring that prefetches the tip codes of the post-order leaf stream.
template <int BATCH>
struct TipStream {
  uint32_t pre[BATCH];
  uint32_t* ring;
  const int32_t* ids;
  int n_ids;
  const uint64_t* tips;
  int64_t stride;
  int64_t pc;
  int next, pos, bd, tid;
  __device__ __forceinline__ void issue(int start) {
#pragma unroll
    for (int j = 0; j < BATCH; ++j) {
      const int idx = start + j;
      pre[j] = idx < n_ids ? tip_code(tips, stride, ids[idx], pc) : 0u;
    }
  }
  __device__ __forceinline__ void init() {
    issue(0);
#pragma unroll
    for (int j = 0; j < BATCH; ++j) ring[j * bd + tid] = pre[j];
    issue(BATCH);
    next = 2 * BATCH;
    pos = 0;
  }
  __device__ __forceinline__ uint32_t get() {
    if (pos == BATCH) {
#pragma unroll
      for (int j = 0; j < BATCH; ++j) ring[j * bd + tid] = pre[j];
      issue(next);
      next += BATCH;
      pos = 0;
    }
    return ring[(pos++) * bd + tid];
  }
};

constexpr int kTipBatch = 8;

// ---------------------------------------------------------------------------
// Whole-tree post-order pass
// ---------------------------------------------------------------------------
template <typename real, int R, bool FUSED>
__global__ __launch_bounds__(256) void k_prune_down(
    const int4* __restrict__ prog, int n_ops, const int32_t* __restrict__ leaf_ids, int n_leaf_ids,
    const uint64_t* __restrict__ tips, int64_t stride, int64_t n_pat, int64_t pitch,
    const real* __restrict__ Pm, const double* __restrict__ pi, const double* __restrict__ probs,
    real* __restrict__ part, int32_t* __restrict__ ex, const uint32_t* __restrict__ weights,
    double* __restrict__ site_lnl, double* __restrict__ block_sums, int N, int stack_depth) {
  constexpr int RK = R * K4;
  extern __shared__ unsigned char lds_raw[];
  const int tid = threadIdx.x, bd = blockDim.x;
  const int64_t p = (int64_t)blockIdx.x * bd + tid;
  const bool act = p < n_pat;
  const int64_t pc = act ? p : 0;

  // LDS: tip ring [kTipBatch][bd] u32 | stack values [depth][RK][bd] | stack exponents [depth][bd]
  uint32_t* ring = reinterpret_cast<uint32_t*>(lds_raw);
  real* stk = reinterpret_cast<real*>(lds_raw + (size_t)kTipBatch * bd * sizeof(uint32_t));
  int32_t* stke = reinterpret_cast<int32_t*>(stk + (size_t)stack_depth * RK * bd);
  TipStream<kTipBatch> ts;
  ts.ring = ring;
  ts.ids = leaf_ids;
  ts.n_ids = n_leaf_ids;
  ts.tips = tips;
  ts.stride = stride;
  ts.pc = pc;
  ts.bd = bd;
  ts.tid = tid;
  ts.init();

  real acc[RK];
  int acc_e = 0;
  int sp = 0;
#pragma unroll
  for (int k = 0; k < RK; ++k) acc[k] = (real)0;

  for (int it = 0; it < n_ops; ++it) {
    const int4 op = prog[it];
    const int fl = op.w;
    if (FUSED && (fl & DOWN_SPILL)) {
#pragma unroll
      for (int k = 0; k < RK; ++k) stk[((size_t)sp * RK + k) * bd + tid] = acc[k];
      stke[(size_t)sp * bd + tid] = acc_e;
      ++sp;
    }
    const int sa = fl & 3, sb = (fl >> 2) & 3;
    real xa[RK], xb[RK];
    int ea = 0, eb = 0;
    // operand A
    if (sa == SRC_LEAF) {
      real t[4];
      tipvec(ts.get(), t);
#pragma unroll
      for (int r = 0; r < R; ++r) matvec4(Pm + ((size_t)op.y * R + r) * 16, t, xa + r * 4);
    } else {
      real src[RK];
      if (sa == SRC_ACC) {
#pragma unroll
        for (int k = 0; k < RK; ++k) src[k] = acc[k];
        ea = acc_e;
      } else if (FUSED) {
        --sp;
#pragma unroll
        for (int k = 0; k < RK; ++k) src[k] = stk[((size_t)sp * RK + k) * bd + tid];
        ea = stke[(size_t)sp * bd + tid];
      } else {
        const real* s = part + (size_t)(op.y - N) * RK * pitch + pc;
#pragma unroll
        for (int k = 0; k < RK; ++k) src[k] = s[(size_t)k * pitch];
        ea = ex[(size_t)(op.y - N) * pitch + pc];
      }
#pragma unroll
      for (int r = 0; r < R; ++r) matvec4(Pm + ((size_t)op.y * R + r) * 16, src + r * 4, xa + r * 4);
    }
    // operand B (a leaf or the accumulator)
    if (sb == SRC_LEAF) {
      real t[4];
      tipvec(ts.get(), t);
#pragma unroll
      for (int r = 0; r < R; ++r) matvec4(Pm + ((size_t)op.z * R + r) * 16, t, xb + r * 4);
    } else {
#pragma unroll
      for (int r = 0; r < R; ++r) matvec4(Pm + ((size_t)op.z * R + r) * 16, acc + r * 4, xb + r * 4);
      eb = acc_e;
    }
#pragma unroll
    for (int k = 0; k < RK; ++k) acc[k] = xa[k] * xb[k];
    acc_e = ea + eb;
    rescale<real, RK>(acc, &acc_e);
    if (!FUSED && act) {
      real* d = part + (size_t)(op.x - N) * RK * pitch + p;
#pragma unroll
      for (int k = 0; k < RK; ++k) d[(size_t)k * pitch] = acc[k];
      ex[(size_t)(op.x - N) * pitch + p] = acc_e;
    }
  }
  // root: ln L_p = ln sum_r prob_r sum_i pi_i L[r][i] + e ln 2
  double lk = 0.0;
#pragma unroll
  for (int r = 0; r < R; ++r) {
    double s = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) s += pi[i] * (double)acc[r * 4 + i];
    lk += probs[r] * s;
  }
  const double slnl = log(lk) + (double)acc_e * 0.6931471805599453;
  double contrib = 0.0;
  if (act) {
    if (site_lnl) site_lnl[p] = slnl;
    contrib = (double)weights[p] * slnl;
  }
  // one partial sum per wave of 64 patterns: the reduction order does not
  // depend on the block size
  const double ws = wave_sum_d(contrib);
  if ((tid & 63) == 0) block_sums[p >> 6] = ws;
}

// ---------------------------------------------------------------------------
// Pre-order derivative pass: for every branch c (parent p, sibling s)
//   A_c = M_p (*) (P_s D_s),  M_root = pi,  M_c = A_c^T P_c
//   L = sum_r prob_r A_c . (P_c D_c),  L' and L'' with P' and P''
// per-pattern g = L'/L and h = L''/L - g^2, summed per branch over the wave.
// ---------------------------------------------------------------------------
template <typename real, int R>
__global__ __launch_bounds__(256) void k_prune_deriv(
    const int4* __restrict__ prog, int n_ops, const uint64_t* __restrict__ tips, int64_t stride,
    int64_t n_pat, int64_t pitch, const real* __restrict__ Pm, const real* __restrict__ P1m,
    const real* __restrict__ P2m, const double* __restrict__ pi, const double* __restrict__ probs,
    const real* __restrict__ part, const uint32_t* __restrict__ weights,
    double* __restrict__ wave_sums, int N) {
  constexpr int RK = R * K4;
  extern __shared__ unsigned char lds_raw[];
  real* stk = reinterpret_cast<real*>(lds_raw);
  const int tid = threadIdx.x, bd = blockDim.x;
  const int64_t p = (int64_t)blockIdx.x * bd + tid;
  const bool act = p < n_pat;
  const int64_t pc = act ? p : 0;
  const double wp = act ? (double)weights[p] : 0.0;
  const int M = 2 * N - 1;
  const int64_t wave = ((int64_t)blockIdx.x * bd + tid) >> 6;
  const int lane = tid & 63;

  real accM[RK];
  int sp = 0;
#pragma unroll
  for (int k = 0; k < RK; ++k) accM[k] = (real)0;

  auto load_D = [&](int c, real* D) {
    if (c < N) {
      real t[4];
      tipvec(tip_code(tips, stride, c, pc), t);
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) D[r * 4 + i] = t[i];
    } else {
      const real* s = part + (size_t)(c - N) * RK * pitch + pc;
#pragma unroll
      for (int k = 0; k < RK; ++k) D[k] = s[(size_t)k * pitch];
    }
  };

  for (int it = 0; it < n_ops; ++it) {
    const int4 op = prog[it];
    const int fl = op.w, src = fl & 3;
    const int a = op.y, b = op.z;
    real Mv[RK];
    if (src == SRC_LEAF) {  // root
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int i = 0; i < 4; ++i) Mv[r * 4 + i] = (real)pi[i];
    } else if (src == SRC_ACC) {
#pragma unroll
      for (int k = 0; k < RK; ++k) Mv[k] = accM[k];
    } else {
      --sp;
#pragma unroll
      for (int k = 0; k < RK; ++k) Mv[k] = stk[((size_t)sp * RK + k) * bd + tid];
    }
    real Da[RK], Db[RK], ya[RK], yb[RK];
    load_D(a, Da);
    load_D(b, Db);
#pragma unroll
    for (int r = 0; r < R; ++r) {
      matvec4(Pm + ((size_t)a * R + r) * 16, Da + r * 4, ya + r * 4);
      matvec4(Pm + ((size_t)b * R + r) * 16, Db + r * 4, yb + r * 4);
    }
    double l0 = 0.0, l1a = 0.0, l2a = 0.0, l1b = 0.0, l2b = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      real x1a[4], x2a[4], x1b[4], x2b[4];
      matvec4(P1m + ((size_t)a * R + r) * 16, Da + r * 4, x1a);
      matvec4(P2m + ((size_t)a * R + r) * 16, Da + r * 4, x2a);
      matvec4(P1m + ((size_t)b * R + r) * 16, Db + r * 4, x1b);
      matvec4(P2m + ((size_t)b * R + r) * 16, Db + r * 4, x2b);
      real s0 = 0, s1a = 0, s2a = 0, s1b = 0, s2b = 0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const real ma = Mv[r * 4 + i] * yb[r * 4 + i];  // A_a
        const real mb = Mv[r * 4 + i] * ya[r * 4 + i];  // A_b
        s0 = fma(ma, ya[r * 4 + i], s0);
        s1a = fma(ma, x1a[i], s1a);
        s2a = fma(ma, x2a[i], s2a);
        s1b = fma(mb, x1b[i], s1b);
        s2b = fma(mb, x2b[i], s2b);
      }
      const double pr = probs[r];
      l0 += pr * (double)s0;
      l1a += pr * (double)s1a;
      l2a += pr * (double)s2a;
      l1b += pr * (double)s1b;
      l2b += pr * (double)s2b;
    }
    double ga = 0, ha = 0, gb = 0, hb = 0;
    if (act) {
      const double inv = 1.0 / l0;
      const double g1 = l1a * inv, g2 = l1b * inv;
      ga = wp * g1;
      ha = wp * (l2a * inv - g1 * g1);
      gb = wp * g2;
      hb = wp * (l2b * inv - g2 * g2);
    }
    ga = wave_sum_d(ga);
    ha = wave_sum_d(ha);
    gb = wave_sum_d(gb);
    hb = wave_sum_d(hb);
    if (lane == 0) {
      double* ws = wave_sums + (size_t)wave * M * 2;
      ws[2 * a] = ga;
      ws[2 * a + 1] = ha;
      ws[2 * b] = gb;
      ws[2 * b + 1] = hb;
    }
    // messages to internal children: M_c = (M (*) P_s D_s)^T P_c
    real Ma[RK], Mb[RK];
    int dummy = 0;
    if (!(fl & UP_A_LEAF)) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        real A[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) A[i] = Mv[r * 4 + i] * yb[r * 4 + i];
        vecmat4(Pm + ((size_t)a * R + r) * 16, A, Ma + r * 4);
      }
      rescale<real, RK>(Ma, &dummy);
    }
    if (!(fl & UP_B_LEAF)) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        real A[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) A[i] = Mv[r * 4 + i] * ya[r * 4 + i];
        vecmat4(Pm + ((size_t)b * R + r) * 16, A, Mb + r * 4);
      }
      rescale<real, RK>(Mb, &dummy);
    }
    if (fl & UP_PUSH) {
      const real* o = (fl & UP_CONT_B) ? Ma : Mb;
#pragma unroll
      for (int k = 0; k < RK; ++k) stk[((size_t)sp * RK + k) * bd + tid] = o[k];
      ++sp;
    }
    if (!(fl & UP_A_LEAF) || !(fl & UP_B_LEAF)) {
      const real* c = (fl & UP_CONT_B) ? Mb : Ma;
#pragma unroll
      for (int k = 0; k < RK; ++k) accM[k] = c[k];
    }
  }
}

// Neumaier sum of n doubles with one block (deterministic order).
__global__ void k_reduce_sum(const double* __restrict__ x, int64_t n, int64_t stride_x,
                             double* __restrict__ out) {
  __shared__ double ss[256], cc[256];
  double s = 0.0, c = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) {
    const double v = x[i * stride_x];
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  ss[threadIdx.x] = s;
  cc[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = 0.0, Cc = 0.0;
    for (int k = 0; k < (int)blockDim.x; ++k) {
      const double v = ss[k];
      const double t = S + v;
      if (fabs(S) >= fabs(v)) Cc += (S - t) + v;
      else Cc += (v - t) + S;
      S = t;
      Cc += cc[k];
    }
    *out = S + Cc;
  }
}

// one block per branch: d1[c], d2[c] = sums over waves; root -> 0; the second
// root child takes the first one's values (one unrooted edge).
__global__ void k_reduce_derivs(const double* __restrict__ ws, int64_t n_waves, int M, int root,
                                int ra, int rb, double* __restrict__ d1, double* __restrict__ d2) {
  const int c = blockIdx.x;
  __shared__ double s1[256], c1[256], s2[256], c2[256];
  const int src = (c == rb) ? ra : c;
  double a = 0, ac = 0, b = 0, bc = 0;
  if (c != root) {
    for (int64_t w = threadIdx.x; w < n_waves; w += blockDim.x) {
      const double v1 = ws[((size_t)w * M + src) * 2], v2 = ws[((size_t)w * M + src) * 2 + 1];
      double t = a + v1;
      if (fabs(a) >= fabs(v1)) ac += (a - t) + v1; else ac += (v1 - t) + a;
      a = t;
      t = b + v2;
      if (fabs(b) >= fabs(v2)) bc += (b - t) + v2; else bc += (v2 - t) + b;
      b = t;
    }
  }
  s1[threadIdx.x] = a;
  c1[threadIdx.x] = ac;
  s2[threadIdx.x] = b;
  c2[threadIdx.x] = bc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double A = 0, AC = 0, B = 0, BC = 0;
    for (int k = 0; k < (int)blockDim.x; ++k) {
      double v = s1[k], t = A + v;
      if (fabs(A) >= fabs(v)) AC += (A - t) + v; else AC += (v - t) + A;
      A = t;
      AC += c1[k];
      v = s2[k];
      t = B + v;
      if (fabs(B) >= fabs(v)) BC += (B - t) + v; else BC += (v - t) + B;
      B = t;
      BC += c2[k];
    }
    d1[c] = A + AC;
    d2[c] = B + BC;
  }
}

// ---------------------------------------------------------------------------
// Per-node API kernels
// ---------------------------------------------------------------------------
template <typename real, int R>
__global__ void k_lk_leaf(const uint64_t* __restrict__ tip, int64_t n_pat, int64_t pitch,
                          real* __restrict__ out, int32_t* __restrict__ eout) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pat) return;
  real t[4];
  tipvec(tip_code(tip, 0, 0, p), t);
#pragma unroll
  for (int r = 0; r < R; ++r)
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(size_t)(r * 4 + i) * pitch + p] = t[i];
  eout[p] = 0;
}

template <typename real, int R>
__global__ void k_lk_median2(const real* __restrict__ Pxy, const real* __restrict__ x,
                             const int32_t* __restrict__ exx, const real* __restrict__ y,
                             const int32_t* __restrict__ exy, int64_t n_pat, int64_t pitch,
                             real* __restrict__ out, int32_t* __restrict__ eout) {
  constexpr int RK = R * 4;
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pat) return;
  real v[RK];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    real a[4], b[4], xa[4], yb[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      a[i] = x[(size_t)(r * 4 + i) * pitch + p];
      b[i] = y[(size_t)(r * 4 + i) * pitch + p];
    }
    matvec4(Pxy + (size_t)r * 16, a, xa);
    matvec4(Pxy + (size_t)(R + r) * 16, b, yb);
#pragma unroll
    for (int i = 0; i < 4; ++i) v[r * 4 + i] = xa[i] * yb[i];
  }
  int e = exx[p] + exy[p];
  rescale<real, RK>(v, &e);
#pragma unroll
  for (int k = 0; k < RK; ++k) out[(size_t)k * pitch + p] = v[k];
  eout[p] = e;
}

template <typename real, int R>
__global__ void k_lk_root(const real* __restrict__ L, const int32_t* __restrict__ e, int64_t n_pat,
                          int64_t pitch, const double* __restrict__ pi,
                          const double* __restrict__ probs, const uint32_t* __restrict__ weights,
                          double* __restrict__ block_sums) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  if (p < n_pat) {
    double lk = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double s = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) s += pi[i] * (double)L[(size_t)(r * 4 + i) * pitch + p];
      lk += probs[r] * s;
    }
    contrib = (double)(weights ? weights[p] : 1u) * (log(lk) + (double)e[p] * 0.6931471805599453);
  }
  const double ws = wave_sum_d(contrib);
  if ((threadIdx.x & 63) == 0) block_sums[p >> 6] = ws;
}

int env_block(const char* name, int dflt) {
  const char* s = getenv(name);
  if (!s) return dflt;
  const int v = atoi(s);
  return (v == 64 || v == 128 || v == 256) ? v : dflt;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Workspace plan of phy_lk_tree
struct Plan {
  size_t off_P = 0, off_P1 = 0, off_P2 = 0, off_part = 0, off_ex = 0, off_bsum = 0, off_wsum = 0;
  size_t total = 0;
  int64_t pitch = 0, n_blocks = 0, n_waves = 0;
  int bs = 0;
  size_t esz = 8;
};

Plan make_plan(const phy_topo* t, const phy_model* m, int64_t n_pat, int32_t form, int32_t derivs) {
  Plan pl;
  const int N = t->n_leaves, M = 2 * N - 1, R = m->R, K = m->K;
  pl.esz = (form & PHY_F32) ? 4 : 8;
  const bool fused = (form & PHY_FUSED) != 0;
  pl.bs = env_block(fused ? "PHY_PRUNE_FUSED_BLOCK" : "PHY_PRUNE_BLOCK", fused ? 64 : 128);
  if (derivs) pl.bs = env_block("PHY_PRUNE_DERIV_BLOCK", 64);
  pl.pitch = ((n_pat + 255) / 256) * 256;
  if (pl.pitch == 0) pl.pitch = 256;
  pl.n_blocks = pl.pitch / pl.bs;
  pl.n_waves = pl.pitch / 64;
  size_t o = 0;
  const size_t pm = (size_t)M * R * K * K * pl.esz;
  pl.off_P = o; o = align256(o + pm);
  if (derivs) {
    pl.off_P1 = o; o = align256(o + pm);
    pl.off_P2 = o; o = align256(o + pm);
  }
  if (!fused) {
    pl.off_part = o; o = align256(o + (size_t)(N - 1) * R * K * pl.pitch * pl.esz);
    pl.off_ex = o; o = align256(o + (size_t)(N - 1) * pl.pitch * 4);
  }
  pl.off_bsum = o; o = align256(o + (size_t)pl.n_waves * 8);
  if (derivs) {
    pl.off_wsum = o; o = align256(o + (size_t)pl.n_waves * M * 2 * 8);
  }
  pl.total = o;
  return pl;
}

template <typename real, int R>
int launch_tree(phy_ctx* ctx, const phy_topo* t, const phy_model* m, bool fused, bool derivs,
                int64_t n_pat, const uint64_t* tips, int64_t stride, const uint32_t* weights,
                const double* blen, double* lnl, double* site_lnl, double* d1, double* d2,
                char* work, const Plan& pl) {
  const int N = t->n_leaves, M = 2 * N - 1, K = 4;
  const double* cst = m->d_consts;
  const double* d_pi = cst + 2 * K * K + K + R;
  const double* d_probs = d_pi + K;
  real* P = reinterpret_cast<real*>(work + pl.off_P);
  real* P1 = derivs ? reinterpret_cast<real*>(work + pl.off_P1) : nullptr;
  real* P2 = derivs ? reinterpret_cast<real*>(work + pl.off_P2) : nullptr;
  {
    const int64_t n = (int64_t)M * R * K * K;
    hipLaunchKernelGGL((k_pmats<real>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       cst, K, R, blen, 0.0, 0.0, M, P, P1, P2);
    PHY_LAUNCHED(ctx);
  }
  real* part = fused ? nullptr : reinterpret_cast<real*>(work + pl.off_part);
  int32_t* ex = fused ? nullptr : reinterpret_cast<int32_t*>(work + pl.off_ex);
  double* bsum = reinterpret_cast<double*>(work + pl.off_bsum);
  const int bs = pl.bs;
  const size_t ring = (size_t)kTipBatch * bs * sizeof(uint32_t);
  if (fused) {
    const size_t lds = ring + (size_t)t->down_depth * bs * (R * K * sizeof(real) + sizeof(int32_t));
    if (lds > 64 * 1024)
      return set_err(ctx, PHY_ERANGE, "pruning: fused LDS stack exceeds 64 KiB for this tree");
    hipLaunchKernelGGL((k_prune_down<real, R, true>), dim3((unsigned)pl.n_blocks), dim3(bs), lds,
                       ctx->stream, t->d_down_prog, N - 1, t->d_down_fifo, N, tips, stride, n_pat,
                       pl.pitch, P, d_pi, d_probs, part, ex, weights, site_lnl, bsum, N, t->down_depth);
  } else {
    hipLaunchKernelGGL((k_prune_down<real, R, false>), dim3((unsigned)pl.n_blocks), dim3(bs), ring,
                       ctx->stream, t->d_down_prog, N - 1, t->d_down_fifo, N, tips, stride, n_pat,
                       pl.pitch, P, d_pi, d_probs, part, ex, weights, site_lnl, bsum, N, 0);
  }
  PHY_LAUNCHED(ctx);
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, ctx->stream, bsum, pl.n_waves, (int64_t)1, lnl);
  PHY_LAUNCHED(ctx);
  if (derivs) {
    double* wsum = reinterpret_cast<double*>(work + pl.off_wsum);
    const size_t lds = (size_t)t->up_depth * R * K * bs * sizeof(real);
    if (lds > 64 * 1024) return set_err(ctx, PHY_ERANGE, "pruning: derivative stack too deep");
    hipLaunchKernelGGL((k_prune_deriv<real, R>), dim3((unsigned)pl.n_blocks), dim3(bs), lds,
                       ctx->stream, t->d_up_prog, N - 1, tips, stride, n_pat, pl.pitch, P, P1, P2,
                       d_pi, d_probs, part, weights, wsum, N);
    PHY_LAUNCHED(ctx);
    const int ra = t->children[2 * (t->root - N)], rb = t->children[2 * (t->root - N) + 1];
    hipLaunchKernelGGL(k_reduce_derivs, dim3((unsigned)M), dim3(256), 0, ctx->stream, wsum,
                       pl.n_waves, M, t->root, ra, rb, d1, d2);
    PHY_LAUNCHED(ctx);
  }
  return PHY_OK;
}

template <typename real>
int dispatch_tree(phy_ctx* ctx, const phy_topo* t, const phy_model* m, bool fused, bool derivs,
                  int64_t n_pat, const uint64_t* tips, int64_t stride, const uint32_t* w,
                  const double* blen, double* lnl, double* sl, double* d1, double* d2, char* work,
                  const Plan& pl) {
  switch (m->R) {
    case 1: return launch_tree<real, 1>(ctx, t, m, fused, derivs, n_pat, tips, stride, w, blen, lnl, sl, d1, d2, work, pl);
    case 2: return launch_tree<real, 2>(ctx, t, m, fused, derivs, n_pat, tips, stride, w, blen, lnl, sl, d1, d2, work, pl);
    case 4: return launch_tree<real, 4>(ctx, t, m, fused, derivs, n_pat, tips, stride, w, blen, lnl, sl, d1, d2, work, pl);
    case 8: return launch_tree<real, 8>(ctx, t, m, fused, derivs, n_pat, tips, stride, w, blen, lnl, sl, d1, d2, work, pl);
    default: return set_err(ctx, PHY_ERANGE, "pruning: n_rates must be 1, 2, 4 or 8");
  }
}

int model_scratch(phy_model* m, size_t bytes) {
  if (m->scratch_bytes >= bytes) return PHY_OK;
  if (m->d_scratch) PHY_HIP(m->ctx, hipFree(m->d_scratch));
  m->d_scratch = nullptr;
  m->scratch_bytes = 0;
  PHY_HIP(m->ctx, hipMalloc(&m->d_scratch, bytes));
  m->scratch_bytes = bytes;
  return PHY_OK;
}

}  // namespace

extern "C" {

int phy_model_create(phy_ctx* ctx, int32_t K, int32_t R, const double* pi, const double* Q,
                     const double* rates, const double* probs, phy_model** out) {
  if (!ctx || !out) return PHY_EINVAL;
  *out = nullptr;
  if (!(K == 4 || K == 20)) return set_err(ctx, PHY_ERANGE, "model: n_states must be 4 or 20");
  if (R < 1 || R > 8) return set_err(ctx, PHY_ERANGE, "model: n_rates must be 1..8");
  if (!pi || !Q || !rates || !probs) return set_err(ctx, PHY_EINVAL, "model: null argument");
  phy_model* m = new phy_model();
  m->ctx = ctx;
  m->K = K;
  m->R = R;
  m->pi.assign(pi, pi + K);
  m->Q.assign(Q, Q + K * K);
  m->rates.assign(rates, rates + R);
  m->probs.assign(probs, probs + R);
  m->U.resize(K * K);
  m->Ui.resize(K * K);
  m->lam.resize(K);
  std::string err;
  if (!phy::eigen_reversible(K, Q, m->U.data(), m->lam.data(), m->Ui.data(), &err)) {
    delete m;
    return set_err(ctx, PHY_ENUMERIC, "model: " + err);
  }
  std::vector<double> c;
  c.insert(c.end(), m->U.begin(), m->U.end());
  c.insert(c.end(), m->lam.begin(), m->lam.end());
  c.insert(c.end(), m->Ui.begin(), m->Ui.end());
  c.insert(c.end(), m->rates.begin(), m->rates.end());
  c.insert(c.end(), m->pi.begin(), m->pi.end());
  c.insert(c.end(), m->probs.begin(), m->probs.end());
  if (hipMalloc((void**)&m->d_consts, c.size() * sizeof(double)) != hipSuccess ||
      hipMemcpy(m->d_consts, c.data(), c.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    phy_model_destroy(m);
    return set_err(ctx, PHY_EDEVICE, "model: device upload failed");
  }
  *out = m;
  return PHY_OK;
}

void phy_model_destroy(phy_model* m) {
  if (!m) return;
  if (m->d_consts) (void)hipFree(m->d_consts);
  if (m->d_scratch) (void)hipFree(m->d_scratch);
  delete m;
}

int phy_model_pmatrices(phy_ctx* ctx, const phy_model* m, const double* d_t, int32_t nb, double* P,
                        double* P1, double* P2) {
  if (!ctx || !m || !d_t || !P || nb < 0) return PHY_EINVAL;
  if ((P1 == nullptr) != (P2 == nullptr)) return set_err(ctx, PHY_EINVAL, "pmatrices: P1 and P2 go together");
  if (nb == 0) return PHY_OK;
  const int64_t n = (int64_t)nb * m->R * m->K * m->K;
  hipLaunchKernelGGL((k_pmats<double>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                     m->d_consts, m->K, m->R, d_t, 0.0, 0.0, nb, P, P1, P2);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

int phy_lk_tree_work_size(const phy_topo* t, const phy_model* m, int64_t n_pat, int32_t form,
                          int32_t derivs, size_t* bytes) {
  if (!t || !m || !bytes || n_pat < 0) return PHY_EINVAL;
  if (derivs && (form & PHY_FUSED)) return PHY_EINVAL;
  *bytes = make_plan(t, m, n_pat, form, derivs).total;
  return PHY_OK;
}

int phy_lk_tree_partials(const phy_topo* t, const phy_model* m, int64_t n_pat, int32_t form,
                         int32_t derivs, void* work, int32_t v, void** d_partial, int32_t** d_exp,
                         int64_t* pitch) {
  if (!t || !m || !work || (form & PHY_FUSED)) return PHY_EINVAL;
  const int N = t->n_leaves;
  if (v < N || v > 2 * N - 2) return PHY_EINVAL;
  const Plan pl = make_plan(t, m, n_pat, form, derivs);
  char* w = static_cast<char*>(work);
  if (d_partial) *d_partial = w + pl.off_part + (size_t)(v - N) * m->R * m->K * pl.pitch * pl.esz;
  if (d_exp) *d_exp = reinterpret_cast<int32_t*>(w + pl.off_ex) + (size_t)(v - N) * pl.pitch;
  if (pitch) *pitch = pl.pitch;
  return PHY_OK;
}

int phy_lk_tree(phy_ctx* ctx, const phy_topo* t, const phy_model* m, int32_t form, int64_t n_pat,
                const uint64_t* tips, int64_t stride, const uint32_t* weights, const double* blen,
                double* lnl, double* site_lnl, double* d1, double* d2, void* work,
                size_t work_bytes) {
  if (!ctx) return PHY_EINVAL;
  if (!t || !t->d_down_prog || !m) return set_err(ctx, PHY_EINVAL, "lk_tree: topology/model missing");
  if (m->K != 4) return set_err(ctx, PHY_EINVAL, "lk_tree: this entry point is for 4-state models");
  if (form & ~(PHY_F32 | PHY_FUSED)) return set_err(ctx, PHY_EINVAL, "lk_tree: unknown form bits");
  const bool fused = (form & PHY_FUSED) != 0;
  const bool derivs = d1 != nullptr || d2 != nullptr;
  if (derivs && (!d1 || !d2)) return set_err(ctx, PHY_EINVAL, "lk_tree: d1 and d2 go together");
  if (derivs && fused) return set_err(ctx, PHY_EINVAL, "lk_tree: derivatives need stored partials (no PHY_FUSED)");
  if (n_pat < 1) return set_err(ctx, PHY_EINVAL, "lk_tree: need at least one pattern");
  if (stride < phy_words(n_pat)) return set_err(ctx, PHY_EINVAL, "lk_tree: tip stride < words");
  if (!tips || !weights || !blen || !lnl || !work) return set_err(ctx, PHY_EINVAL, "lk_tree: null argument");
  const Plan pl = make_plan(t, m, n_pat, form, derivs);
  if (work_bytes < pl.total) return set_err(ctx, PHY_EINVAL, "lk_tree: workspace too small");
  char* w = static_cast<char*>(work);
  if (form & PHY_F32)
    return dispatch_tree<float>(ctx, t, m, fused, derivs, n_pat, tips, stride, weights, blen, lnl,
                                site_lnl, d1, d2, w, pl);
  return dispatch_tree<double>(ctx, t, m, fused, derivs, n_pat, tips, stride, weights, blen, lnl,
                               site_lnl, d1, d2, w, pl);
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Per-node API
// ---------------------------------------------------------------------------
template <typename real, int R>
static int node_leaf(phy_ctx* ctx, int64_t n, const uint64_t* tip, void* out, int32_t* eout,
                     int64_t pitch) {
  hipLaunchKernelGGL((k_lk_leaf<real, R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     ctx->stream, tip, n, pitch, (real*)out, eout);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

template <typename real, int R>
static int node_median2(phy_ctx* ctx, phy_model* m, int64_t n, const void* x, const int32_t* ex,
                        double tx, const void* y, const int32_t* ey, double ty, void* out,
                        int32_t* eout, int64_t pitch) {
  const size_t need = (size_t)2 * R * 16 * sizeof(real);
  int rc = model_scratch(m, std::max<size_t>(need, 4096));
  if (rc) return rc;
  real* Pxy = (real*)m->d_scratch;
  const int64_t nm = (int64_t)2 * R * 16;
  hipLaunchKernelGGL((k_pmats<real>), dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, ctx->stream,
                     m->d_consts, 4, R, (const double*)nullptr, tx, ty, 2, Pxy, (real*)nullptr,
                     (real*)nullptr);
  PHY_LAUNCHED(ctx);
  hipLaunchKernelGGL((k_lk_median2<real, R>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     ctx->stream, Pxy, (const real*)x, ex, (const real*)y, ey, n, pitch, (real*)out,
                     eout);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

template <typename real, int R>
static int node_root(phy_ctx* ctx, phy_model* m, int64_t n, const void* L, const int32_t* e,
                     int64_t pitch, const uint32_t* w, double* lnl) {
  const int64_t nb = (n + 255) / 256;
  int rc = model_scratch(m, std::max<size_t>((size_t)nb * 4 * 8, 4096));
  if (rc) return rc;
  double* bs = (double*)m->d_scratch;  // block sums
  const int K = 4;
  const double* d_pi = m->d_consts + 2 * K * K + K + R;
  hipLaunchKernelGGL((k_lk_root<real, R>), dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                     (const real*)L, e, n, pitch, d_pi, d_pi + K, w, bs);
  PHY_LAUNCHED(ctx);
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, ctx->stream, bs, nb * 4, (int64_t)1, lnl);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

#define PHY_DISPATCH_R(fn, real, ...)                                            \
  switch (m->R) {                                                                \
    case 1: return fn<real, 1>(__VA_ARGS__);                                     \
    case 2: return fn<real, 2>(__VA_ARGS__);                                     \
    case 4: return fn<real, 4>(__VA_ARGS__);                                     \
    case 8: return fn<real, 8>(__VA_ARGS__);                                     \
    default: return set_err(ctx, PHY_ERANGE, "n_rates must be 1, 2, 4 or 8");    \
  }

extern "C" {

int phy_lk_leaf(phy_ctx* ctx, const phy_model* m, int32_t form, int64_t n, const uint64_t* tip,
                void* out, int32_t* eout, int64_t pitch) {
  if (!ctx || !m) return PHY_EINVAL;
  if (m->K != 4 || n < 1 || pitch < n || !tip || !out || !eout)
    return set_err(ctx, PHY_EINVAL, "lk_leaf: bad argument");
  if (form & PHY_F32) { PHY_DISPATCH_R(node_leaf, float, ctx, n, tip, out, eout, pitch); }
  PHY_DISPATCH_R(node_leaf, double, ctx, n, tip, out, eout, pitch);
}

int phy_lk_median2(phy_ctx* ctx, const phy_model* mc, int32_t form, int64_t n, const void* x,
                   const int32_t* ex, double tx, const void* y, const int32_t* ey, double ty,
                   void* out, int32_t* eout, int64_t pitch) {
  if (!ctx || !mc) return PHY_EINVAL;
  phy_model* m = const_cast<phy_model*>(mc);
  if (m->K != 4 || n < 1 || pitch < n || !x || !ex || !y || !ey || !out || !eout)
    return set_err(ctx, PHY_EINVAL, "lk_median2: bad argument");
  if (form & PHY_F32) { PHY_DISPATCH_R(node_median2, float, ctx, m, n, x, ex, tx, y, ey, ty, out, eout, pitch); }
  PHY_DISPATCH_R(node_median2, double, ctx, m, n, x, ex, tx, y, ey, ty, out, eout, pitch);
}

int phy_lk_root(phy_ctx* ctx, const phy_model* mc, int32_t form, int64_t n, const void* L,
                const int32_t* e, int64_t pitch, const uint32_t* w, double* lnl) {
  if (!ctx || !mc) return PHY_EINVAL;
  phy_model* m = const_cast<phy_model*>(mc);
  if (m->K != 4 || n < 1 || pitch < n || !L || !e || !lnl)
    return set_err(ctx, PHY_EINVAL, "lk_root: bad argument");
  if (form & PHY_F32) { PHY_DISPATCH_R(node_root, float, ctx, m, n, L, e, pitch, w, lnl); }
  PHY_DISPATCH_R(node_root, double, ctx, m, n, L, e, pitch, w, lnl);
}

}  // extern "C"
