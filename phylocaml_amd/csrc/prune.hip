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
#include <algorithm>
#include <climits>
#include <atomic>
#include <memory>
#include <cmath>
#include <mutex>
#include <thread>
#include <type_traits>

#include "common.hpp"
#include "stream.hpp"
#include "model_host.hpp"

// Workspace of phy_lk_tree_noalloc (per-branch P matrices and per-64-pattern
// partial sums of the fused lnL pass): a fixed device array, one copy per
// GPU, held until the process ends.  Host code never passes `scratch` itself
// (in host code the name is a stand-in, not a device address): it uses the
// address phy_noalloc_scratch_init resolves for each device.  Calls on one
// GPU share the array, so calls that overlap on one GPU race.
constexpr size_t kNoAllocScratchBytes = (size_t)8 << 20;  // 8 MiB
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

int phy_noalloc_scratch_init(int device, void** out) {
  static std::mutex mu;
  static void* addr[64] = {};
  static bool done[64] = {};
  if (!out || device < 0 || device >= 64) return PHY_EINVAL;
  std::lock_guard<std::mutex> lock(mu);
  if (!done[device]) {
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess || cur != device) return PHY_EDEVICE;
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return PHY_EDEVICE;
    }
    addr[device] = p;
    done[device] = true;
  }
  *out = addr[device];
  return PHY_OK;
}

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

// up to three branch lengths passed by value (the per-node operations)
struct HostLens {
  double t[3];
};

// P, P' and P'' of `nb` branches x R rates: one thread per (branch, rate, i, j).
// t_b = d_t[b] if d_t, else hl.t[b] (b < 3).
template <typename real>
__global__ void k_pmats(const double* __restrict__ cst, int K, int R, const double* __restrict__ d_t,
                        HostLens hl, int nb, real* __restrict__ P, real* __restrict__ P1,
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
  const double t = d_t ? d_t[b] : hl.t[b < 3 ? b : 2];
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
  if constexpr (sizeof(real) == 4) {
    // bit j of the 4-bit code to byte j (the four shifted copies of the
    // multiply do not overlap, so no carries), then one byte conversion each
    uint32_t spread = ((code & 15u) * 0x00204081u) & 0x01010101u;
    // hide the known-zero bits, so that each byte converts in one instruction
    asm("" : "+v"(spread));
    x[0] = (real)(spread & 255u);
    x[1] = (real)((spread >> 8) & 255u);
    x[2] = (real)((spread >> 16) & 255u);
    x[3] = (real)(spread >> 24);
  } else {
#pragma unroll
    for (int j = 0; j < 4; ++j) x[j] = (real)((code >> j) & 1u);
  }
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

// Cross-lane data movement through DPP (no LDS round trip).  Controls:
// quad_perm(1,0,3,2) = lane ^ 1, quad_perm(2,3,0,1) = lane ^ 2,
// row_half_mirror = lane ^ 7 within 8, row_mirror = lane ^ 15 within 16.
// exponent of an all-zero partial in a maximum over rate classes: below any
// real exponent, and far enough from INT_MIN that e - kNoExp cannot overflow
constexpr int kNoExp = INT_MIN / 2;
constexpr int DPP_XOR1 = 0xB1, DPP_XOR2 = 0x4E, DPP_HALF_MIRROR = 0x141, DPP_MIRROR = 0x140;

template <int CTRL>
__device__ __forceinline__ float dpp(float x) {
  return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), CTRL,
                                                               0xF, 0xF, false));
}
template <int CTRL>
__device__ __forceinline__ double dpp(double x) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)u, CTRL, 0xF, 0xF, false);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_update_dpp(0, (int)(uint32_t)(u >> 32), CTRL, 0xF, 0xF, false);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}
__device__ __forceinline__ double readlane(double x, int l) {
  const uint64_t u = __builtin_bit_cast(uint64_t, x);
  const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)u, l);
  const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(u >> 32), l);
  return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

// Sum over the 64 lanes, the same fixed order for every wave; the result is
// wave-uniform.
__device__ __forceinline__ double wave_sum_d(double v) {
  v += dpp<DPP_XOR1>(v);
  v += dpp<DPP_XOR2>(v);
  v += dpp<DPP_HALF_MIRROR>(v);
  v += dpp<DPP_MIRROR>(v);
  return (readlane(v, 0) + readlane(v, 16)) + (readlane(v, 32) + readlane(v, 48));
}

// XCD-aware tile order of the split form's launches.  Workgroups are dealt
// round-robin to the 8 XCDs (linear id L to XCD L % 8), so consecutive
// 64-pattern tiles would sit in 8 different L2s, each fetching the 128-byte
// lines of tip nibbles that 4 neighbouring tiles share.  Block L takes tile
// t = 32 (i / 4) + 4 x + i % 4 (x = L % 8, i = L / 8): the 4 tiles of a line
// run on one XCD, and tiles still start in about their order.  A bijection
// on [0, 32 k); the grid is padded to a multiple of 32.
__device__ __forceinline__ int64_t xcd_tile(int64_t L) {
  const int64_t x = L & 7, i = L >> 3;
  return ((i >> 2) << 5) + (x << 2) + (i & 3);
}

// Tip codes of the post-order leaf stream, 8 per packed register: the loads
// of batch k+2 are in flight while batch k is consumed by shifts (no LDS).
// The id stream is padded (phy_fifo_padded): up to 24 ids past the last one
// consumed are read.
struct TipNibbles {
  const uint32_t* tips32;  // the nibble blocks read as uint32 (8 sites each)
  int64_t stride32;        // uint32 per tip row
  int64_t col;             // uint32 column of this lane's pattern
  int shift;               // 4 * (pattern % 8)
  const int32_t* ids;
  uint32_t pre[8];
  uint32_t cur, nxt;
  int pos, next_start;

  __device__ __forceinline__ void issue(int start) {
    const int4* iv = reinterpret_cast<const int4*>(ids + start);
    const int4 a = iv[0], b = iv[1];
    const int id[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) pre[j] = tips32[(int64_t)id[j] * stride32 + col];
  }
  __device__ __forceinline__ uint32_t pack() const {
    uint32_t v = 0;
#pragma unroll
    for (int j = 0; j < 8; ++j) v |= ((pre[j] >> shift) & 15u) << (4 * j);
    return v;
  }
  __device__ __forceinline__ void init(const uint64_t* tips, int64_t stride, int64_t pc,
                                       const int32_t* id_stream) {
    tips32 = reinterpret_cast<const uint32_t*>(tips);
    stride32 = 2 * stride;
    col = pc >> 3;
    shift = 4 * (int)(pc & 7);
    ids = id_stream;
    issue(0);
    cur = pack();
    issue(8);
    nxt = pack();
    issue(16);
    next_start = 16;
    pos = 0;
  }
  __device__ __forceinline__ uint32_t get() {
    const uint32_t v = (cur >> (4 * pos)) & 15u;
    if (++pos == 8) {
      cur = nxt;
      nxt = pack();
      next_start += 8;
      issue(next_start);
      pos = 0;
    }
    return v;
  }
};
static_assert(24 + 8 <= (int)kFifoSlack, "tip stream look-ahead vs id padding");

// ---------------------------------------------------------------------------
// Whole-tree post-order pass.
//
// A block is R waves over the same 64 patterns; wave r evaluates rate class r
// (lane = pattern).  Each (pattern, rate) carries its own power-of-two
// exponent, so rate classes never interact before the root: the transition
// matrices of a wave are wave-uniform (scalar loads into SGPRs), waves run
// without barriers, and only the root combines the R classes through LDS:
//   L_p = sum_r prob_r S_r 2^(e_r - m) * 2^m,  m = max_r e_r  (exact scaling).
// Stored partial of internal node v: part[(((v - N) * R + r) * 4 + i) * pitch + p]
// (pattern-contiguous rows), exponents ex[((v - N) * R + r) * pitch + p].
// ---------------------------------------------------------------------------
template <typename real, int R>
__device__ __forceinline__ void load_mat(const real* __restrict__ P, real* m) {
#pragma unroll
  for (int k = 0; k < 16; ++k) m[k] = P[k];
}

// exact power-of-two rescale of one lane's 4 values (an all-zero vector
// keeps its values and exponent: frexp(0) gives exponent 0)
template <typename real>
__device__ __forceinline__ void rescale4(real* v, int* e) {
  const real mx = fmax(fmax(v[0], v[1]), fmax(v[2], v[3]));
  if (mx < Thr<real>::lo) {
    int ex;
    (void)frexp(mx, &ex);
#pragma unroll
    for (int k = 0; k < 4; ++k) v[k] = ldexp(v[k], -ex);
    *e += ex;
  }
}

// State of one wave of the post-order pass: accumulators of its G rate
// classes, stack, tip stream.  A block is R / G waves over the same 64
// patterns; wave w evaluates rate classes w*G .. w*G+G-1, each with its own
// per-(pattern, rate) exponent, so the classes never interact before the
// root.  With G > 1 a leaf's 0/1 expansion, the program bookkeeping and the
// stack addressing are shared by G classes.  The bottom LV levels of the
// fused stack live in registers (selected by the wave-uniform stack pointer),
// so the LDS holds LV levels less.
// MODE 0 runs the whole tree.  The split form runs it in two launches: MODE 1
// evaluates one part y (a subtree) and stores its root partial in slot y of
// the part rows; MODE 2 evaluates the top program, whose SRC_EXT operands
// read those rows (slot of A in bits 8-15, of B in bits 16-23 of the
// entry's flags).  Part row of slot k, class r:
// xt[((k * R + r) * 4 + i) * pitch + p], exponent xte[(k * R + r) * pitch + p].
// STORE (with FUSED): the operand stack stays in LDS and every partial is
// written too, so the stored form writes its partials and reads none back;
// split, its parts write their roots' partials like any other, and the top
// program's SRC_EXT operands read the cut nodes' own rows (no part rows).
template <typename real, int R, int G, bool FUSED, int LV = 0, int MODE = 0, bool STORE = false>
struct DownState {
  static_assert(R % G == 0, "rate classes per wave must divide R");
  static_assert(FUSED || LV == 0, "register stack levels belong to the fused form");
  static_assert(MODE == 0 || (FUSED && G == 1), "the split form is fused, one class per wave");
  static constexpr int kR = R, kG = G;
  static constexpr int BD = 64 * (R / G);
  static constexpr int NB = LV > 0 ? LV : 1;
  real acc[G][4];
  int acc_e[G];
  real bot[NB][G][4];
  int bot_e[NB][G];
  int sp = 0;  // wave-uniform
  real* stk;       // [depth - LV][G * 4][BD]
  int32_t* stke;   // [depth - LV][G][BD]
  int tid, r0, N;
  int64_t tile, pitch, pc, p;  // tile: the block's 64 patterns
  bool act;
  real* part;
  int32_t* ex;
  real* xt;
  int32_t* xte;
  TipNibbles ts;

  // one class's stored part partial (slot k; with STORE the cut node v's
  // own stored partial) times P
  __device__ __forceinline__ void ext_product(int k, int v, const real* P, real* y, int* e, int g) {
    const real* sv = STORE ? part + (((size_t)(v - N) * R + r0 + g) * 4) * pitch + pc
                           : xt + (((size_t)k * R + r0 + g) * 4) * pitch + pc;
    real x[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) x[i] = sv[(size_t)i * pitch];
    *e = STORE ? ex[((size_t)(v - N) * R + r0 + g) * pitch + pc] : xte[((size_t)k * R + r0 + g) * pitch + pc];
    matvec4(P, x, y);
  }
  // store the accumulator (the part's root partial) in slot k
  __device__ __forceinline__ void ext_store(int k) const {
#pragma unroll
    for (int g = 0; g < G; ++g) {
      real* d = xt + (((size_t)k * R + r0 + g) * 4) * pitch + p;
#pragma unroll
      for (int i = 0; i < 4; ++i) d[(size_t)i * pitch] = acc[g][i];
      xte[((size_t)k * R + r0 + g) * pitch + p] = acc_e[g];
    }
  }

  __device__ __forceinline__ void spill_if(int fl) {
    if (FUSED && (fl & DOWN_SPILL)) {
      if (sp < LV) {
        // value selects on the uniform level, so the levels stay registers
        static_for<0, LV>([&](auto l) {
          const bool hit = sp == l;
#pragma unroll
          for (int g = 0; g < G; ++g) {
#pragma unroll
            for (int k = 0; k < 4; ++k) bot[l][g][k] = hit ? acc[g][k] : bot[l][g][k];
            bot_e[l][g] = hit ? acc_e[g] : bot_e[l][g];
          }
        });
      } else {
        const int lv = sp - LV;
#pragma unroll
        for (int g = 0; g < G; ++g) {
#pragma unroll
          for (int k = 0; k < 4; ++k) stk[(lv * G * 4 + g * 4 + k) * BD + tid] = acc[g][k];
          stke[(lv * G + g) * BD + tid] = acc_e[g];
        }
      }
      ++sp;
    }
  }
  // y[g] = P[g] x[g] for operand A (leaf, accumulator, stack, stored child a
  // or part row ka)
  __device__ __forceinline__ void product_a(int sa, int a, int ka, const real (*P)[16], real (*y)[4],
                                            int* e) {
    if (MODE == 2 && sa == SRC_EXT) {
      ext_product(ka, a, P[0], y[0], &e[0], 0);
    } else if (sa == SRC_LEAF) {
      real x[4];
      tipvec(ts.get(), x);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        matvec4(P[g], x, y[g]);
        e[g] = 0;
      }
    } else if (sa == SRC_ACC) {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        matvec4(P[g], acc[g], y[g]);
        e[g] = acc_e[g];
      }
    } else if (FUSED) {
      --sp;
      if (sp < LV) {
#pragma unroll
        for (int g = 0; g < G; ++g) {
          real x[4];
          int xe = bot_e[0][g];
#pragma unroll
          for (int k = 0; k < 4; ++k) x[k] = bot[0][g][k];
          static_for<1, LV>([&](auto l) {
            const bool hit = sp == l;
#pragma unroll
            for (int k = 0; k < 4; ++k) x[k] = hit ? bot[l][g][k] : x[k];
            xe = hit ? bot_e[l][g] : xe;
          });
          matvec4(P[g], x, y[g]);
          e[g] = xe;
        }
      } else {
        const int lv = sp - LV;
#pragma unroll
        for (int g = 0; g < G; ++g) {
          real x[4];
#pragma unroll
          for (int k = 0; k < 4; ++k) x[k] = stk[(lv * G * 4 + g * 4 + k) * BD + tid];
          e[g] = stke[(lv * G + g) * BD + tid];
          matvec4(P[g], x, y[g]);
        }
      }
    } else {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        const real* sv = part + (((size_t)(a - N) * R + r0 + g) * 4) * pitch + pc;
        real x[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) x[k] = sv[(size_t)k * pitch];
        e[g] = ex[((size_t)(a - N) * R + r0 + g) * pitch + pc];
        matvec4(P[g], x, y[g]);
      }
    }
  }
  // operand B (leaf, accumulator or the part root b, row kb)
  __device__ __forceinline__ void product_b(int sb, int b, int kb, const real (*P)[16], real (*y)[4],
                                            int* e) {
    if (MODE == 2 && sb == SRC_EXT) {
      ext_product(kb, b, P[0], y[0], &e[0], 0);
    } else if (sb == SRC_LEAF) {
      real x[4];
      tipvec(ts.get(), x);
#pragma unroll
      for (int g = 0; g < G; ++g) {
        matvec4(P[g], x, y[g]);
        e[g] = 0;
      }
    } else {
#pragma unroll
      for (int g = 0; g < G; ++g) {
        matvec4(P[g], acc[g], y[g]);
        e[g] = acc_e[g];
      }
    }
  }
  __device__ __forceinline__ void combine(int v, const real (*xa)[4], const real (*xb)[4],
                                          const int* ea, const int* eb) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
#pragma unroll
      for (int k = 0; k < 4; ++k) acc[g][k] = xa[g][k] * xb[g][k];
      acc_e[g] = ea[g] + eb[g];
      rescale4(acc[g], &acc_e[g]);
      if ((!FUSED || STORE) && act) {
        real* d = part + (((size_t)(v - N) * R + r0 + g) * 4) * pitch + p;
#pragma unroll
        for (int k = 0; k < 4; ++k) d[(size_t)k * pitch] = acc[g][k];
        ex[((size_t)(v - N) * R + r0 + g) * pitch + p] = acc_e[g];
      }
    }
  }
};

// root: S_r = sum_i pi_i L[r][i]; combine the R classes in wave 0
template <class State>
__device__ __forceinline__ void down_root(const State& st, unsigned char* lds_raw,
                                          const double* __restrict__ pi,
                                          const double* __restrict__ probs,
                                          const uint32_t* __restrict__ weights,
                                          double* __restrict__ site_lnl, double* __restrict__ wave_sums) {
  constexpr int R = State::kR, G = State::kG;
  const int lane = st.tid & 63;
  double* rootS = reinterpret_cast<double*>(lds_raw);
  int* rootE = reinterpret_cast<int*>(rootS + R * 64);
  double S[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    S[g] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) S[g] += pi[i] * (double)st.acc[g][i];
  }
  __syncthreads();  // every wave is done with its stack rows
#pragma unroll
  for (int g = 0; g < G; ++g) {
    rootS[(st.r0 + g) * 64 + lane] = S[g];
    rootE[(st.r0 + g) * 64 + lane] = st.acc_e[g];
  }
  __syncthreads();
  if (st.r0 == 0) {
    // m over the classes with a non-zero sum only: a class whose partial is
    // all zero (a zero rate, P = I, under a non-constant pattern) never
    // rescales, so its exponent says nothing about the pattern's scale
    int m = INT_MIN;
#pragma unroll
    for (int q = 0; q < R; ++q)
      if (rootS[q * 64 + lane] > 0.0) m = max(m, rootE[q * 64 + lane]);
    if (m == INT_MIN) m = 0;
    double L = 0.0;
#pragma unroll
    for (int q = 0; q < R; ++q) L += probs[q] * ldexp(rootS[q * 64 + lane], rootE[q * 64 + lane] - m);
    const double slnl = log(L) + (double)m * 0.6931471805599453;
    double contrib = 0.0;
    if (st.act) {
      if (site_lnl) site_lnl[st.p] = slnl;
      contrib = (double)weights[st.p] * slnl;
    }
    // one partial sum per 64 patterns, independent of the launch shape
    const double ws = wave_sum_d(contrib);
    if (lane == 0) wave_sums[st.tile] = ws;
  }
}

template <typename real, int R, int G, bool FUSED, int LV, int MODE, bool STORE>
__device__ __forceinline__ void down_init(DownState<real, R, G, FUSED, LV, MODE, STORE>& st, unsigned char* lds_raw,
                                          const int32_t* leaf_ids, const uint64_t* tips,
                                          int64_t stride, int64_t n_pat, int64_t pitch, real* part,
                                          int32_t* ex, int N, int stack_depth, int64_t tile) {
  st.tid = threadIdx.x;
  st.r0 = __builtin_amdgcn_readfirstlane((st.tid >> 6) * G);  // first rate class of this wave
  st.tile = tile;
  st.p = tile * 64 + (st.tid & 63);
  st.act = st.p < n_pat;
  st.pc = st.act ? st.p : 0;
  st.pitch = pitch;
  st.N = N;
  st.part = part;
  st.ex = ex;
  st.xt = nullptr;
  st.xte = nullptr;
#pragma unroll
  for (int g = 0; g < G; ++g) {
    st.acc_e[g] = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) st.acc[g][k] = (real)0;
  }
  // LDS: (fused) stack [depth - LV][G*4][BD] real + [depth - LV][G][BD]
  // int; at the root, when the stack is empty, the same bytes hold the rate
  // exchange [R][64] double S + [R][64] int e
  const int rows = FUSED ? max(stack_depth - LV, 0) : 0;
  st.stk = reinterpret_cast<real*>(lds_raw);
  st.stke = reinterpret_cast<int32_t*>(st.stk + (size_t)rows * G * 4 * st.BD);
  st.ts.init(tips, stride, st.pc, leaf_ids);
}

// The wave's 4x4 P of both children for each of its G classes are
// wave-uniform (SGPRs): 16 FMAs with an SGPR operand per product.
template <class State, typename real>
__device__ __forceinline__ void down_program(State& st, const int4* __restrict__ prog, int n_ops,
                                             const real* __restrict__ Pm) {
  constexpr int R = State::kR, G = State::kG;
  const int r0 = st.r0;
  for_each_entry(prog, n_ops, [&](int, const int4 op) {
    real PA[G][16], PB[G][16];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      load_mat<real, R>(Pm + ((size_t)op.y * R + r0 + g) * 16, PA[g]);
      load_mat<real, R>(Pm + ((size_t)op.z * R + r0 + g) * 16, PB[g]);
    }
    const int fl = op.w;
    st.spill_if(fl);
    real xa[G][4], xb[G][4];
    int ea[G], eb[G];
    st.product_a(fl & 3, op.y, (fl >> 8) & 255, PA, xa, ea);
    st.product_b((fl >> 2) & 3, op.z, (fl >> 16) & 255, PB, xb, eb);
    st.combine(op.x, xa, xb, ea, eb);
  });
}

template <typename real, int R, int G, bool FUSED, int LV, int MODE = 0, bool STORE = false>
__device__ __forceinline__ void prune_down_body(
    const int4* __restrict__ prog, int n_ops, const int32_t* __restrict__ leaf_ids,
    const uint64_t* __restrict__ tips, int64_t stride, int64_t n_pat, int64_t pitch,
    const real* __restrict__ Pm, const double* __restrict__ pi, const double* __restrict__ probs,
    real* __restrict__ part, int32_t* __restrict__ ex, const uint32_t* __restrict__ weights,
    double* __restrict__ site_lnl, double* __restrict__ wave_sums, int N, int stack_depth,
    real* __restrict__ xt = nullptr, int32_t* __restrict__ xte = nullptr) {
  extern __shared__ unsigned char lds_raw[];
  int64_t tile = blockIdx.x;
  if constexpr (MODE == 2) {
    tile = xcd_tile(blockIdx.x);
    if (tile >= pitch / 64) return;  // padding of the grid (whole block)
  }
  DownState<real, R, G, FUSED, LV, MODE, STORE> st;
  down_init(st, lds_raw, leaf_ids, tips, stride, n_pat, pitch, part, ex, N, stack_depth, tile);
  st.xt = xt;
  st.xte = xte;
  down_program(st, prog, n_ops, Pm);
  down_root(st, lds_raw, pi, probs, weights, site_lnl, wave_sums);
}

// Whole tree (MODE 0) or the top program of the split form (MODE 2).
template <typename real, int R, int G, bool FUSED, int LV, int MODE = 0, bool STORE = false>
__global__ __launch_bounds__(512) void k_prune_down(
    const int4* __restrict__ prog, int n_ops, const int32_t* __restrict__ leaf_ids,
    const uint64_t* __restrict__ tips, int64_t stride, int64_t n_pat, int64_t pitch,
    const real* __restrict__ Pm, const double* __restrict__ pi, const double* __restrict__ probs,
    real* __restrict__ part, int32_t* __restrict__ ex, const uint32_t* __restrict__ weights,
    double* __restrict__ site_lnl, double* __restrict__ wave_sums, int N, int stack_depth,
    real* __restrict__ xt, int32_t* __restrict__ xte) {
  prune_down_body<real, R, G, FUSED, LV, MODE, STORE>(prog, n_ops, leaf_ids, tips, stride, n_pat, pitch,
                                                  Pm, pi, probs, part, ex, weights, site_lnl,
                                                  wave_sums, N, stack_depth, xt, xte);
}

// The parts of the split form side by side: tile t = y * (pitch / 64) + x
// runs part y over patterns 64x .. 64x+63 and stores the part's root partial
// in slot y (tiles in XCD-aware order over a 1-D grid padded to a multiple
// of 32).  The part programs and leaf streams sit at the offsets of meta
// (prog_off | n_ops | fifo_off, n_parts each).  With STORE (the stored
// form) every partial of the part goes to its node's rows instead, the
// part root's among them, and the top program reads it there.
template <typename real, int R, int LV, bool STORE = false>
__global__ __launch_bounds__(512) void k_prune_part(
    const int4* __restrict__ prog, const int32_t* __restrict__ meta, int n_parts,
    const int32_t* __restrict__ fifo, const uint64_t* __restrict__ tips, int64_t stride,
    int64_t n_pat, int64_t pitch, const real* __restrict__ Pm, real* __restrict__ xt,
    int32_t* __restrict__ xte, real* __restrict__ part, int32_t* __restrict__ ex, int N,
    int stack_depth) {
  extern __shared__ unsigned char lds_raw[];
  const int64_t nb = pitch / 64;
  const int64_t t = xcd_tile(blockIdx.x);
  if (t >= nb * n_parts) return;  // padding of the grid (whole block)
  const int y = (int)(t / nb);
  const int4* pg = prog + meta[y];
  const int n_ops = meta[n_parts + y];
  const int32_t* li = fifo + meta[2 * n_parts + y];
  DownState<real, R, 1, true, LV, 1, STORE> st;
  down_init(st, lds_raw, li, tips, stride, n_pat, pitch, part, ex, N, stack_depth, t - (int64_t)y * nb);
  st.xt = xt;
  st.xte = xte;
  down_program(st, pg, n_ops, Pm);
  if (!STORE) st.ext_store(y);
}

// Many candidate topologies over the same tips and weights (fused form):
// blockIdx.y selects candidate c, whose program, leaf order, P matrices and
// per-wave sums sit at fixed strides.
template <typename real, int R, int G>
__global__ __launch_bounds__(512) void k_prune_down_batch(
    const int4* __restrict__ prog, const int32_t* __restrict__ leaf_ids, int N,
    const uint64_t* __restrict__ tips, int64_t stride, int64_t n_pat, int64_t pitch,
    const real* __restrict__ Pm, const double* __restrict__ pi, const double* __restrict__ probs,
    const uint32_t* __restrict__ weights, double* __restrict__ wave_sums, int64_t n_waves,
    int stack_depth) {
  const int64_t c = blockIdx.y;
  const int M = 2 * N - 1;
  const int4* pg = prog + c * (N - 1);
  const int32_t* li = leaf_ids + c * (int64_t)phy_fifo_padded(N);
  const real* P = Pm + c * (int64_t)M * R * 16;
  prune_down_body<real, R, G, true, 0>(pg, N - 1, li, tips, stride, n_pat, pitch, P, pi, probs, nullptr,
                                    nullptr, weights, nullptr, wave_sums + c * n_waves, N, stack_depth);
}

// Helpers of the derivative pass, whose threads are (pattern, rate) quads.
// sum over the R lanes of one pattern (lanes p*R .. p*R+R-1); every lane of
// the group gets the same bits (a + b == b + a exactly)
template <int R>
__device__ __forceinline__ double rate_sum(double x) {
  if (R >= 2) x += dpp<DPP_XOR1>(x);
  if (R >= 4) x += dpp<DPP_XOR2>(x);
  if (R >= 8) x += dpp<DPP_HALF_MIRROR>(x);
  return x;
}

// Sum over the lanes of a 16-lane row that are congruent mod S (S = 1, 2, 4,
// 8): each such lane receives the sum, via DPP row rotations.
template <int S>
__device__ __forceinline__ double row_sum_stride(double v) {
  if (S <= 8) v += dpp<0x128>(v);  // row_ror:8
  if (S <= 4) v += dpp<0x124>(v);  // row_ror:4
  if (S <= 2) v += dpp<0x122>(v);  // row_ror:2
  if (S <= 1) v += dpp<0x121>(v);  // row_ror:1
  return v;
}

// ---------------------------------------------------------------------------
// Pre-order derivative pass: for every branch c (parent p, sibling s)
//   A_c = M_p (*) (P_s D_s),  M_root = pi,  M_c = A_c^T P_c
//   L = sum_r prob_r A_c . (P_c D_c),  L' and L'' with P' and P''
// per-pattern g = L'/L and h = L''/L - g^2, summed per branch over the wave.
//
// Everything runs in the eigenbasis of Q = U diag(lam) Ui: with
// e_k = exp(lam_k rate_r t_c) = 1 + m_k, u = Ui D_c, w = e (*) u and
// g = U^T A_c,
//   P_c D_c = U w = D_c + U (m (*) u),
//   A_c . (P^(n) D_c) = sum_k g_k w_k (lam_k rate_r)^n,
//   M_c = Ui^T (e (*) g) = A_c + Ui^T (m (*) g)
// (U Ui = I; the forms with m keep full relative precision on short
// branches),
// so a lane needs only the 4 factors e of each child branch per node; U and
// Ui are wave-uniform.  Threads are (pattern, rate) quads as in the node
// kernels: lane p*R + r, rate sums by DPP, one exponent per pattern.
// ---------------------------------------------------------------------------

// exp(lam_k rate_r t_b) - 1 of nb branches (0 where rate_r t_b < 1e-10, the
// identity rule of lib/mlmodel.c:332-341), plus real copies of U, Ui and
// lam_k rate_r: tab = [U 16 | Ui 16 | lr R*4].  The factors are kept as
// e - 1 (expm1, full relative precision) so that P D = D + U ((e - 1) (*)
// Ui D) keeps the relative precision of P's off-diagonal entries (~ Q t) on
// short branches, which U (e (*) Ui D) with e ~ 1 would lose to
// cancellation (a relative error of eps / (|lambda| t)).
template <typename real>
__global__ void k_eigen_exp(const double* __restrict__ cst, int R, const double* __restrict__ d_t,
                            int nb, real* __restrict__ E, real* __restrict__ tab) {
  constexpr int K = 4;
  const int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const double* lam = cst + cst_lam(K);
  if (idx < 32 + 4 * R) {
    double v;
    if (idx < 16) v = cst[cst_U(K) + idx];
    else if (idx < 32) v = cst[cst_Ui(K) + idx - 16];
    else v = lam[(idx - 32) % 4] * cst[cst_rates(K) + (idx - 32) / 4];
    tab[idx] = (real)v;
  }
  if (idx >= (int64_t)nb * R * K) return;
  const int b = (int)(idx / (R * K)), r = (int)(idx / K) % R, k = (int)(idx % K);
  const double rt = cst[cst_rates(K) + r] * d_t[b];
  E[idx] = (real)((rt >= 1e-10) ? expm1(lam[k] * rt) : 0.0);
}

template <typename real>
__device__ __forceinline__ void matvec_u(const real* __restrict__ A, const real* x, real* y) {
#pragma unroll
  for (int i = 0; i < 4; ++i)
    y[i] = fma(A[4 * i], x[0], fma(A[4 * i + 1], x[1], fma(A[4 * i + 2], x[2], A[4 * i + 3] * x[3])));
}
template <typename real>
__device__ __forceinline__ void vecmat_u(const real* __restrict__ A, const real* x, real* y) {
#pragma unroll
  for (int j = 0; j < 4; ++j)
    y[j] = fma(A[j], x[0], fma(A[4 + j], x[1], fma(A[8 + j], x[2], A[12 + j] * x[3])));
}

template <typename real, int R>
__global__ __launch_bounds__(256, sizeof(real) == 8 ? 3 : 1) void k_prune_deriv(
    const int4* __restrict__ prog, int n_ops, const uint64_t* __restrict__ tips, int64_t stride,
    int64_t n_pat, int64_t pitch, const real* __restrict__ tab, const real* __restrict__ E,
    const double* __restrict__ pi, const double* __restrict__ probs, const real* __restrict__ part,
    const int32_t* __restrict__ ex, const uint32_t* __restrict__ weights,
    double* __restrict__ wsum, int64_t n_slots, int N, int up_depth) {
  constexpr int QPL = R >= 4 ? 1 : 4 / R;  // derivative quantities per lane
  extern __shared__ unsigned char lds_raw[];
  real* stk = reinterpret_cast<real*>(lds_raw);  // [depth][4][bd]
  const int tid = threadIdx.x, bd = blockDim.x;
  int32_t* stke = reinterpret_cast<int32_t*>(stk + (size_t)up_depth * 4 * bd);  // [depth][bd]
  const int r = tid % R;
  const int64_t p = (int64_t)blockIdx.x * (bd / R) + tid / R;
  const bool act = p < n_pat;
  const int64_t pc = act ? p : 0;
  const double wp = act ? (double)weights[p] : 0.0;
  const int64_t wave = ((int64_t)blockIdx.x * bd + tid) >> 6;
  const int lane = tid & 63;

  // FP32 fetches an entry's eigen factors one node ahead with its partials
  // and keeps a table of (lambda_k r)^2; FP64 reads the factors (a small
  // table that stays in cache) when the entry starts and forms the squares
  // as two factors, which keeps it within 168 registers (three waves per
  // SIMD instead of two: 7.2 against 8.1 ms at 512 x 200k; FP32 is faster
  // the first way, 4.55 against 4.68 ms)
  constexpr bool kAhead = sizeof(real) == 4;
  real U[16], Ui[16], l1[4], l2[4];
#pragma unroll
  for (int k = 0; k < 16; ++k) {
    U[k] = tab[k];
    Ui[k] = tab[16 + k];
  }
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    l1[k] = tab[32 + 4 * r + k];
    l2[k] = l1[k] * l1[k];
  }
  const double pr = probs[r];

  auto efac = [&](int c, real* ev) {
    const real* es = E + ((size_t)c * R + r) * 4;
#pragma unroll
    for (int k = 0; k < 4; ++k) ev[k] = es[k];
  };
  // children partials (4 values + exponent each) of an entry, fetched raw
  // one node ahead, with the eigen factors in FP32
  auto fetch = [&](int c, real* Dv, int* e, real* ev) {
    if (c < N) {
      tipvec(tip_code(tips, stride, c, pc), Dv);
      *e = 0;
    } else {
      const real* sv = part + (((size_t)(c - N) * R + r) * 4) * pitch + pc;
#pragma unroll
      for (int k = 0; k < 4; ++k) Dv[k] = sv[(size_t)k * pitch];
      *e = ex[((size_t)(c - N) * R + r) * pitch + pc];
    }
    if constexpr (kAhead) efac(c, ev);
  };
  // Every factor keeps the power-of-two exponent of its own rate class: the
  // children partials the exponents of the post-order pass, the messages M
  // their own (M_root = pi: 0).  A class's terms s = M . ((P_a D_a) (*)
  // (P_b D_b)) are products of three rescaled factors, so they stay normal
  // in the working precision; the classes of a pattern meet only in the FP64
  // rate sum, aligned to the largest exponent among the classes whose term
  // is non-zero.  (An all-zero class -- a zero rate under a non-constant
  // pattern -- never rescales, and a class that dominates one child can be
  // zero in the product, so no alignment of a single factor is safe.)

  real Dc[8], Ec[8];  // children partials and eigen factors of the current entry
  int Xc[2];
  {
    const int4 o = prog[0];
    fetch(o.y, Dc, &Xc[0], Ec);
    fetch(o.z, Dc + 4, &Xc[1], Ec + 4);
  }
  real accM[4] = {(real)0, (real)0, (real)0, (real)0};
  int accE = 0;
  int sp = 0;
  for_each_entry_ahead(prog, n_ops, [&](int i, const int4 op, const int4 on) {
    real Dn[8], En[8];
    int Xn[2] = {0, 0};
    const bool more = i + 1 < n_ops;
    if (more) {
      fetch(on.y, Dn, &Xn[0], En);
      fetch(on.z, Dn + 4, &Xn[1], En + 4);
    }
    const int fl = op.w, src = fl & 3;
    const int a = op.y, b = op.z;
    if constexpr (!kAhead) {
      efac(a, Ec);
      efac(b, Ec + 4);
    }
    real Mv[4];
    int eM;
    if (src == SRC_LEAF) {
#pragma unroll
      for (int k = 0; k < 4; ++k) Mv[k] = (real)pi[k];
      eM = 0;
    } else if (src == SRC_ACC) {
#pragma unroll
      for (int k = 0; k < 4; ++k) Mv[k] = accM[k];
      eM = accE;
    } else {
      --sp;
#pragma unroll
      for (int k = 0; k < 4; ++k) Mv[k] = stk[((size_t)sp * 4 + k) * bd + tid];
      eM = stke[(size_t)sp * bd + tid];
    }
    const real* Da = Dc;
    const real* Db = Dc + 4;
    const real* ea = Ec;
    const real* eb = Ec + 4;
    // ea / eb hold e - 1 (see k_eigen_exp)
    real ua[4], ub[4], wa[4], wb[4], ya[4], yb[4], Aa[4], Ab[4], ga[4], gb[4];
    matvec_u(Ui, Da, ua);
    matvec_u(Ui, Db, ub);
    {
      real ma[4], mb[4], ta[4], tb[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        ma[k] = ea[k] * ua[k];
        mb[k] = eb[k] * ub[k];
        wa[k] = ua[k] + ma[k];
        wb[k] = ub[k] + mb[k];
      }
      matvec_u(U, ma, ta);
      matvec_u(U, mb, tb);
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        ya[k] = Da[k] + ta[k];
        yb[k] = Db[k] + tb[k];
      }
    }
    real s0 = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      Aa[k] = Mv[k] * yb[k];
      Ab[k] = Mv[k] * ya[k];
      s0 = fma(Aa[k], ya[k], s0);
    }
    vecmat_u(U, Aa, ga);
    vecmat_u(U, Ab, gb);
    real s1a = 0, s2a = 0, s1b = 0, s2b = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const real qa = ga[k] * wa[k], qb = gb[k] * wb[k];
      s1a = fma(qa, l1[k], s1a);
      s2a = kAhead ? fma(qa, l2[k], s2a) : fma(qa * l1[k], l1[k], s2a);
      s1b = fma(qb, l1[k], s1b);
      s2b = kAhead ? fma(qb, l2[k], s2b) : fma(qb * l1[k], l1[k], s2b);
    }
    // class weight prob_r 2^(E_r - m), E_r = eM + e_a + e_b, m the largest E
    // over the pattern's classes with a non-zero term (0 for the others)
    const int E = eM + Xc[0] + Xc[1];
    const bool nz = s0 != (real)0;
    int m = nz ? E : kNoExp;
    if (R >= 2) m = max(m, __builtin_amdgcn_update_dpp(0, m, DPP_XOR1, 0xF, 0xF, false));
    if (R >= 4) m = max(m, __builtin_amdgcn_update_dpp(0, m, DPP_XOR2, 0xF, 0xF, false));
    if (R >= 8) m = max(m, __builtin_amdgcn_update_dpp(0, m, DPP_HALF_MIRROR, 0xF, 0xF, false));
    const double sc = nz ? ldexp(pr, E - m) : 0.0;
    const double L0 = rate_sum<R>(sc * (double)s0);
    const double L1a = rate_sum<R>(sc * (double)s1a), L2a = rate_sum<R>(sc * (double)s2a);
    const double L1b = rate_sum<R>(sc * (double)s1b), L2b = rate_sum<R>(sc * (double)s2b);
    // The four per-pattern quantities q = (g_a, h_a, g_b, h_b) are spread
    // over the R lanes of the pattern (lane r holds q = r * QPL + j), summed
    // over the lanes of each 16-lane row that hold the same q (DPP row
    // rotations), then over the four rows, and written to the wave's slot
    // (one slot per wave keeps the sums the reduction reads 4x smaller).
    {
      const double inv = 1.0 / L0;
      const double ga1 = L1a * inv, gb1 = L1b * inv;
      double qv[QPL];
#pragma unroll
      for (int j = 0; j < QPL; ++j) {
        const int q = r * QPL + j;
        double v = 0.0;
        if (act && q < 4) {
          const double g = (q < 2) ? ga1 : gb1;
          const double l2 = (q < 2) ? L2a : L2b;
          v = wp * ((q & 1) ? (l2 * inv - g * g) : g);
        }
        v = row_sum_stride<R>(v);
        v += __shfl_xor(v, 16, 64);
        qv[j] = v + __shfl_xor(v, 32, 64);
      }
      if (lane < R) {
        const int64_t slot = wave;
#pragma unroll
        for (int j = 0; j < QPL; ++j) {
          const int q = r * QPL + j;
          if (q < 4) {
            const int edge = (q < 2) ? a : b;
            wsum[((size_t)edge * 2 + (q & 1)) * n_slots + slot] = qv[j];
          }
        }
      }
    }
    // messages to internal children: M_c = Ui^T (e_c (*) g_c), with the
    // exponent of A_c (the parent's message times the sibling's partial)
    real Ma[4], Mb[4];
    int eMa = eM + Xc[1], eMb = eM + Xc[0];
    if (!(fl & UP_A_LEAF)) {
      real t[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) t[k] = ea[k] * ga[k];
      vecmat_u(Ui, t, Ma);
#pragma unroll
      for (int k = 0; k < 4; ++k) Ma[k] += Aa[k];
      rescale4(Ma, &eMa);
    }
    if (!(fl & UP_B_LEAF)) {
      real t[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) t[k] = eb[k] * gb[k];
      vecmat_u(Ui, t, Mb);
#pragma unroll
      for (int k = 0; k < 4; ++k) Mb[k] += Ab[k];
      rescale4(Mb, &eMb);
    }
    if (fl & UP_PUSH) {
      const bool cb = fl & UP_CONT_B;
      const real* o = cb ? Ma : Mb;
#pragma unroll
      for (int k = 0; k < 4; ++k) stk[((size_t)sp * 4 + k) * bd + tid] = o[k];
      stke[(size_t)sp * bd + tid] = cb ? eMa : eMb;
      ++sp;
    }
    if (!(fl & UP_A_LEAF) || !(fl & UP_B_LEAF)) {
      const bool cb = fl & UP_CONT_B;
      const real* c = cb ? Mb : Ma;
#pragma unroll
      for (int k = 0; k < 4; ++k) accM[k] = c[k];
      accE = cb ? eMb : eMa;
    }
    if (more) {
#pragma unroll
      for (int k = 0; k < 8; ++k) {
        Dc[k] = Dn[k];
        if constexpr (kAhead) Ec[k] = En[k];
      }
      Xc[0] = Xn[0];
      Xc[1] = Xn[1];
    }
  });
}

// Neumaier sum of n doubles with one block (deterministic order).
// Compensated sums as (s, c) pairs: Neumaier within a thread, then pairs
// combined by an error-free two-sum in a fixed butterfly over the wave and a
// fixed order over the waves (deterministic, launch-shape independent).
struct CSum {
  double s = 0.0, c = 0.0;
  __device__ __forceinline__ void add(double v) {
    const double t = s + v;
    if (fabs(s) >= fabs(v)) c += (s - t) + v;
    else c += (v - t) + s;
    s = t;
  }
  __device__ __forceinline__ void merge(double s2, double c2) {
    const double t = s + s2, bb = t - s;
    c += c2 + ((s - (t - bb)) + (s2 - bb));
    s = t;
  }
};

// block-wide compensated sum (blockDim.x a multiple of 64, at most 1024);
// thread 0 gets the total
__device__ __forceinline__ double block_csum(CSum a) {
  __shared__ double ws[16], wc[16];
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) a.merge(__shfl_xor(a.s, off, 64), __shfl_xor(a.c, off, 64));
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    ws[w] = a.s;
    wc[w] = a.c;
  }
  __syncthreads();
  double out = 0.0;
  if (threadIdx.x == 0) {
    CSum t;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t.merge(ws[k], wc[k]);
    out = t.s + t.c;
  }
  return out;
}

__global__ void k_reduce_sum(const double* __restrict__ x, int64_t n, int64_t stride_x,
                             double* __restrict__ out) {
  CSum a;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) a.add(x[i * stride_x]);
  const double t = block_csum(a);
  if (threadIdx.x == 0) *out = t;
}

// one block per row: out[row] = compensated sum of x[row * n .. row * n + n)
__global__ void k_reduce_rows(const double* __restrict__ x, int64_t n, double* __restrict__ out) {
  const double* r = x + (int64_t)blockIdx.x * n;
  CSum a;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) a.add(r[i]);
  const double t = block_csum(a);
  if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// one block per branch: d1[c], d2[c] = sums over waves; root -> 0; the second
// root child takes the first one's values (one unrooted edge).
__global__ void k_reduce_derivs(const double* __restrict__ ws, int64_t n_waves, int M, int root,
                                int ra, int rb, double* __restrict__ d1, double* __restrict__ d2) {
  // ws[(edge * 2 + q) * n_waves + wave]: q = 0 first, 1 second derivative
  const int c = blockIdx.x;
  const int src = (c == rb) ? ra : c;
  CSum a, b;
  if (c != root) {
    const double* w1 = ws + (size_t)src * 2 * n_waves;
    const double* w2 = w1 + n_waves;
    for (int64_t w = threadIdx.x; w < n_waves; w += blockDim.x) {
      a.add(w1[w]);
      b.add(w2[w]);
    }
  }
  const double ta = block_csum(a);
  __syncthreads();  // block_csum's shared slots are reused
  const double tb = block_csum(b);
  if (threadIdx.x == 0) {
    d1[c] = ta;
    d2[c] = tb;
  }
}

// ---------------------------------------------------------------------------
// Incremental SPR scoring.
//
// The move (s, v) prunes the subtree of s -- its parent p is removed and its
// sibling x hangs from g = parent(p) with t_x + t_p (x becomes the root when
// p is) -- and regrafts it at the middle of the branch above v.  With T' the
// tree the prune leaves, D'_v the partial of v in T' and A'_v the message of
// the pre-order pass of T' (L(T') = sum_r prob_r A'_v . (P(t_v) D'_v)), the
// candidate's likelihood per pattern and rate class is, exactly,
//   L_r = sum_j [P(t_v / 2)^T A'_v]_j [P(t_v / 2) D'_v]_j X_j,  X = P(t_s) D_s.
// So a repair of the base tree's partials (only the ancestors g .. root of p
// change) and one pre-order pass over T' score the insertion at every branch
// of T': about 2N node operations per pruned subtree instead of N per move.
// The half-branch products take the forms of the derivative pass,
//   P(h)^T A = A + Ui^T (m (*) U^T A),  P(h) D = D + U (m (*) Ui D),
// m = exp(lam rate h) - 1, which keep short branches precise.
//
// Partials live in one pool of rows in the stored layout of the post-order
// pass (row i: R*4 value rows and R exponent rows of `pitch` patterns): rows
// 0 .. N-2 hold the base tree's internal nodes (row v - N), the rows after
// them the repaired ancestors.  ref[k][c] is the row of internal node c in
// the pruned tree of subtree k.  meta[k] = {path offset, path length, x (-1
// when p is the root), s, program entries, stack depth, 0, 0}.  Threads are
// (pattern, rate) pairs, lane p*R + r, as in the derivative pass.
// ---------------------------------------------------------------------------
constexpr int kIncrMeta = 8;

// y = P D = D + U (f (*) Ui D) for the factors f = exp(lam rate t) - 1; u
// receives Ui D
template <typename real>
__device__ __forceinline__ void eig_apply(const real* U, const real* Ui, const real* f, const real* D,
                                          real* u, real* y) {
  real m[4], t[4];
  matvec_u(Ui, D, u);
#pragma unroll
  for (int k = 0; k < 4; ++k) m[k] = f[k] * u[k];
  matvec_u(U, m, t);
#pragma unroll
  for (int k = 0; k < 4; ++k) y[k] = D[k] + t[k];
}

template <typename real, int R>
__device__ __forceinline__ void pool_read(const real* __restrict__ part, const int32_t* __restrict__ ex,
                                          int row, int r, int64_t pitch, int64_t pc, real* D, int* e) {
  const real* sv = part + (((size_t)row * R + r) * 4) * pitch + pc;
#pragma unroll
  for (int k = 0; k < 4; ++k) D[k] = sv[(size_t)k * pitch];
  *e = ex[((size_t)row * R + r) * pitch + pc];
}

// Repair along the path g .. root of each pruned subtree (grid.y): ops
// {y, a, b, flags} bottom-up, flags bit 0 (bit 1): a (b) is the previous
// op's node, held in registers.  The branch of x takes the merged factors
// Ex[k]; the other branches keep the base tree's E.
template <typename real, int R>
__global__ __launch_bounds__(256) void k_spr_repair(
    const int4* __restrict__ ops, const int32_t* __restrict__ meta, const int32_t* __restrict__ ref,
    const real* __restrict__ tab, const real* __restrict__ E, const real* __restrict__ Ex,
    const uint64_t* __restrict__ tips, int64_t stride, int64_t n_pat, int64_t pitch,
    real* __restrict__ part, int32_t* __restrict__ ex, int N, int M) {
  const int k = blockIdx.y;
  const int32_t* mk = meta + kIncrMeta * k;
  const int n = mk[1];
  const int tid = threadIdx.x;
  const int r = tid % R;
  const int64_t p = (int64_t)blockIdx.x * (blockDim.x / R) + tid / R;
  if (n == 0 || p >= n_pat) return;  // no barriers below
  const int32_t* rf = ref + (size_t)k * M;
  const int x = mk[2];
  const int4* op = ops + mk[0];
  real U[16], Ui[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    U[i] = tab[i];
    Ui[i] = tab[16 + i];
  }
  real acc[4] = {(real)0, (real)0, (real)0, (real)0};
  int acce = 0;
  for (int i = 0; i < n; ++i) {
    const int4 o = op[i];
    real y[2][4];
    int ye[2];
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int node = c ? o.z : o.y;
      real D[4], f[4], u[4];
      int e;
      if ((o.w >> c) & 1) {
#pragma unroll
        for (int j = 0; j < 4; ++j) D[j] = acc[j];
        e = acce;
      } else if (node < N) {
        tipvec(tip_code(tips, stride, node, p), D);
        e = 0;
      } else {
        pool_read<real, R>(part, ex, rf[node], r, pitch, p, D, &e);
      }
      const real* fs = (node == x) ? Ex + ((size_t)k * R + r) * 4 : E + ((size_t)node * R + r) * 4;
#pragma unroll
      for (int j = 0; j < 4; ++j) f[j] = fs[j];
      eig_apply(U, Ui, f, D, u, y[c]);
      ye[c] = e;
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[j] = y[0][j] * y[1][j];
    acce = ye[0] + ye[1];
    rescale4(acc, &acce);
    const int row = rf[o.x];
    real* d = part + (((size_t)row * R + r) * 4) * pitch + p;
#pragma unroll
    for (int j = 0; j < 4; ++j) d[(size_t)j * pitch] = acc[j];
    ex[((size_t)row * R + r) * pitch + p] = acce;
  }
}

// Pre-order pass over each pruned tree (grid.y = subtree k) scoring the
// insertion of X at the middle of every branch: per wave and target c, the
// weighted sum of the per-pattern lnL goes to wsum[(k * M + c) * n_waves +
// wave].  The message flow is the derivative pass's; each entry adds the
// insertion at both children.
template <typename real, int R>
__global__ __launch_bounds__(256) void k_spr_insert(
    const int4* __restrict__ prog, int64_t prog_stride, const int32_t* __restrict__ meta,
    const int32_t* __restrict__ ref, const real* __restrict__ tab, const real* __restrict__ E,
    const real* __restrict__ Eh, const real* __restrict__ Ex, const double* __restrict__ pi,
    const double* __restrict__ probs, const uint64_t* __restrict__ tips, int64_t stride,
    int64_t n_pat, int64_t pitch, const real* __restrict__ part, const int32_t* __restrict__ ex,
    const uint32_t* __restrict__ weights, double* __restrict__ wsum, int64_t n_waves, int N, int M,
    int up_depth) {
  constexpr int QPL = R >= 2 ? 1 : 2;  // insertion targets per lane
  extern __shared__ unsigned char lds_raw[];
  const int k = blockIdx.y;
  const int32_t* mk = meta + kIncrMeta * k;
  const int n_ops = mk[4];
  if (n_ops == 0) return;  // the whole block: no barriers below
  const int tid = threadIdx.x, bd = blockDim.x;
  real* stk = reinterpret_cast<real*>(lds_raw);                                  // [depth][4][bd]
  int32_t* stke = reinterpret_cast<int32_t*>(stk + (size_t)up_depth * 4 * bd);  // [depth][bd]
  const int r = tid % R;
  const int64_t p = (int64_t)blockIdx.x * (bd / R) + tid / R;
  const bool act = p < n_pat;
  const int64_t pc = act ? p : 0;
  const double wp = act ? (double)weights[p] : 0.0;
  const int64_t wave = ((int64_t)blockIdx.x * bd + tid) >> 6;
  const int lane = tid & 63;
  const int32_t* rf = ref + (size_t)k * M;
  const int x = mk[2], s = mk[3];
  const int4* pg = prog + (size_t)k * prog_stride;
  double* ws = wsum + (size_t)k * M * n_waves;

  real U[16], Ui[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    U[i] = tab[i];
    Ui[i] = tab[16 + i];
  }
  const double pr = probs[r];

  // X = P(t_s) D_s: the subtree of s and its branch are untouched
  real X[4];
  int eX;
  {
    real D[4], f[4], u[4];
    if (s < N) {
      tipvec(tip_code(tips, stride, s, pc), D);
      eX = 0;
    } else {
      pool_read<real, R>(part, ex, s - N, r, pitch, pc, D, &eX);
    }
#pragma unroll
    for (int j = 0; j < 4; ++j) f[j] = E[((size_t)s * R + r) * 4 + j];
    eig_apply(U, Ui, f, D, u, X);
  }

  // partial, exponent, full and half factors of a child, fetched one node
  // ahead
  auto fetch = [&](int c, real* Dv, int* e, real* ev, real* hv) {
    if (c < N) {
      tipvec(tip_code(tips, stride, c, pc), Dv);
      *e = 0;
    } else {
      pool_read<real, R>(part, ex, rf[c], r, pitch, pc, Dv, e);
    }
    const real* es = (c == x) ? Ex + ((size_t)k * R + r) * 4 : E + ((size_t)c * R + r) * 4;
    const real* hs = Eh + ((size_t)c * R + r) * 4;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      ev[j] = es[j];
      hv[j] = hs[j];
    }
  };
  // insertion at the middle of the branch above a child with partial D
  // (exponent eD, u = Ui D, half factors h) under the arriving message A
  // (exponent eA, gA = U^T A): the class's term and, after the FP64 rate sum
  // aligned to the largest exponent among the non-zero terms, L and m with
  // the pattern's likelihood L 2^m
  auto insert = [&](const real* A, const real* gA, int eA, const real* D, const real* u, int eD,
                    const real* h, double* L, int* mo) {
    real t[4], left[4], q[4], right[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) t[j] = h[j] * gA[j];
    vecmat_u(Ui, t, left);
#pragma unroll
    for (int j = 0; j < 4; ++j) left[j] += A[j];
    int eL = eA;
    rescale4(left, &eL);
#pragma unroll
    for (int j = 0; j < 4; ++j) q[j] = h[j] * u[j];
    matvec_u(U, q, right);
    real sv = (real)0;
#pragma unroll
    for (int j = 0; j < 4; ++j) sv = fma(left[j] * (D[j] + right[j]), X[j], sv);
    const int Ev = eL + eD + eX;
    const bool nz = sv != (real)0;
    int m = nz ? Ev : kNoExp;
    if (R >= 2) m = max(m, __builtin_amdgcn_update_dpp(0, m, DPP_XOR1, 0xF, 0xF, false));
    if (R >= 4) m = max(m, __builtin_amdgcn_update_dpp(0, m, DPP_XOR2, 0xF, 0xF, false));
    if (R >= 8) m = max(m, __builtin_amdgcn_update_dpp(0, m, DPP_HALF_MIRROR, 0xF, 0xF, false));
    const double sc = nz ? ldexp(pr, Ev - m) : 0.0;
    *L = rate_sum<R>(sc * (double)sv);
    *mo = m;
  };

  real Dc[8], Ec[8], Hc[8];
  int Xc[2];
  {
    const int4 o = pg[0];
    fetch(o.y, Dc, &Xc[0], Ec, Hc);
    fetch(o.z, Dc + 4, &Xc[1], Ec + 4, Hc + 4);
  }
  real accM[4] = {(real)0, (real)0, (real)0, (real)0};
  int accE = 0;
  int sp = 0;
  for_each_entry_ahead(pg, n_ops, [&](int i, const int4 op, const int4 on) {
    real Dn[8], En[8], Hn[8];
    int Xn[2] = {0, 0};
    const bool more = i + 1 < n_ops;
    if (more) {
      fetch(on.y, Dn, &Xn[0], En, Hn);
      fetch(on.z, Dn + 4, &Xn[1], En + 4, Hn + 4);
    }
    const int fl = op.w, src = fl & 3;
    const int a = op.y, b = op.z;
    real Mv[4];
    int eM;
    if (src == SRC_LEAF) {
#pragma unroll
      for (int j = 0; j < 4; ++j) Mv[j] = (real)pi[j];
      eM = 0;
    } else if (src == SRC_ACC) {
#pragma unroll
      for (int j = 0; j < 4; ++j) Mv[j] = accM[j];
      eM = accE;
    } else {
      --sp;
#pragma unroll
      for (int j = 0; j < 4; ++j) Mv[j] = stk[((size_t)sp * 4 + j) * bd + tid];
      eM = stke[(size_t)sp * bd + tid];
    }
    real ua[4], ub[4], ya[4], yb[4], Aa[4], Ab[4], ga[4], gb[4];
    eig_apply(U, Ui, Ec, Dc, ua, ya);
    eig_apply(U, Ui, Ec + 4, Dc + 4, ub, yb);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      Aa[j] = Mv[j] * yb[j];
      Ab[j] = Mv[j] * ya[j];
    }
    vecmat_u(U, Aa, ga);
    vecmat_u(U, Ab, gb);
    double Lq[2];
    int mq[2];
    insert(Aa, ga, eM + Xc[1], Dc, ua, Xc[0], Hc, &Lq[0], &mq[0]);
    insert(Ab, gb, eM + Xc[0], Dc + 4, ub, Xc[1], Hc + 4, &Lq[1], &mq[1]);
    // weighted per-pattern lnL of the two targets, spread over the R lanes
    // of the pattern (lane r takes target r * QPL + j), summed over each
    // 16-lane row's lanes of one target, then over the wave's four rows
    {
      double qv[QPL];
#pragma unroll
      for (int j = 0; j < QPL; ++j) {
        const int q = r * QPL + j;
        double v = 0.0;
        if (act && q < 2) {
          const double L = q ? Lq[1] : Lq[0];
          const int m = q ? mq[1] : mq[0];
          v = wp * (log(L) + (double)m * 0.6931471805599453);
        }
        qv[j] = row_sum_stride<R>(v);
      }
      double S[2];
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        const double vq = (QPL == 1) ? qv[0] : qv[q];
        const int l0 = (QPL == 1) ? q : 0;
        S[q] = (readlane(vq, l0) + readlane(vq, 16 + l0)) + (readlane(vq, 32 + l0) + readlane(vq, 48 + l0));
      }
      if (lane == 0) {
        ws[(size_t)a * n_waves + wave] = S[0];
        ws[(size_t)b * n_waves + wave] = S[1];
      }
    }
    // messages to internal children, as in the derivative pass
    real Ma[4], Mb[4];
    int eMa = eM + Xc[1], eMb = eM + Xc[0];
    if (!(fl & (UP_A_LEAF | UP_A_CUT))) {
      real t[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) t[j] = Ec[j] * ga[j];
      vecmat_u(Ui, t, Ma);
#pragma unroll
      for (int j = 0; j < 4; ++j) Ma[j] += Aa[j];
      rescale4(Ma, &eMa);
    }
    if (!(fl & (UP_B_LEAF | UP_B_CUT))) {
      real t[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) t[j] = Ec[4 + j] * gb[j];
      vecmat_u(Ui, t, Mb);
#pragma unroll
      for (int j = 0; j < 4; ++j) Mb[j] += Ab[j];
      rescale4(Mb, &eMb);
    }
    if (fl & UP_PUSH) {
      const bool cb = fl & UP_CONT_B;
      const real* o = cb ? Ma : Mb;
#pragma unroll
      for (int j = 0; j < 4; ++j) stk[((size_t)sp * 4 + j) * bd + tid] = o[j];
      stke[(size_t)sp * bd + tid] = cb ? eMa : eMb;
      ++sp;
    }
    if (!(fl & (UP_A_LEAF | UP_A_CUT)) || !(fl & (UP_B_LEAF | UP_B_CUT))) {
      const bool cb = fl & UP_CONT_B;
      const real* c = cb ? Mb : Ma;
#pragma unroll
      for (int j = 0; j < 4; ++j) accM[j] = c[j];
      accE = cb ? eMb : eMa;
    }
    if (more) {
#pragma unroll
      for (int j = 0; j < 8; ++j) {
        Dc[j] = Dn[j];
        Ec[j] = En[j];
        Hc[j] = Hn[j];
      }
      Xc[0] = Xn[0];
      Xc[1] = Xn[1];
    }
  });
}

// The insertion pass with one rate class per wave: a block is R waves over
// the same 64 patterns (lane = pattern), so a wave's P matrices are
// wave-uniform (scalar loads) and partial rows are read whole.  P matrices
// instead of eigenbasis factors: Pf (full branches), Ph (half branches), Px
// (the merged branch of x, per subtree).  The classes of a target meet per
// entry in LDS -- exponents aligned over the non-zero terms, the FP64 rate
// sum -- where wave 0 finishes target a and wave 1 target b (wave 0 both
// when R = 1); double-buffered by entry parity, one barrier per entry.
// wsum[(k * M + c) * n_blk + block]: per target and block of 64 patterns.
template <typename real, int R>
__global__ __launch_bounds__(64 * R) void k_spr_insert_w(
    const int4* __restrict__ prog, int64_t prog_stride, const int32_t* __restrict__ meta,
    const int32_t* __restrict__ ref, const real* __restrict__ Pf, const real* __restrict__ Ph,
    const real* __restrict__ Px, const double* __restrict__ pi, const double* __restrict__ probs,
    const uint64_t* __restrict__ tips, int64_t stride, int64_t n_pat, int64_t pitch,
    const real* __restrict__ part, const int32_t* __restrict__ ex, const uint32_t* __restrict__ weights,
    double* __restrict__ wsum, int64_t n_blk, int N, int M, int up_depth) {
  constexpr int BD = 64 * R;
  extern __shared__ unsigned char lds_raw[];
  const int k = blockIdx.y;
  const int32_t* mk = meta + kIncrMeta * k;
  const int n_ops = mk[4];
  if (n_ops == 0) return;  // the whole block: before any barrier
  const int tid = threadIdx.x, lane = tid & 63;
  const int r = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t p = (int64_t)blockIdx.x * 64 + lane;
  const bool act = p < n_pat;
  const int64_t pc = act ? p : 0;
  real* stk = reinterpret_cast<real*>(lds_raw);                                  // [depth][4][BD]
  int32_t* stke = reinterpret_cast<int32_t*>(stk + (size_t)up_depth * 4 * BD);  // [depth][BD]
  double* xv = reinterpret_cast<double*>(stke + (size_t)up_depth * BD);         // [2][2][R][64]
  int32_t* xe = reinterpret_cast<int32_t*>(xv + 2 * 2 * R * 64);                // [2][2][R][64]
  const int32_t* rf = ref + (size_t)k * M;
  const int x = mk[2], s = mk[3];
  const int4* pg = prog + (size_t)k * prog_stride;
  double* ws = wsum + (size_t)k * M * n_blk;
  const real* pxk = Px + ((size_t)k * R + r) * 16;
  auto pfull = [&](int c) { return c == x ? pxk : Pf + ((size_t)c * R + r) * 16; };

  // X = P(t_s) D_s: the subtree of s and its branch are untouched
  real X[4];
  int eX = 0;
  {
    real D[4];
    if (s < N) {
      tipvec(tip_code(tips, stride, s, pc), D);
    } else {
      const real* sv = part + (((size_t)(s - N) * R + r) * 4) * pitch + pc;
#pragma unroll
      for (int j = 0; j < 4; ++j) D[j] = sv[(size_t)j * pitch];
      eX = ex[((size_t)(s - N) * R + r) * pitch + pc];
    }
    real Ps[16];
    load_mat<real, R>(Pf + ((size_t)s * R + r) * 16, Ps);
    matvec4(Ps, D, X);
  }
  auto fetch = [&](int c, real* Dv, int* e) {
    if (c < N) {
      tipvec(tip_code(tips, stride, c, pc), Dv);
      *e = 0;
    } else {
      const int row = rf[c];
      const real* sv = part + (((size_t)row * R + r) * 4) * pitch + pc;
#pragma unroll
      for (int j = 0; j < 4; ++j) Dv[j] = sv[(size_t)j * pitch];
      *e = ex[((size_t)row * R + r) * pitch + pc];
    }
  };
  // the class's insertion term at the middle of the branch above a child
  // with partial D, under the message A arriving at it
  auto insert = [&](const real* Phc, const real* A, const real* D) {
    real left[4], right[4];
    vecmat4(Phc, A, left);
    matvec4(Phc, D, right);
    real sv = (real)0;
#pragma unroll
    for (int j = 0; j < 4; ++j) sv = fma(left[j] * right[j], X[j], sv);
    return sv;
  };

  real Dc[8];
  int Xc[2];
  {
    const int4 o = pg[0];
    fetch(o.y, Dc, &Xc[0]);
    fetch(o.z, Dc + 4, &Xc[1]);
  }
  real accM[4] = {(real)0, (real)0, (real)0, (real)0};
  int accE = 0;
  int sp = 0;
  for_each_entry_ahead(pg, n_ops, [&](int i, const int4 op, const int4 on) {
    real Dn[8];
    int Xn[2] = {0, 0};
    const bool more = i + 1 < n_ops;
    if (more) {
      fetch(on.y, Dn, &Xn[0]);
      fetch(on.z, Dn + 4, &Xn[1]);
    }
    const int fl = op.w, src = fl & 3;
    const int a = op.y, b = op.z;
    real Mv[4];
    int eM;
    if (src == SRC_LEAF) {
#pragma unroll
      for (int j = 0; j < 4; ++j) Mv[j] = (real)pi[j];
      eM = 0;
    } else if (src == SRC_ACC) {
#pragma unroll
      for (int j = 0; j < 4; ++j) Mv[j] = accM[j];
      eM = accE;
    } else {
      --sp;
#pragma unroll
      for (int j = 0; j < 4; ++j) Mv[j] = stk[((size_t)sp * 4 + j) * BD + tid];
      eM = stke[(size_t)sp * BD + tid];
    }
    real PA[16], PB[16];
    load_mat<real, R>(pfull(a), PA);
    load_mat<real, R>(pfull(b), PB);
    real ya[4], yb[4], Aa[4], Ab[4];
    matvec4(PA, Dc, ya);
    matvec4(PB, Dc + 4, yb);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      Aa[j] = Mv[j] * yb[j];
      Ab[j] = Mv[j] * ya[j];
    }
    int eAa = eM + Xc[1], eAb = eM + Xc[0];
    rescale4(Aa, &eAa);
    rescale4(Ab, &eAb);
    real sa, sb;
    {
      real PH[16];
      load_mat<real, R>(Ph + ((size_t)a * R + r) * 16, PH);
      sa = insert(PH, Aa, Dc);
      load_mat<real, R>(Ph + ((size_t)b * R + r) * 16, PH);
      sb = insert(PH, Ab, Dc + 4);
    }
    // exchange; wave 0 finishes target a, wave 1 (wave 0 when R = 1) b
    {
      const int buf = i & 1;
      double* xb0 = xv + (size_t)buf * 2 * R * 64;
      int32_t* eb0 = xe + (size_t)buf * 2 * R * 64;
      xb0[r * 64 + lane] = (double)sa;
      eb0[r * 64 + lane] = eAa + Xc[0] + eX;
      xb0[(R + r) * 64 + lane] = (double)sb;
      eb0[(R + r) * 64 + lane] = eAb + Xc[1] + eX;
      __syncthreads();
#pragma unroll
      for (int q = 0; q < 2; ++q) {
        if (r != (R == 1 ? 0 : q)) continue;
        const double* sq = xb0 + (size_t)q * R * 64;
        const int32_t* eq = eb0 + (size_t)q * R * 64;
        int m = kNoExp;
#pragma unroll
        for (int rr = 0; rr < R; ++rr)
          if (sq[rr * 64 + lane] != 0.0) m = max(m, eq[rr * 64 + lane]);
        // the class sum and its log in the working precision (single: the
        // per-pattern term to ~1e-7, far inside the path's 1e-6 on lnL);
        // the exponent offset and the weighted sums in FP64
        real L = (real)0;
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
          const real v = (real)sq[rr * 64 + lane];
          if (v != (real)0) L += (real)probs[rr] * ldexp(v, eq[rr * 64 + lane] - m);
        }
        const double v = act ? (double)weights[p] * ((double)log(L) + (double)m * 0.6931471805599453) : 0.0;
        const double S = wave_sum_d(v);
        if (lane == 0) ws[(size_t)(q ? b : a) * n_blk + blockIdx.x] = S;
      }
    }
    // messages to internal children: M_c = P_c^T A_c
    real Ma[4], Mb[4];
    int eMa = eAa, eMb = eAb;
    if (!(fl & (UP_A_LEAF | UP_A_CUT))) {
      vecmat4(PA, Aa, Ma);
      rescale4(Ma, &eMa);
    }
    if (!(fl & (UP_B_LEAF | UP_B_CUT))) {
      vecmat4(PB, Ab, Mb);
      rescale4(Mb, &eMb);
    }
    if (fl & UP_PUSH) {
      const bool cb = fl & UP_CONT_B;
      const real* o = cb ? Ma : Mb;
#pragma unroll
      for (int j = 0; j < 4; ++j) stk[((size_t)sp * 4 + j) * BD + tid] = o[j];
      stke[(size_t)sp * BD + tid] = cb ? eMa : eMb;
      ++sp;
    }
    if (!(fl & (UP_A_LEAF | UP_A_CUT)) || !(fl & (UP_B_LEAF | UP_B_CUT))) {
      const bool cb = fl & UP_CONT_B;
      const real* c = cb ? Mb : Ma;
#pragma unroll
      for (int j = 0; j < 4; ++j) accM[j] = c[j];
      accE = cb ? eMb : eMa;
    }
    if (more) {
#pragma unroll
      for (int j = 0; j < 8; ++j) Dc[j] = Dn[j];
      Xc[0] = Xn[0];
      Xc[1] = Xn[1];
    }
  });
}

// pairs per LDS exchange of k_tbr_insert_w
constexpr int kTbrChunk = 4;

// ---------------------------------------------------------------------------
// Incremental TBR scoring.  The move (s, w, v) reroots the subtree of s at
// the middle of the branch above w (w in the subtree; its old root s
// suppressed, s's two branches one edge) and regrafts it as the SPR move
// (s, v), or keeps it in place (v = -1).  Per rerooting the subtree's root
// vector is
//   X_w = P(t_s) R_w,  R_w = (P(t_w/2) D_w) (*) (P(t_w/2) B_w),
// B_w the conditional likelihood of the rest of the subtree seen at w's
// parent, from a pre-order pass over the subtree that starts from C_s = 1:
//   B_a = (P_b D_b) (*) C_n,  C_a = P_a B_a   (entry n with children a, b);
// X_s = P(t_s) D_s (no rerooting).  The insertion pass over the pruned tree
// then scores every needed (target, rerooting) pair; keeping the attachment
// is the insertion on x's merged branch at its original split (P(t_p) above,
// P(t_x) below), or, when p is the root, the root of the pruned tree joined
// to the subtree (pi, then P(t_x)).
//
// tbr meta[k]: {subtree program entries, s, first X row of the slot in its
// batch, subtree stack depth, x (the sibling of s), g >= 0, 0, 0}.
// wslot[k][c]: X row of the rerooting at c (-1: not needed); a slot's
// rerootings take consecutive rows from its first.
// ---------------------------------------------------------------------------
template <typename real, int R>
__global__ __launch_bounds__(64 * R) void k_tbr_reroot(
    const int4* __restrict__ prog, int64_t prog_stride, const int32_t* __restrict__ meta,
    const int32_t* __restrict__ wslot, const real* __restrict__ Pf, const real* __restrict__ Ph,
    const uint64_t* __restrict__ tips, int64_t stride, int64_t n_pat, int64_t pitch,
    const real* __restrict__ part, const int32_t* __restrict__ ex, real* __restrict__ xv,
    int32_t* __restrict__ xe, int N, int M, int up_depth) {
  constexpr int BD = 64 * R;
  extern __shared__ unsigned char lds_raw[];
  const int k = blockIdx.y;
  const int32_t* mk = meta + kIncrMeta * k;
  const int n_ops = mk[0], s = mk[1];
  const int tid = threadIdx.x, lane = tid & 63;
  const int r = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t p = (int64_t)blockIdx.x * 64 + lane;
  const bool act = p < n_pat;
  const int64_t pc = act ? p : 0;
  const int32_t* ws = wslot + (size_t)k * M;
  real* stk = reinterpret_cast<real*>(lds_raw);                                  // [depth][4][BD]
  int32_t* stke = reinterpret_cast<int32_t*>(stk + (size_t)up_depth * 4 * BD);  // [depth][BD]
  real Ps[16];
  load_mat<real, R>(Pf + ((size_t)s * R + r) * 16, Ps);
  auto fetch = [&](int c, real* Dv, int* e) {
    if (c < N) {
      tipvec(tip_code(tips, stride, c, pc), Dv);
      *e = 0;
    } else {
      const real* sv = part + (((size_t)(c - N) * R + r) * 4) * pitch + pc;
#pragma unroll
      for (int j = 0; j < 4; ++j) Dv[j] = sv[(size_t)j * pitch];
      *e = ex[((size_t)(c - N) * R + r) * pitch + pc];
    }
  };
  auto store = [&](int row, const real* R4, int e) {
    real X[4];
    int eX = e;
    matvec4(Ps, R4, X);
    rescale4(X, &eX);
    if (act) {
      real* d = xv + (((size_t)row * R + r) * 4) * pitch + p;
#pragma unroll
      for (int j = 0; j < 4; ++j) d[(size_t)j * pitch] = X[j];
      xe[((size_t)row * R + r) * pitch + p] = eX;
    }
  };
  if (ws[s] >= 0) {  // no rerooting
    real D[4];
    int e;
    fetch(s, D, &e);
    store(ws[s], D, e);
  }
  if (n_ops == 0) return;  // no barriers in this kernel
  real Dc[8];
  int Xc[2];
  {
    const int4 o = prog[(size_t)k * prog_stride];
    fetch(o.y, Dc, &Xc[0]);
    fetch(o.z, Dc + 4, &Xc[1]);
  }
  real accC[4] = {(real)0, (real)0, (real)0, (real)0};
  int accE = 0;
  int sp = 0;
  for_each_entry_ahead(prog + (size_t)k * prog_stride, n_ops, [&](int i, const int4 op, const int4 on) {
    real Dn[8];
    int Xn[2] = {0, 0};
    const bool more = i + 1 < n_ops;
    if (more) {
      fetch(on.y, Dn, &Xn[0]);
      fetch(on.z, Dn + 4, &Xn[1]);
    }
    const int fl = op.w, src = fl & 3;
    const int a = op.y, b = op.z;
    real Cv[4];
    int eC;
    if (src == SRC_LEAF) {  // the subtree's root: nothing outside it
#pragma unroll
      for (int j = 0; j < 4; ++j) Cv[j] = (real)1;
      eC = 0;
    } else if (src == SRC_ACC) {
#pragma unroll
      for (int j = 0; j < 4; ++j) Cv[j] = accC[j];
      eC = accE;
    } else {
      --sp;
#pragma unroll
      for (int j = 0; j < 4; ++j) Cv[j] = stk[((size_t)sp * 4 + j) * BD + tid];
      eC = stke[(size_t)sp * BD + tid];
    }
    real PA[16], PB[16];
    load_mat<real, R>(Pf + ((size_t)a * R + r) * 16, PA);
    load_mat<real, R>(Pf + ((size_t)b * R + r) * 16, PB);
    real ya[4], yb[4], Ba[4], Bb[4];
    matvec4(PA, Dc, ya);
    matvec4(PB, Dc + 4, yb);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      Ba[j] = yb[j] * Cv[j];
      Bb[j] = ya[j] * Cv[j];
    }
    int eBa = eC + Xc[1], eBb = eC + Xc[0];
    rescale4(Ba, &eBa);
    rescale4(Bb, &eBb);
    // rerootings at a and b
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int c = q ? b : a;
      const int row = ws[c];
      if (row < 0) continue;
      real PH[16], u[4], v[4], Rw[4];
      load_mat<real, R>(Ph + ((size_t)c * R + r) * 16, PH);
      matvec4(PH, q ? Dc + 4 : Dc, u);
      matvec4(PH, q ? Bb : Ba, v);
#pragma unroll
      for (int j = 0; j < 4; ++j) Rw[j] = u[j] * v[j];
      int eR = (q ? Xc[1] + eBb : Xc[0] + eBa);
      rescale4(Rw, &eR);
      store(row, Rw, eR);
    }
    // outside vectors at internal children: C_c = P_c B_c
    real Ca[4], Cb[4];
    int eCa = eBa, eCb = eBb;
    if (!(fl & UP_A_LEAF)) {
      matvec4(PA, Ba, Ca);
      rescale4(Ca, &eCa);
    }
    if (!(fl & UP_B_LEAF)) {
      matvec4(PB, Bb, Cb);
      rescale4(Cb, &eCb);
    }
    if (fl & UP_PUSH) {
      const bool cb = fl & UP_CONT_B;
      const real* o = cb ? Ca : Cb;
#pragma unroll
      for (int j = 0; j < 4; ++j) stk[((size_t)sp * 4 + j) * BD + tid] = o[j];
      stke[(size_t)sp * BD + tid] = cb ? eCa : eCb;
      ++sp;
    }
    if (!(fl & UP_A_LEAF) || !(fl & UP_B_LEAF)) {
      const bool cb = fl & UP_CONT_B;
      const real* c = cb ? Cb : Ca;
#pragma unroll
      for (int j = 0; j < 4; ++j) accC[j] = c[j];
      accE = cb ? eCb : eCa;
    }
    if (more) {
#pragma unroll
      for (int j = 0; j < 8; ++j) Dc[j] = Dn[j];
      Xc[0] = Xn[0];
      Xc[1] = Xn[1];
    }
  });
}

// The insertion pass of TBR moves over each pruned tree: at every entry,
// for each child c with needed pairs (tmask[k][c], one bit per rerooting
// row w of the subtree), q = (P(t_c/2)^T A_c) (*) (P(t_c/2) D_c) and the
// class term q . X_w of each needed w; the classes of a pair meet in LDS
// (kTbrChunk pairs per exchange, double-buffered by exchange parity, one
// barrier per exchange, wave r finishing pairs r, r + R, ... of it).  Keeping the attachment (tmask[k][M]) is the
// child x of g with P(t_p) above and P(t_x) below (Pk[k]), or, when p is the
// root, the first entry (x, the pruned tree's root) with pi above.
// wsum row of (k, c, w): rb[k][c] + the rank of w among c's needed bits.
template <typename real, int R>
__global__ __launch_bounds__(64 * R) void k_tbr_insert_w(
    const int4* __restrict__ prog, int64_t prog_stride, const int32_t* __restrict__ meta,
    const int32_t* __restrict__ tmeta, const int32_t* __restrict__ ref, const uint32_t* __restrict__ tmask,
    const int32_t* __restrict__ rbase, const real* __restrict__ Pf, const real* __restrict__ Ph,
    const real* __restrict__ Px, const real* __restrict__ Pk, const real* __restrict__ xv,
    const int32_t* __restrict__ xe, const double* __restrict__ pi, const double* __restrict__ probs,
    const uint64_t* __restrict__ tips, int64_t stride, int64_t n_pat, int64_t pitch,
    const real* __restrict__ part, const int32_t* __restrict__ ex, const uint32_t* __restrict__ weights,
    double* __restrict__ wsum, int64_t n_blk, int N, int M, int up_depth) {
  constexpr int BD = 64 * R;
  extern __shared__ unsigned char lds_raw[];
  const int k = blockIdx.y;
  const int32_t* mk = meta + kIncrMeta * k;
  const int32_t* tk = tmeta + kIncrMeta * k;
  const int n_ops = mk[4];
  const int tid = threadIdx.x, lane = tid & 63;
  const int r = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int64_t p = (int64_t)blockIdx.x * 64 + lane;
  const bool act = p < n_pat;
  const int64_t pc = act ? p : 0;
  real* stk = reinterpret_cast<real*>(lds_raw);                                  // [depth][4][BD]
  int32_t* stke = reinterpret_cast<int32_t*>(stk + (size_t)up_depth * 4 * BD);  // [depth][BD]
  double* xvl = reinterpret_cast<double*>(stke + (size_t)up_depth * BD);        // [2][chunk][R][64]
  int32_t* xel = reinterpret_cast<int32_t*>(xvl + 2 * kTbrChunk * R * 64);      // [2][chunk][R][64]
  const int32_t* rf = ref + (size_t)k * M;
  const uint32_t* tm = tmask + (size_t)k * (M + 1);
  const int32_t* rb = rbase + (size_t)k * (M + 1);
  const int x = mk[2];
  const bool has_g = tk[5] != 0;
  const int xrow0 = tk[2];
  const int4* pg = prog + (size_t)k * prog_stride;
  const real* pxk = Px + ((size_t)k * R + r) * 16;
  auto pfull = [&](int c) { return c == x ? pxk : Pf + ((size_t)c * R + r) * 16; };
  int chunk = 0;  // exchanges so far (uniform)
  // the class terms of the needed pairs (c, w) of q, kTbrChunk pairs per
  // exchange: every wave forms its class's terms of the chunk (the X rows
  // loaded together), one barrier, then wave r finishes pairs r, r + R, ...
  auto pairs = [&](uint32_t mask, int row0, const real* q, int eq) {
    mask = __builtin_amdgcn_readfirstlane(mask);  // the same word in every lane
    while (mask) {
      const int buf = chunk & 1;
      int n = 0;
#pragma unroll
      for (int j = 0; j < kTbrChunk; ++j) {
        if (mask) {
          const int w = __builtin_ctz(mask);
          mask &= mask - 1;
          const int xr = xrow0 + w;
          const real* xs = xv + (((size_t)xr * R + r) * 4) * pitch + pc;
          real sv = (real)0;
#pragma unroll
          for (int i = 0; i < 4; ++i) sv = fma(q[i], xs[(size_t)i * pitch], sv);
          const int slot = ((buf * kTbrChunk + j) * R + r) * 64 + lane;
          xvl[slot] = (double)sv;
          xel[slot] = eq + xe[((size_t)xr * R + r) * pitch + pc];
          ++n;
        }
      }
      __syncthreads();
      for (int j = r; j < n; j += R) {
        const double* sq = xvl + (buf * kTbrChunk + j) * R * 64;
        const int32_t* eqq = xel + (buf * kTbrChunk + j) * R * 64;
        int m = kNoExp;
#pragma unroll
        for (int rr = 0; rr < R; ++rr)
          if (sq[rr * 64 + lane] != 0.0) m = max(m, eqq[rr * 64 + lane]);
        real L = (real)0;
#pragma unroll
        for (int rr = 0; rr < R; ++rr) {
          const real v = (real)sq[rr * 64 + lane];
          if (v != (real)0) L += (real)probs[rr] * ldexp(v, eqq[rr * 64 + lane] - m);
        }
        const double v = act ? (double)weights[p] * ((double)log(L) + (double)m * 0.6931471805599453) : 0.0;
        const double S = wave_sum_d(v);
        if (lane == 0) wsum[(size_t)(row0 + j) * n_blk + blockIdx.x] = S;
      }
      row0 += n;
      ++chunk;
    }
  };
  auto fetch = [&](int c, real* Dv, int* e) {
    if (c < N) {
      tipvec(tip_code(tips, stride, c, pc), Dv);
      *e = 0;
    } else {
      const int row = rf[c];
      const real* sv = part + (((size_t)row * R + r) * 4) * pitch + pc;
#pragma unroll
      for (int j = 0; j < 4; ++j) Dv[j] = sv[(size_t)j * pitch];
      *e = ex[((size_t)row * R + r) * pitch + pc];
    }
  };
  if (n_ops == 0) {  // block-uniform
    // the pruned tree is the leaf x alone (p the root): only keeping the
    // attachment, pi above x's tip through P(t_x)
    if (!has_g && tm[M]) {
      real D[4], PK[16], right[4], q[4];
      tipvec(tip_code(tips, stride, tk[4], pc), D);
      load_mat<real, R>(Pk + (((size_t)k * 2 + 1) * R + r) * 16, PK);
      matvec4(PK, D, right);
#pragma unroll
      for (int j = 0; j < 4; ++j) q[j] = (real)pi[j] * right[j];
      pairs(tm[M], rb[M], q, 0);
    }
    return;
  }
  real Dc[8];
  int Xc[2];
  {
    const int4 o = pg[0];
    fetch(o.y, Dc, &Xc[0]);
    fetch(o.z, Dc + 4, &Xc[1]);
  }
  real accM[4] = {(real)0, (real)0, (real)0, (real)0};
  int accE = 0;
  int sp = 0;
  for_each_entry_ahead(pg, n_ops, [&](int i, const int4 op, const int4 on) {
    real Dn[8];
    int Xn[2] = {0, 0};
    const bool more = i + 1 < n_ops;
    if (more) {
      fetch(on.y, Dn, &Xn[0]);
      fetch(on.z, Dn + 4, &Xn[1]);
    }
    const int fl = op.w, src = fl & 3;
    const int a = op.y, b = op.z;
    real Mv[4];
    int eM;
    if (src == SRC_LEAF) {
#pragma unroll
      for (int j = 0; j < 4; ++j) Mv[j] = (real)pi[j];
      eM = 0;
    } else if (src == SRC_ACC) {
#pragma unroll
      for (int j = 0; j < 4; ++j) Mv[j] = accM[j];
      eM = accE;
    } else {
      --sp;
#pragma unroll
      for (int j = 0; j < 4; ++j) Mv[j] = stk[((size_t)sp * 4 + j) * BD + tid];
      eM = stke[(size_t)sp * BD + tid];
    }
    real PA[16], PB[16];
    load_mat<real, R>(pfull(a), PA);
    load_mat<real, R>(pfull(b), PB);
    real ya[4], yb[4], Aa[4], Ab[4];
    matvec4(PA, Dc, ya);
    matvec4(PB, Dc + 4, yb);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      Aa[j] = Mv[j] * yb[j];
      Ab[j] = Mv[j] * ya[j];
    }
    int eAa = eM + Xc[1], eAb = eM + Xc[0];
    rescale4(Aa, &eAa);
    rescale4(Ab, &eAb);
    // keeping the attachment when p is the root: the pruned tree's root x
    // (this first entry) joined to the subtree, pi above
    if (i == 0 && !has_g && tm[M]) {
      real D[4], right[4], q[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) D[j] = ya[j] * yb[j];
      int eD = Xc[0] + Xc[1];
      rescale4(D, &eD);
      real PK[16];
      load_mat<real, R>(Pk + (((size_t)k * 2 + 1) * R + r) * 16, PK);
      matvec4(PK, D, right);
#pragma unroll
      for (int j = 0; j < 4; ++j) q[j] = (real)pi[j] * right[j];
      pairs(tm[M], rb[M], q, eD);
    }
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const int c = t ? b : a;
      const real* A = t ? Ab : Aa;
      const real* D = t ? Dc + 4 : Dc;
      const int eA = t ? eAb : eAa, eD = Xc[t];
      if (tm[c]) {
        real PH[16], left[4], right[4], q[4];
        load_mat<real, R>(Ph + ((size_t)c * R + r) * 16, PH);
        vecmat4(PH, A, left);
        matvec4(PH, D, right);
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = left[j] * right[j];
        int eq = eA + eD;
        rescale4(q, &eq);
        pairs(tm[c], rb[c], q, eq);
      }
      // keeping the attachment: x's merged branch at its original split
      if (c == x && has_g && tm[M]) {
        real PU[16], PD[16], left[4], right[4], q[4];
        load_mat<real, R>(Pk + (((size_t)k * 2 + 0) * R + r) * 16, PU);
        load_mat<real, R>(Pk + (((size_t)k * 2 + 1) * R + r) * 16, PD);
        vecmat4(PU, A, left);
        matvec4(PD, D, right);
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = left[j] * right[j];
        int eq = eA + eD;
        rescale4(q, &eq);
        pairs(tm[M], rb[M], q, eq);
      }
    }
    // messages to internal children: M_c = P_c^T A_c
    real Ma[4], Mb[4];
    int eMa = eAa, eMb = eAb;
    if (!(fl & (UP_A_LEAF | UP_A_CUT))) {
      vecmat4(PA, Aa, Ma);
      rescale4(Ma, &eMa);
    }
    if (!(fl & (UP_B_LEAF | UP_B_CUT))) {
      vecmat4(PB, Ab, Mb);
      rescale4(Mb, &eMb);
    }
    if (fl & UP_PUSH) {
      const bool cb = fl & UP_CONT_B;
      const real* o = cb ? Ma : Mb;
#pragma unroll
      for (int j = 0; j < 4; ++j) stk[((size_t)sp * 4 + j) * BD + tid] = o[j];
      stke[(size_t)sp * BD + tid] = cb ? eMa : eMb;
      ++sp;
    }
    if (!(fl & (UP_A_LEAF | UP_A_CUT)) || !(fl & (UP_B_LEAF | UP_B_CUT))) {
      const bool cb = fl & UP_CONT_B;
      const real* c = cb ? Mb : Ma;
#pragma unroll
      for (int j = 0; j < 4; ++j) accM[j] = c[j];
      accE = cb ? eMb : eMa;
    }
    if (more) {
#pragma unroll
      for (int j = 0; j < 8; ++j) Dc[j] = Dn[j];
      Xc[0] = Xn[0];
      Xc[1] = Xn[1];
    }
  });
}

// LDS of k_tbr_insert_w: the message stack and the double-buffered exchange
size_t tbr_insert_lds(int R, size_t esz, int depth) {
  return (size_t)std::max(depth, 1) * 64 * R * (4 * esz + 4) + (size_t)2 * kTbrChunk * R * 64 * 12;
}
// LDS of k_tbr_reroot: the outside-vector stack
size_t tbr_reroot_lds(int R, size_t esz, int depth) {
  return (size_t)std::max(depth, 1) * 64 * R * (4 * esz + 4);
}

// LDS of k_spr_insert_w: the message stack and the double-buffered exchange
size_t insert_w_lds(int R, size_t esz, int depth) {
  return (size_t)std::max(depth, 1) * 64 * R * (4 * esz + 4) + (size_t)2 * 2 * R * 64 * 12;
}

// out[idx[j]] = compensated sum of row rows[j] of x (n values per row); one
// block per j
__global__ void k_reduce_gather(const double* __restrict__ x, int64_t n, const int64_t* __restrict__ rows,
                                const int64_t* __restrict__ idx, double* __restrict__ out) {
  const int64_t j = blockIdx.x;
  const double* r = x + rows[j] * n;
  CSum a;
  for (int64_t i = threadIdx.x; i < n; i += blockDim.x) a.add(r[i]);
  const double t = block_csum(a);
  if (threadIdx.x == 0) out[idx[j]] = t;
}

// ---------------------------------------------------------------------------
// Per-node API kernels.  A value is R*K rows of `pitch` values, row r*K + i
// holding L[r][i][p], and R rows of int32 exponents e[r][p] (the layout of
// the whole-tree stored partials): every (pattern, rate class) is rescaled
// on its own, and the classes of a pattern meet only in FP64 sums aligned to
// the largest exponent among the classes with a non-zero term.
// One thread per pattern.
// ---------------------------------------------------------------------------
template <typename real, int R>
__global__ void k_lk_leaf(const uint64_t* __restrict__ tip, int64_t n_pat, int64_t pitch,
                          real* __restrict__ out, int32_t* __restrict__ eout) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n_pat) return;
  real t[4];
  tipvec(tip_code(tip, 0, 0, p), t);
#pragma unroll
  for (int r = 0; r < R; ++r) {
#pragma unroll
    for (int i = 0; i < 4; ++i) out[(size_t)(r * 4 + i) * pitch + p] = t[i];
    eout[(size_t)r * pitch + p] = 0;
  }
}

// y = P(rate_r t) x for class r of a stored value at column p
template <typename real>
__device__ __forceinline__ int class_product(const real* __restrict__ P, const real* __restrict__ x,
                                             const int32_t* __restrict__ ex, int r, int64_t pitch,
                                             int64_t p, real* y) {
  real a[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) a[i] = x[(size_t)(r * 4 + i) * pitch + p];
  matvec4(P, a, y);
  return ex[(size_t)r * pitch + p];
}

// Per class, the FP64 weight prob_r 2^(E_r - m) of a term s_r with exponent
// E_r, m the largest E_r over the classes whose term is non-zero; 0 for a
// zero term.  Returns m (kNoExp when every term is zero).
template <int R>
__device__ __forceinline__ int class_weights(const double* s, const int* E,
                                             const double* __restrict__ probs, double* wt) {
  int m = kNoExp;
#pragma unroll
  for (int r = 0; r < R; ++r)
    if (s[r] != 0.0) m = max(m, E[r]);
#pragma unroll
  for (int r = 0; r < R; ++r) wt[r] = s[r] != 0.0 ? ldexp(probs[r], E[r] - m) : 0.0;
  return m;
}

// median_2 (and median_3 when z != nullptr): per class
//   L_out[r] = (P_x x[r]) (*) (P_y y[r]) [(*) (P_z z[r])],
// e_out[r] = e_x[r] + e_y[r] [+ e_z[r]] + its own rescaling exponent.
// With block_sums, also the root sum of the result: one weighted log-likelihood
// partial sum per 64 patterns.
template <typename real, int R>
__global__ void k_lk_median(const real* __restrict__ Pm, const real* __restrict__ x,
                            const int32_t* __restrict__ ex, const real* __restrict__ y,
                            const int32_t* __restrict__ ey, const real* __restrict__ z,
                            const int32_t* __restrict__ ez, int64_t n_pat, int64_t pitch,
                            real* __restrict__ out, int32_t* __restrict__ eout,
                            const double* __restrict__ pi, const double* __restrict__ probs,
                            const uint32_t* __restrict__ weights, double* __restrict__ block_sums) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const bool act = p < n_pat;
  const int64_t pc = act ? p : 0;
  double S[R];
  int E[R];
#pragma unroll
  for (int r = 0; r < R; ++r) {
    real xa[4], yb[4];
    int e = class_product(Pm + (size_t)r * 16, x, ex, r, pitch, pc, xa);
    e += class_product(Pm + (size_t)(R + r) * 16, y, ey, r, pitch, pc, yb);
    real v[4];
#pragma unroll
    for (int i = 0; i < 4; ++i) v[i] = xa[i] * yb[i];
    if (z) {
      real zc[4];
      e += class_product(Pm + (size_t)(2 * R + r) * 16, z, ez, r, pitch, pc, zc);
#pragma unroll
      for (int i = 0; i < 4; ++i) v[i] *= zc[i];
    }
    rescale4(v, &e);
    if (out && act) {
#pragma unroll
      for (int i = 0; i < 4; ++i) out[(size_t)(r * 4 + i) * pitch + p] = v[i];
      eout[(size_t)r * pitch + p] = e;
    }
    S[r] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i) S[r] += pi[i] * (double)v[i];
    E[r] = e;
  }
  if (block_sums) {
    double wt[R];
    const int m = class_weights<R>(S, E, probs, wt);
    double L = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) L += wt[r] * S[r];
    const double contrib =
        act ? (double)(weights ? weights[p] : 1u) * (log(L) + (double)m * 0.6931471805599453) : 0.0;
    const double ws = wave_sum_d(contrib);
    if ((threadIdx.x & 63) == 0) block_sums[p >> 6] = ws;
  }
}

// root_cost: sum_p w_p ln(sum_r prob_r sum_i pi_i L[r][i][p] 2^e[r][p])
template <typename real, int R>
__global__ void k_lk_root(const real* __restrict__ L, const int32_t* __restrict__ e, int64_t n_pat,
                          int64_t pitch, const double* __restrict__ pi,
                          const double* __restrict__ probs, const uint32_t* __restrict__ weights,
                          double* __restrict__ block_sums) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double contrib = 0.0;
  if (p < n_pat) {
    double S[R], wt[R];
    int E[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      S[r] = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) S[r] += pi[i] * (double)L[(size_t)(r * 4 + i) * pitch + p];
      E[r] = e[(size_t)r * pitch + p];
    }
    const int m = class_weights<R>(S, E, probs, wt);
    double lk = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) lk += wt[r] * S[r];
    contrib = (double)(weights ? weights[p] : 1u) * (log(lk) + (double)m * 0.6931471805599453);
  }
  const double ws = wave_sum_d(contrib);
  if ((threadIdx.x & 63) == 0) block_sums[p >> 6] = ws;
}

// One edge of length t between the conditional partials x and y (the two
// sides of the edge; reversible model, so pi sits on either side):
//   L_p = sum_r prob_r sum_ij pi_i x[r][i] P_ij(rate_r t) y[r][j],
// and L', L'' with P' = rate_r Q P and P'' = (rate_r Q)^2 P.  Per 64
// patterns, three partial sums: sum w ln L, sum w L'/L and
// sum w (L''/L - (L'/L)^2), at block_sums[q * n_wave_slots + wave].
template <typename real, int R>
__global__ void k_lk_edge(const real* __restrict__ Pm, const real* __restrict__ P1m,
                          const real* __restrict__ P2m, const real* __restrict__ x,
                          const int32_t* __restrict__ ex, const real* __restrict__ y,
                          const int32_t* __restrict__ ey, int64_t n_pat, int64_t pitch,
                          const double* __restrict__ pi, const double* __restrict__ probs,
                          const uint32_t* __restrict__ weights, double* __restrict__ block_sums,
                          int64_t n_wave_slots) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  double q0 = 0.0, q1 = 0.0, q2 = 0.0;
  if (p < n_pat) {
    double S0[R], S1[R], S2[R], wt[R];
    int E[R];
#pragma unroll
    for (int r = 0; r < R; ++r) {
      real y0[4], y1[4], y2[4], xv[4];
      int e = class_product(Pm + (size_t)r * 16, y, ey, r, pitch, p, y0);
      real yv[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) yv[i] = y[(size_t)(r * 4 + i) * pitch + p];
      matvec4(P1m + (size_t)r * 16, yv, y1);
      matvec4(P2m + (size_t)r * 16, yv, y2);
#pragma unroll
      for (int i = 0; i < 4; ++i) xv[i] = x[(size_t)(r * 4 + i) * pitch + p];
      e += ex[(size_t)r * pitch + p];
      S0[r] = S1[r] = S2[r] = 0.0;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const double a = pi[i] * (double)xv[i];
        S0[r] += a * (double)y0[i];
        S1[r] += a * (double)y1[i];
        S2[r] += a * (double)y2[i];
      }
      E[r] = e;
    }
    const int m = class_weights<R>(S0, E, probs, wt);
    double L0 = 0.0, L1 = 0.0, L2 = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      L0 += wt[r] * S0[r];
      L1 += wt[r] * S1[r];
      L2 += wt[r] * S2[r];
    }
    const double w = (double)(weights ? weights[p] : 1u);
    const double g = L1 / L0;
    q0 = w * (log(L0) + (double)m * 0.6931471805599453);
    q1 = w * g;
    q2 = w * (L2 / L0 - g * g);
  }
  q0 = wave_sum_d(q0);
  q1 = wave_sum_d(q1);
  q2 = wave_sum_d(q2);
  if ((threadIdx.x & 63) == 0) {
    const int64_t wv = p >> 6;
    block_sums[wv] = q0;
    block_sums[n_wave_slots + wv] = q1;
    block_sums[2 * n_wave_slots + wv] = q2;
  }
}

int env_block(const char* name, int dflt) {
  const char* s = getenv(name);
  if (!s) return dflt;
  const int v = atoi(s);
  return (v == 64 || v == 128 || v == 256) ? v : dflt;
}

inline size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Rate classes per wave of the post-order pass.  Two classes per wave share
// the tip expansion and the program bookkeeping, and their four 4x4 P fill 64
// SGPRs, so only single precision can take two.  The fused stack costs the
// same LDS per 64 patterns either way, so with two classes per wave an
// LDS-bound launch keeps half as many waves resident: the fused form
// measures faster with one class per wave, the stored form with two.
// PHY_PRUNE_RPW=1|2 overrides the choice.
int rates_per_wave(size_t esz, int R, bool fused) {
  if (esz != 4 || R % 2) return 1;
  const char* s = getenv("PHY_PRUNE_RPW");
  if (s && (atoi(s) == 1 || atoi(s) == 2)) return atoi(s);
  return fused ? 1 : 2;
}


// Workspace plan of phy_lk_tree
struct Plan {
  size_t off_P = 0, off_E = 0, off_tab = 0, off_part = 0, off_ex = 0, off_bsum = 0, off_wsum = 0;
  size_t total = 0;
  int64_t pitch = 0, n_blocks = 0, n_waves = 0, dblocks = 0, n_dwaves = 0;
  int bs = 0, dbs = 0;  // block size of the post-order pass / derivative pass
  int rpw = 1;          // rate classes per wave of the post-order pass
  size_t esz = 8, lds = 0, dlds = 0;
  bool fused = false;
  bool lds_stack = false;  // stored form with the operand stack in LDS
  // split form of the fused pass: the parts side by side, then the top
  bool split = false;
  int n_parts = 0;
  size_t off_xt = 0, off_xte = 0;
};

constexpr size_t kLdsMax = 64 * 1024;

// LDS of the post-order pass: the (fused) stack, less the levels held in
// registers, which also holds the rate exchange at the root (the same for any
// number of rate classes per wave)
size_t down_lds(int R, size_t esz, bool fused, int depth, int reg_levels = 0) {
  const size_t bs = 64 * (size_t)R;
  const size_t root = (size_t)R * 64 * 12;
  const size_t rows = (size_t)std::max(depth - reg_levels, 0);
  return std::max(root, fused ? rows * bs * (4 * esz + 4) : 0);
}
// LDS of the derivative pass: a message stack of 4 values + an exponent per
// thread and level
size_t deriv_lds(int R, size_t esz, int bs, int depth) {
  return (size_t)std::max(depth, 1) * bs * (4 * esz + 4);
}

// largest block size <= want whose LDS fits (0 if none)
template <class F>
int fit_block(int want, F lds_of, int min_bs = 64) {
  for (int bs = want; bs >= min_bs; bs /= 2)
    if (lds_of(bs) <= kLdsMax) return bs;
  return 0;
}

// The fused pass runs in split form (parts side by side, then the top) when
// the topology has parts; PHY_PRUNE_SPLIT=0 runs the whole tree per block.
bool split_env() {
  const char* s = getenv("PHY_PRUNE_SPLIT");
  return !(s && s[0] == '0' && s[1] == 0);
}

Plan make_plan(const phy_topo* t, const phy_model* m, int64_t n_pat, int32_t form, int32_t derivs,
               bool allow_split = true) {
  Plan pl;
  const int N = t->n_leaves, M = 2 * N - 1, R = m->R, K = m->K;
  pl.esz = (form & PHY_F32) ? 4 : 8;
  const bool fused = (form & PHY_FUSED) != 0;
  const size_t esz = pl.esz;
  const int depth = t->down_depth;
  // a fused stack that does not fit in LDS falls back to stored partials
  // (same arithmetic and result, more HBM traffic)
  const bool stack_fits = down_lds(R, esz, true, depth) <= kLdsMax;
  pl.fused = fused && stack_fits;
  // the stored form keeps its operand stack in LDS too when it fits, so it
  // writes its partials and reads none back (PHY_PRUNE_STORED_LDS=0, or two
  // rate classes per wave, reads the children from HBM)
  const char* sl = getenv("PHY_PRUNE_STORED_LDS");
  const bool want_stack = !pl.fused && stack_fits && !(sl && sl[0] == '0' && sl[1] == 0);
  pl.rpw = rates_per_wave(esz, R, pl.fused || want_stack);
  pl.lds_stack = want_stack && pl.rpw == 1;
  pl.lds = down_lds(R, esz, pl.fused || pl.lds_stack, depth);  // upper bound: no level in registers
  pl.bs = 64 * R / pl.rpw;  // R / rpw waves over 64 patterns
  if (derivs) {
    pl.dbs = fit_block(env_block("PHY_PRUNE_DERIV_BLOCK", 256),
                       [&](int bs) { return deriv_lds(R, esz, bs, t->up_depth); });
    pl.dlds = pl.dbs ? deriv_lds(R, esz, pl.dbs, t->up_depth) : 0;
  }
  pl.pitch = ((n_pat + 255) / 256) * 256;
  if (pl.pitch == 0) pl.pitch = 256;
  pl.n_blocks = pl.pitch / 64;
  pl.dblocks = pl.dbs ? pl.pitch * R / pl.dbs : 0;
  pl.n_waves = pl.pitch / 64;          // partial lnL sums: one per 64 patterns
  pl.n_dwaves = pl.pitch * R / 64;  // derivative sums: one per wave
  size_t o = 0;
  const size_t pm = (size_t)M * R * K * K * pl.esz;
  pl.off_P = o; o = align256(o + pm);
  if (derivs) {
    pl.off_E = o; o = align256(o + (size_t)M * R * K * pl.esz);
    pl.off_tab = o; o = align256(o + (size_t)(2 * K * K + R * K) * pl.esz);
  }
  if (!pl.fused) {
    pl.off_part = o; o = align256(o + (size_t)(N - 1) * R * K * pl.pitch * pl.esz);
    pl.off_ex = o; o = align256(o + (size_t)(N - 1) * R * pl.pitch * 4);
  }
  pl.split = allow_split && (pl.fused || pl.lds_stack) && pl.rpw == 1 && t->pparts.size() >= 2 && t->d_ptop_prog &&
             down_lds(R, esz, true, std::max(t->psplit.depth, t->ptop_depth)) <= kLdsMax && split_env();
  if (pl.split) pl.n_parts = (int)t->pparts.size();
  if (pl.split && pl.fused) {
    pl.off_xt = o; o = align256(o + (size_t)pl.n_parts * R * K * pl.pitch * pl.esz);
    pl.off_xte = o; o = align256(o + (size_t)pl.n_parts * R * pl.pitch * 4);
  }
  pl.off_bsum = o; o = align256(o + (size_t)pl.n_waves * 8);
  if (derivs) {
    pl.off_wsum = o; o = align256(o + (size_t)pl.n_dwaves * M * 2 * 8);
  }
  pl.total = o;
  return pl;
}

// Fused stack levels held in registers (0, 1 or 2) for one rate class per
// wave: the choice that keeps the most waves resident per CU (registers per
// thread against LDS per block), the fewest levels on a tie.  Cached per
// stack depth; PHY_PRUNE_REG=0|1|2 overrides it.
template <typename real, int R>
int register_levels(int depth, int bs) {
  if (depth < 1) return 0;
  const char* s = getenv("PHY_PRUNE_REG");
  if (s && s[0] >= '0' && s[0] <= '2' && s[1] == 0) return std::min(s[0] - '0', depth);
  static std::atomic<int> cache[64];
  static std::atomic<int> cache_dev[64];
  int dev = 0;
  (void)hipGetDevice(&dev);
  const int slot = std::min(depth, 63);
  if (depth < 64 && cache[slot].load() > 0 && cache_dev[slot].load() == dev)
    return cache[slot].load() - 1;
  const void* kern[3] = {reinterpret_cast<const void*>(&k_prune_down<real, R, 1, true, 0>),
                         reinterpret_cast<const void*>(&k_prune_down<real, R, 1, true, 1>),
                         reinterpret_cast<const void*>(&k_prune_down<real, R, 1, true, 2>)};
  int best = 0, best_waves = -1;
  for (int lv = 0; lv <= std::min(2, depth); ++lv) {
    const size_t lds = down_lds(R, sizeof(real), true, depth, lv);
    int nb = 0;
    if (lds > kLdsMax ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern[lv], bs, lds) != hipSuccess)
      continue;
    const int waves = nb * (bs / 64);
    if (waves > best_waves) best = lv, best_waves = waves;
  }
  if (depth < 64) {
    cache_dev[slot].store(dev);
    cache[slot].store(best + 1);
  }
  return best;
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
  if (derivs && !pl.dbs) return set_err(ctx, PHY_ERANGE, "pruning: derivative LDS stack too deep");
  fused = pl.fused;
  {
    const int64_t n = (int64_t)M * R * K * K;
    hipLaunchKernelGGL((k_pmats<real>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream,
                       cst, K, R, blen, HostLens{}, M, P, (real*)nullptr, (real*)nullptr);
    PHY_LAUNCHED(ctx);
  }
  real* part = fused ? nullptr : reinterpret_cast<real*>(work + pl.off_part);
  int32_t* ex = fused ? nullptr : reinterpret_cast<int32_t*>(work + pl.off_ex);
  double* bsum = reinterpret_cast<double*>(work + pl.off_bsum);
  auto down = [&](auto kern, size_t lds) {
    hipLaunchKernelGGL(kern, dim3((unsigned)pl.n_blocks), dim3(pl.bs), lds, ctx->stream,
                       t->d_down_prog, N - 1, t->d_down_fifo, tips, stride, n_pat, pl.pitch, P, d_pi,
                       d_probs, part, ex, weights, site_lnl, bsum, N, t->down_depth, (real*)nullptr,
                       (int32_t*)nullptr);
  };
  constexpr int G2 = (R % 2 == 0 && sizeof(real) == 4) ? 2 : 1;
  const int lv = ((fused || pl.lds_stack) && pl.rpw != 2) ? register_levels<real, R>(t->down_depth, pl.bs) : 0;
  {
    KernelSpan span(ctx, "prune_down");
    if (pl.split) {
      const SplitProg& sd = t->psplit;
      const int np = pl.n_parts;
      real* xt = fused ? reinterpret_cast<real*>(work + pl.off_xt) : nullptr;
      int32_t* xte = fused ? reinterpret_cast<int32_t*>(work + pl.off_xte) : nullptr;
      const int lvp = register_levels<real, R>(sd.depth, pl.bs);
      const size_t ldsp = down_lds(R, sizeof(real), true, sd.depth, lvp);
      auto pad32 = [](int64_t n) { return (unsigned)((n + 31) / 32 * 32); };
      auto parts = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(pad32(pl.n_blocks * np)), dim3(pl.bs), ldsp, ctx->stream,
                           sd.d_prog, sd.d_meta, np, sd.d_fifo, tips, stride, n_pat, pl.pitch, P, xt, xte,
                           part, ex, N, sd.depth);
      };
      if (fused) {
        if (lvp == 2)
          parts(k_prune_part<real, R, 2>);
        else if (lvp == 1)
          parts(k_prune_part<real, R, 1>);
        else
          parts(k_prune_part<real, R, 0>);
      } else {
        if (lvp == 2)
          parts(k_prune_part<real, R, 2, true>);
        else if (lvp == 1)
          parts(k_prune_part<real, R, 1, true>);
        else
          parts(k_prune_part<real, R, 0, true>);
      }
      PHY_LAUNCHED(ctx);
      const int dt = std::max(t->ptop_depth, 1);
      auto top = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3(pad32(pl.n_blocks)), dim3(pl.bs), down_lds(R, sizeof(real), true, dt),
                           ctx->stream, t->d_ptop_prog, (int)t->ptop_prog.size() / 4, t->d_ptop_fifo, tips,
                           stride, n_pat, pl.pitch, P, d_pi, d_probs, part, ex, weights, site_lnl, bsum, N, dt,
                           xt, xte);
      };
      if (fused)
        top(k_prune_down<real, R, 1, true, 0, 2>);
      else
        top(k_prune_down<real, R, 1, true, 0, 2, true>);
    } else if (pl.lds_stack) {
      const size_t lds = down_lds(R, sizeof(real), true, t->down_depth, lv);
      if (lv == 2)
        down(k_prune_down<real, R, 1, true, 2, 0, true>, lds);
      else if (lv == 1)
        down(k_prune_down<real, R, 1, true, 1, 0, true>, lds);
      else
        down(k_prune_down<real, R, 1, true, 0, 0, true>, lds);
    } else if (!fused) {
      pl.rpw == 2 ? down(k_prune_down<real, R, G2, false, 0>, pl.lds)
                  : down(k_prune_down<real, R, 1, false, 0>, pl.lds);
    } else if (pl.rpw == 2) {
      down(k_prune_down<real, R, G2, true, 0>, pl.lds);
    } else {
      const size_t lds = down_lds(R, sizeof(real), true, t->down_depth, lv);
      if (lv == 2)
        down(k_prune_down<real, R, 1, true, 2>, lds);
      else if (lv == 1)
        down(k_prune_down<real, R, 1, true, 1>, lds);
      else
        down(k_prune_down<real, R, 1, true, 0>, lds);
    }
  }
  PHY_LAUNCHED(ctx);
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, ctx->stream, bsum, pl.n_waves, (int64_t)1, lnl);
  PHY_LAUNCHED(ctx);
  if (derivs) {
    double* wsum = reinterpret_cast<double*>(work + pl.off_wsum);
    real* E = reinterpret_cast<real*>(work + pl.off_E);
    real* tab = reinterpret_cast<real*>(work + pl.off_tab);
    const int64_t ne = std::max<int64_t>((int64_t)M * R * K, 32 + 4 * R);
    hipLaunchKernelGGL((k_eigen_exp<real>), dim3((unsigned)((ne + 255) / 256)), dim3(256), 0,
                       ctx->stream, cst, R, blen, M, E, tab);
    PHY_LAUNCHED(ctx);
    {
      KernelSpan span(ctx, "prune_deriv");
      hipLaunchKernelGGL((k_prune_deriv<real, R>), dim3((unsigned)pl.dblocks), dim3(pl.dbs), pl.dlds,
                         ctx->stream, t->d_up_prog, N - 1, tips, stride, n_pat, pl.pitch, tab, E,
                         d_pi, d_probs, part, ex, weights, wsum, pl.n_dwaves, N,
                         std::max(t->up_depth, 1));
    }
    PHY_LAUNCHED(ctx);
    const int ra = t->children[2 * (t->root - N)], rb = t->children[2 * (t->root - N) + 1];
    hipLaunchKernelGGL(k_reduce_derivs, dim3((unsigned)M), dim3(256), 0, ctx->stream, wsum,
                       pl.n_dwaves, M, t->root, ra, rb, d1, d2);
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
                     m->d_consts, m->K, m->R, d_t, HostLens{}, nb, P, P1, P2);
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
  if (d_exp) *d_exp = reinterpret_cast<int32_t*>(w + pl.off_ex) + (size_t)(v - N) * m->R * pl.pitch;
  if (pitch) *pitch = pl.pitch;
  return PHY_OK;
}

int phy_lk_tree_noalloc(phy_ctx* ctx, const phy_topo* t, const phy_model* m, int32_t form,
                        int64_t n_pat, const uint64_t* tips, int64_t stride,
                        const uint32_t* weights, const double* blen, double* lnl,
                        double* site_lnl) {
  if (!ctx) return PHY_EINVAL;
  if (!t || !t->d_down_prog || !m) return set_err(ctx, PHY_EINVAL, "lk_tree_noalloc: topology/model missing");
  if (m->K != 4) return set_err(ctx, PHY_EINVAL, "lk_tree_noalloc: this entry point is for 4-state models");
  if (form & ~(PHY_F32 | PHY_FUSED)) return set_err(ctx, PHY_EINVAL, "lk_tree_noalloc: unknown form bits");
  if (!(form & PHY_FUSED)) return set_err(ctx, PHY_EINVAL, "lk_tree_noalloc: the fused form only (PHY_FUSED)");
  if (n_pat < 1) return set_err(ctx, PHY_EINVAL, "lk_tree_noalloc: need at least one pattern");
  if (stride < phy_words(n_pat)) return set_err(ctx, PHY_EINVAL, "lk_tree_noalloc: tip stride < words");
  if (!tips || !weights || !blen || !lnl) return set_err(ctx, PHY_EINVAL, "lk_tree_noalloc: null argument");
  const Plan pl = make_plan(t, m, n_pat, form, 0, false);
  if (!pl.fused)
    return set_err(ctx, PHY_ERANGE, "lk_tree_noalloc: the tree's stack does not fit the fused form");
  if (pl.total > kNoAllocScratchBytes)
    return set_err(ctx, PHY_ERANGE, "lk_tree_noalloc: workspace of " + std::to_string(pl.total) +
                                        " bytes exceeds the fixed scratch of " +
                                        std::to_string(kNoAllocScratchBytes));
  if (!ctx->noalloc_scratch)
    return set_err(ctx, PHY_EDEVICE, "lk_tree_noalloc: the scratch array of this device is unresolved");
  char* w = static_cast<char*>(ctx->noalloc_scratch);
  if (form & PHY_F32)
    return dispatch_tree<float>(ctx, t, m, true, false, n_pat, tips, stride, weights, blen, lnl,
                                site_lnl, nullptr, nullptr, w, pl);
  return dispatch_tree<double>(ctx, t, m, true, false, n_pat, tips, stride, weights, blen, lnl,
                               site_lnl, nullptr, nullptr, w, pl);
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
// Batched SPR candidate scoring
// ---------------------------------------------------------------------------
extern "C" int phy_spr_apply(int32_t N, const int32_t* children, int32_t root, const double* blen,
                             int32_t s, int32_t v, int32_t* children_out, int32_t* root_out,
                             double* blen_out);

namespace {

struct SprPlan {
  size_t off_P, off_bl, off_prog, off_fifo, off_ws, total;
  int64_t pitch, n_blocks, n_waves;
  int bs, depth_cap, rpw;
};

SprPlan spr_plan(int32_t N, const phy_model* m, int64_t n_pat, int32_t form, int32_t chunk) {
  SprPlan pl;
  const int M = 2 * N - 1;
  const size_t esz = (form & PHY_F32) ? 4 : 8;
  pl.rpw = rates_per_wave(esz, m->R, true);
  pl.bs = 64 * m->R / pl.rpw;
  pl.pitch = ((n_pat + 255) / 256) * 256;
  pl.n_blocks = pl.pitch / 64;
  pl.n_waves = pl.pitch / 64;
  pl.depth_cap = 12;
  size_t o = 0;
  pl.off_P = o; o = align256(o + (size_t)chunk * M * m->R * 16 * esz);
  pl.off_bl = o; o = align256(o + (size_t)chunk * M * 8);
  pl.off_prog = o; o = align256(o + (size_t)chunk * (N - 1) * 16);
  pl.off_fifo = o; o = align256(o + (size_t)chunk * phy_fifo_padded(N) * 4);
  pl.off_ws = o; o = align256(o + (size_t)chunk * pl.n_waves * 8);
  pl.total = o;
  return pl;
}

template <typename real, int R>
int spr_launch(phy_ctx* ctx, const phy_model* m, int N, int64_t n_pat, const uint64_t* tips,
               int64_t stride, const uint32_t* w, int C, int depth, char* work, const SprPlan& pl,
               double* out) {
  const int M = 2 * N - 1, K = 4;
  const double* cst = m->d_consts;
  const double* d_pi = cst + 2 * K * K + K + R;
  real* P = reinterpret_cast<real*>(work + pl.off_P);
  const double* bl = reinterpret_cast<const double*>(work + pl.off_bl);
  const int64_t n = (int64_t)C * M * R * 16;
  hipLaunchKernelGGL((k_pmats<real>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, cst,
                     K, R, bl, HostLens{}, C * M, P, (real*)nullptr, (real*)nullptr);
  PHY_LAUNCHED(ctx);
  const size_t lds = down_lds(R, sizeof(real), true, depth);
  if (lds > kLdsMax) return set_err(ctx, PHY_ERANGE, "spr scores: LDS stack too deep");
  double* ws = reinterpret_cast<double*>(work + pl.off_ws);
  auto down = [&](auto kern) {
    hipLaunchKernelGGL(kern, dim3((unsigned)(pl.pitch / 64), (unsigned)C), dim3(pl.bs), lds,
                       ctx->stream, reinterpret_cast<const int4*>(work + pl.off_prog),
                       reinterpret_cast<const int32_t*>(work + pl.off_fifo), N, tips, stride, n_pat,
                       pl.pitch, P, d_pi, d_pi + K, w, ws, pl.n_waves, depth);
  };
  constexpr int G2 = (R % 2 == 0 && sizeof(real) == 4) ? 2 : 1;
  {
    KernelSpan span(ctx, "prune_down_batch");
    if (pl.rpw == 2 && G2 == 2)
      down(k_prune_down_batch<real, R, G2>);
    else
      down(k_prune_down_batch<real, R, 1>);
  }
  PHY_LAUNCHED(ctx);
  hipLaunchKernelGGL(k_reduce_rows, dim3((unsigned)C), dim3(256), 0, ctx->stream, ws, pl.n_waves, out);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

}  // namespace

extern "C" {

int phy_lk_spr_work_size(int32_t N, const phy_model* m, int64_t n_pat, int32_t form, int32_t chunk,
                         size_t* bytes) {
  if (!m || !bytes || N < 3 || n_pat < 1 || chunk < 1) return PHY_EINVAL;
  *bytes = spr_plan(N, m, n_pat, form, chunk).total;
  return PHY_OK;
}

}  // extern "C"

namespace {

// Score n_moves candidate trees of the base tree; apply(i, ch, &root, bl)
// builds candidate i (PHY_OK or an error).
template <class Apply>
int score_candidates(phy_ctx* ctx, const phy_model* m, int32_t form, int64_t n_pat,
                     const uint64_t* tips, int64_t stride, const uint32_t* weights, int32_t N,
                     int64_t n_moves, int32_t chunk, double* d_lnl, void* d_work,
                     size_t work_bytes, const char* what, Apply apply) {
  const SprPlan pl = spr_plan(N, m, n_pat, form, chunk);
  if (work_bytes < pl.total) return set_err(ctx, PHY_EINVAL, std::string(what) + ": workspace too small");
  const int M = 2 * N - 1;
  char* work = static_cast<char*>(d_work);
  std::vector<int32_t> prog, fifo;
  std::vector<double> bl;
  for (int64_t base = 0; base < n_moves; base += chunk) {
    const int C = (int)std::min<int64_t>(chunk, n_moves - base);
    prog.assign((size_t)C * (N - 1) * 4, 0);
    const size_t Fp = phy_fifo_padded(N);
    fifo.assign((size_t)C * Fp, 0);
    bl.assign((size_t)C * M, 0.0);
    std::vector<int> depth(C, 0), bad(C, 0);
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned k = 0; k < nt; ++k)
      th.emplace_back([&, k] {
        std::vector<int32_t> ch((size_t)2 * (N - 1));
        for (int c = (int)k; c < C; c += (int)nt) {
          int32_t r = -1;
          if (apply(base + c, ch.data(), &r, &bl[(size_t)c * M]) != PHY_OK) {
            bad[c] = 1;
            continue;
          }
          phy_topo* t = nullptr;
          if (phy_topo_create(nullptr, N, ch.data(), r, &t) != PHY_OK) {
            bad[c] = 1;
            continue;
          }
          std::copy(t->down_prog.begin(), t->down_prog.end(), prog.begin() + (size_t)c * (N - 1) * 4);
          std::fill(fifo.begin() + (size_t)c * Fp, fifo.begin() + (size_t)(c + 1) * Fp, t->down_fifo[0]);
          std::copy(t->down_fifo.begin(), t->down_fifo.end(), fifo.begin() + (size_t)c * Fp);
          depth[c] = t->down_depth;
          phy_topo_destroy(t);
        }
      });
    for (auto& x : th) x.join();
    for (int c = 0; c < C; ++c)
      if (bad[c]) return set_err(ctx, PHY_EINVAL, std::string(what) + ": invalid move " + std::to_string(base + c));
    const int dmax = *std::max_element(depth.begin(), depth.end());
    PHY_HIP(ctx, hipMemcpyAsync(work + pl.off_prog, prog.data(), prog.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    PHY_HIP(ctx, hipMemcpyAsync(work + pl.off_fifo, fifo.data(), fifo.size() * 4, hipMemcpyHostToDevice, ctx->stream));
    PHY_HIP(ctx, hipMemcpyAsync(work + pl.off_bl, bl.data(), bl.size() * 8, hipMemcpyHostToDevice, ctx->stream));
    int rc;
    switch ((form & PHY_F32 ? 10 : 0) + m->R) {
      case 1: rc = spr_launch<double, 1>(ctx, m, N, n_pat, tips, stride, weights, C, dmax, work, pl, d_lnl + base); break;
      case 4: rc = spr_launch<double, 4>(ctx, m, N, n_pat, tips, stride, weights, C, dmax, work, pl, d_lnl + base); break;
      case 11: rc = spr_launch<float, 1>(ctx, m, N, n_pat, tips, stride, weights, C, dmax, work, pl, d_lnl + base); break;
      case 14: rc = spr_launch<float, 4>(ctx, m, N, n_pat, tips, stride, weights, C, dmax, work, pl, d_lnl + base); break;
      default: rc = set_err(ctx, PHY_ERANGE, std::string(what) + ": n_rates must be 1 or 4");
    }
    if (rc) return rc;
    // the host buffers are reused by the next chunk: wait for the copies
    PHY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  }
  return PHY_OK;
}

int check_scores_args(phy_ctx* ctx, const phy_model* m, int64_t n_pat, const uint64_t* tips,
                      int64_t stride, const uint32_t* weights, int32_t N, const int32_t* children,
                      const double* blen, int64_t n_moves, const int32_t* moves, int32_t chunk,
                      double* d_lnl, void* d_work, int32_t root, const char* what) {
  if (!ctx || !m) return PHY_EINVAL;
  if (m->K != 4) return set_err(ctx, PHY_EINVAL, std::string(what) + ": 4-state models only");
  if (N < 3 || n_pat < 1 || chunk < 1 || n_moves < 0 || !children || !blen || !tips || !weights ||
      !d_lnl || !d_work || (n_moves > 0 && !moves))
    return set_err(ctx, PHY_EINVAL, std::string(what) + ": bad argument");
  if (stride < phy_words(n_pat)) return set_err(ctx, PHY_EINVAL, std::string(what) + ": tip stride < words");
  // the base tree itself (one parent per node, no cycle), before any move
  int64_t n_nni = 0;
  if (phy_nni_moves(N, children, root, &n_nni, nullptr, 0) != PHY_OK)
    return set_err(ctx, PHY_EINVAL, std::string(what) + ": not a rooted binary tree");
  return PHY_OK;
}

}  // namespace

extern "C" {

int phy_lk_spr_scores(phy_ctx* ctx, const phy_model* m, int32_t form, int64_t n_pat,
                      const uint64_t* tips, int64_t stride, const uint32_t* weights, int32_t N,
                      const int32_t* children, int32_t root, const double* blen, int64_t n_moves,
                      const int32_t* moves, int32_t chunk, double* d_lnl, void* d_work,
                      size_t work_bytes) {
  int rc = check_scores_args(ctx, m, n_pat, tips, stride, weights, N, children, blen, n_moves,
                             moves, chunk, d_lnl, d_work, root, "spr scores");
  if (rc) return rc;
  try {
    return score_candidates(ctx, m, form, n_pat, tips, stride, weights, N, n_moves, chunk, d_lnl,
                            d_work, work_bytes, "spr scores",
                            [&](int64_t i, int32_t* ch, int32_t* r, double* bl) {
                              return phy_spr_apply(N, children, root, blen, moves[2 * i],
                                                   moves[2 * i + 1], ch, r, bl);
                            });
  } catch (...) {
    return set_err(ctx, PHY_ENOMEM, "spr scores: host allocation or thread start failed");
  }
}

int phy_lk_tbr_scores(phy_ctx* ctx, const phy_model* m, int32_t form, int64_t n_pat,
                      const uint64_t* tips, int64_t stride, const uint32_t* weights, int32_t N,
                      const int32_t* children, int32_t root, const double* blen, int64_t n_moves,
                      const int32_t* moves, int32_t chunk, double* d_lnl, void* d_work,
                      size_t work_bytes) {
  int rc = check_scores_args(ctx, m, n_pat, tips, stride, weights, N, children, blen, n_moves,
                             moves, chunk, d_lnl, d_work, root, "tbr scores");
  if (rc) return rc;
  try {
    return score_candidates(ctx, m, form, n_pat, tips, stride, weights, N, n_moves, chunk, d_lnl,
                            d_work, work_bytes, "tbr scores",
                            [&](int64_t i, int32_t* ch, int32_t* r, double* bl) {
                              return phy_tbr_apply(N, children, root, blen, moves[3 * i],
                                                   moves[3 * i + 1], moves[3 * i + 2], ch, r, bl);
                            });
  } catch (...) {
    return set_err(ctx, PHY_ENOMEM, "tbr scores: host allocation or thread start failed");
  }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Incremental SPR scoring (host): the base tree's stored post-order pass
// once, then per batch of pruned subtrees the repair of the path g .. root,
// the insertion pre-order pass and the per-move sums.
// ---------------------------------------------------------------------------
namespace {

// rooted tree with parents, leaf counts, depths and Euler intervals
struct IncrTree {
  int32_t N = 0, M = 0, root = -1;
  std::vector<int32_t> ch, parent, size, depth, tin, tout;
  int32_t lmax = 0;  // longest repair path: most ancestors of a non-root internal node

  bool load(int32_t n, const int32_t* children, int32_t r) {
    int64_t n_nni = 0;
    if (phy_nni_moves(n, children, r, &n_nni, nullptr, 0) != PHY_OK) return false;
    N = n;
    M = 2 * n - 1;
    root = r;
    ch.assign(children, children + 2 * (n - 1));
    parent.assign(M, -1);
    for (int32_t k = 0; k < n - 1; ++k) parent[ch[2 * k]] = parent[ch[2 * k + 1]] = n + k;
    size.assign(M, 1);
    depth.assign(M, 0);
    tin.assign(M, 0);
    tout.assign(M, 0);
    std::vector<std::pair<int32_t, int>> st = {{root, 0}};
    int32_t t = 0;
    while (!st.empty()) {
      auto& f = st.back();
      const int32_t v = f.first;
      if (f.second == 0) {
        tin[v] = t++;
        f.second = 1;
        if (v >= N) {
          const int32_t a = ch[2 * (v - N)], b = ch[2 * (v - N) + 1];
          depth[a] = depth[b] = depth[v] + 1;
          st.push_back({b, 0});
          st.push_back({a, 0});
        }
      } else {
        tout[v] = t;
        if (v >= N) size[v] = size[ch[2 * (v - N)]] + size[ch[2 * (v - N) + 1]];
        st.pop_back();
      }
    }
    lmax = 0;
    for (int32_t v = N; v < M; ++v)
      if (v != root) lmax = std::max(lmax, depth[v]);
    return true;
  }
  bool below(int32_t a, int32_t b) const { return tin[a] <= tin[b] && tout[b] <= tout[a]; }  // b under a
  int32_t sibling(int32_t s) const {
    const int32_t p = parent[s];
    return ch[2 * (p - N)] == s ? ch[2 * (p - N) + 1] : ch[2 * (p - N)];
  }
};

// Workspace of a run: lengths and factors, the base tree's P, the partial
// pool, the per-target sums.  Per-subtree factors are sized for every
// pruned subtree a tree can have (M), so the layout depends on the tree
// only through its longest repair path.
struct IncrPlan {
  Plan base;  // stored post-order pass of the base tree
  size_t off_bl = 0, off_E = 0, off_Eh = 0, off_Ex = 0, off_tab = 0, off_P = 0, off_Ph = 0, off_Px = 0;
  size_t off_part = 0, off_ex = 0, off_bsum = 0, off_lnl = 0, off_wsum = 0, off_blh = 0;
  size_t off_xv = 0, off_xe = 0, off_Pk = 0, total = 0;  // TBR: rerooted subtree vectors, keep matrices
  int64_t pitch = 0, n_waves = 0, n_blk = 0, pool_rows = 0;
  int S = 0, lmax = 0, slots = 0;
  int64_t xrows = 0;  // TBR: rerooted subtree vectors per batch
  size_t esz = 8;
};

// Workspace of a run: lengths and factors, the base tree's P, the partial
// pool, the per-target sums; per-slot factors for `slots` pruned subtrees
// (every pruned subtree a tree can have, M, for SPR).  TBR adds the
// rerooted subtree vectors of a batch (xrows rows), the keep
// matrices P(t_p), P(t_x) of each slot, and per-pair sums (max_pairs pairs
// in a batch).
IncrPlan incr_plan(const IncrTree& t, int R, int64_t n_pat, int32_t form, int32_t S, int slots = -1,
                   int64_t xrows = 0, int64_t max_pairs = 0) {
  IncrPlan pl;
  const int N = t.N, M = t.M, K = 4;
  pl.esz = (form & PHY_F32) ? 4 : 8;
  pl.S = S;
  pl.slots = slots < 0 ? M : std::max(slots, 1);
  pl.xrows = xrows;
  pl.lmax = std::max(t.lmax, 1);
  pl.pitch = ((n_pat + 255) / 256) * 256;
  pl.n_waves = pl.pitch * R / 64;
  pl.n_blk = pl.pitch / 64;
  pl.pool_rows = (int64_t)(N - 1) + (int64_t)S * pl.lmax;
  const size_t esz = pl.esz;
  const size_t ns = (size_t)pl.slots;
  size_t o = 0;
  // full | half | merged x of each slot | keep (t_p, t_x) of each slot
  pl.off_bl = o; o = align256(o + ((size_t)2 * M + 3 * ns) * 8);
  pl.off_blh = o; o = align256(o + (size_t)M * 8);  // host lengths of a one-shot call
  pl.off_E = o; o = align256(o + (size_t)M * R * K * esz);
  pl.off_Eh = o; o = align256(o + (size_t)M * R * K * esz);
  pl.off_Ex = o; o = align256(o + ns * R * K * esz);
  pl.off_tab = o; o = align256(o + (size_t)(2 * K * K + R * K) * esz);
  pl.off_P = o; o = align256(o + (size_t)M * R * K * K * esz);
  pl.off_Ph = o; o = align256(o + (size_t)M * R * K * K * esz);
  pl.off_Px = o; o = align256(o + ns * R * K * K * esz);
  pl.off_part = o; o = align256(o + (size_t)pl.pool_rows * R * K * pl.pitch * esz);
  pl.off_ex = o; o = align256(o + (size_t)pl.pool_rows * R * pl.pitch * 4);
  pl.off_bsum = o; o = align256(o + (size_t)(pl.pitch / 64) * 8);
  pl.off_lnl = o; o = align256(o + 8);
  // per-target sums: SPR one row per (subtree of a batch, target), TBR one
  // per needed pair of a batch
  const size_t rows = xrows > 0 ? (size_t)max_pairs * pl.n_blk : (size_t)S * M * std::max(pl.n_waves, pl.n_blk);
  pl.off_wsum = o; o = align256(o + rows * 8);
  if (xrows > 0) {
    pl.off_xv = o; o = align256(o + (size_t)xrows * R * K * pl.pitch * esz);
    pl.off_xe = o; o = align256(o + (size_t)xrows * R * pl.pitch * 4);
    pl.off_Pk = o; o = align256(o + 2 * ns * R * K * K * esz);
  }
  pl.total = o;
  return pl;
}

// Per pruned subtree: the pre-order program of the pruned tree on the
// original node ids, pool rows, repair ops, and x and p (the merged branch
// t_x + t_p; x = -1 when p is the root).
struct IncrSub {
  int32_t s = -1, x = -1, p = -1, n_ops = 0, depth = 0;
  std::vector<int32_t> prog, ref, ops;  // ops: {y, a, b, flags} bottom-up
};

// Internal nodes of the pruned tree of s that the insertion walk enters: the
// proper ancestors, in the pruned tree, of each target (x stands for keeping
// the attachment; x hangs from g there, or is the root when p is).
void incr_enter(const IncrTree& t, int32_t s, const std::vector<int32_t>& targets, std::vector<char>* enter) {
  const int32_t p = t.parent[s], g = t.parent[p];
  enter->assign(t.M, 0);
  for (int32_t c : targets) {
    int32_t y = t.parent[c];
    if (y == p) y = g;
    while (y >= 0 && !(*enter)[y]) {
      (*enter)[y] = 1;
      y = t.parent[y];
      if (y == p) y = g;
    }
  }
}

// enter: the nodes the insertion walk enters, tgt: the targets (nullptr:
// every node).  With both, the repair stops at the highest ancestor whose
// new partial the walk uses (a target, or the sibling of a target or of an
// entered node); the ancestors above keep their base rows, which the walk
// reads only into products it discards.
void incr_prepare(const IncrTree& t, int32_t s, int slot, int lmax, IncrSub* o,
                  const std::vector<char>* enter = nullptr, const std::vector<char>* tgt = nullptr) {
  const int32_t N = t.N, M = t.M;
  const int32_t p = t.parent[s], x = t.sibling(s), g = t.parent[p];
  std::vector<int32_t> ch = t.ch, size = t.size;
  int32_t top = x;
  o->s = s;
  o->x = -1;
  o->p = p;
  o->ops.clear();
  o->ref.resize(M);
  for (int32_t c = 0; c < M; ++c) o->ref[c] = c >= N ? c - N : 0;
  if (g >= 0) {
    (ch[2 * (g - N)] == p ? ch[2 * (g - N)] : ch[2 * (g - N) + 1]) = x;
    top = t.root;
    o->x = x;
    int32_t last = (int32_t)M;  // highest path position to repair
    if (enter && tgt) {
      last = -1;
      int32_t j = 0;
      for (int32_t y = g; y >= 0; y = t.parent[y], ++j) {
        const int32_t up = t.parent[y];
        const int32_t z = up >= 0 ? t.sibling(y) : -1;
        if ((*tgt)[y] || (z >= 0 && ((*tgt)[z] || (*enter)[z]))) last = j;
      }
    }
    int32_t prev = -1, j = 0;
    for (int32_t y = g; y >= 0; y = t.parent[y], ++j) {
      size[y] -= t.size[s];
      if (j > last) continue;
      o->ref[y] = (N - 1) + slot * lmax + j;
      const int32_t a = ch[2 * (y - N)], b = ch[2 * (y - N) + 1];
      o->ops.insert(o->ops.end(), {y, a, b, (a == prev ? 1 : 0) | (b == prev ? 2 : 0)});
      prev = y;
    }
  }
  o->prog.clear();
  o->depth = 0;
  if (top >= N) phy_compile_up_program(N, ch, size, top, &o->prog, &o->depth, enter);
  o->n_ops = (int32_t)o->prog.size() / 4;
}

// Moves grouped by pruned subtree (ascending s): subs, the first move of
// each subtree in the grouped order (n_sub + 1 entries), and per grouped
// move its wsum row (position of s in its batch * M + v) and its index in
// the caller's list.  Invalid moves give their index in *bad.
bool incr_group(const IncrTree& t, int64_t n_moves, const int32_t* moves, int S, std::vector<int32_t>* subs,
                std::vector<int64_t>* start, std::vector<int64_t>* rows, std::vector<int64_t>* idx,
                int64_t* bad) {
  const int32_t M = t.M;
  std::vector<int64_t> count(M, 0);
  for (int64_t i = 0; i < n_moves; ++i) {
    const int32_t s = moves[2 * i], v = moves[2 * i + 1];
    bool ok = s >= 0 && s < M && s != t.root && v >= 0 && v < M;
    if (ok) {
      const int32_t p = t.parent[s], x = t.sibling(s);
      const int32_t new_root = (p == t.root) ? x : t.root;
      ok = !t.below(s, v) && v != p && v != x && v != new_root;
    }
    if (!ok) {
      *bad = i;
      return false;
    }
    ++count[s];
  }
  std::vector<int32_t> pos(M, -1);
  subs->clear();
  for (int32_t s = 0; s < M; ++s)
    if (count[s]) {
      pos[s] = (int32_t)subs->size();
      subs->push_back(s);
    }
  start->assign(subs->size() + 1, 0);
  for (size_t k = 0; k < subs->size(); ++k) (*start)[k + 1] = (*start)[k] + count[(*subs)[k]];
  std::vector<int64_t> fill(start->begin(), start->end() - 1);
  rows->assign(n_moves, 0);
  idx->assign(n_moves, 0);
  for (int64_t i = 0; i < n_moves; ++i) {
    const int32_t s = moves[2 * i], v = moves[2 * i + 1];
    const int64_t j = fill[pos[s]]++;
    (*rows)[j] = (int64_t)(pos[s] % S) * M + v;
    (*idx)[j] = i;
  }
  return true;
}

// lengths of a run from the device blen: full | half | merged x of each
// pruned subtree (t_x + t_p, 0 when p is the root) | with keep (TBR), the
// original split of x's branch of each pruned subtree (t_p, t_x; kx holds x
// also when p is the root)
__global__ void k_incr_lengths(const double* __restrict__ blen, int M, const int32_t* __restrict__ xp,
                               int n_sub, int slots, const int32_t* __restrict__ kx, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < M) {
    out[i] = blen[i];
    out[M + i] = 0.5 * blen[i];
  }
  if (i < n_sub) {
    const int x = xp[2 * i], p = xp[2 * i + 1];
    out[2 * M + i] = x >= 0 ? blen[x] + blen[p] : 0.0;
    if (kx) {
      out[2 * M + slots + 2 * i] = blen[p];
      out[2 * M + slots + 2 * i + 1] = blen[kx[i]];
    }
  }
}

}  // namespace

// A prepared incremental SPR scoring: the moves grouped by pruned subtree,
// each pruned tree's pre-order program, pool-row map and repair ops, the
// base tree's post-order program, all uploaded once; a run issues kernels
// only.
struct phy_spr_plan {
  IncrTree t;
  IncrPlan pl;
  int32_t form = 0, R = 0;
  int64_t n_pat = 0, n_moves = 0;
  int n_sub = 0;
  phy_topo* base = nullptr;               // host programs; device copies in dmem
  std::vector<int64_t> mv_start;          // first grouped move of each subtree (n_sub + 1)
  std::vector<int32_t> batch_depth;       // message stack depth of each batch
  // TBR: the n_sub slots are (pruned subtree, up to 32 rerootings) pairs
  bool tbr = false;
  std::vector<int32_t> batch_rdepth;      // stack depth of the rerooting pass of each batch
  char* dmem = nullptr;
  size_t o_dprog = 0, o_dfifo = 0, o_prog = 0, o_meta = 0, o_ref = 0, o_ops = 0, o_xp = 0, o_rows = 0,
         o_idx = 0, o_tprog = 0, o_tmeta = 0, o_wslot = 0, o_tmask = 0, o_rbase = 0, o_kx = 0;
  ~phy_spr_plan() {
    if (base) {
      base->d_down_prog = nullptr;
      base->d_down_fifo = nullptr;
      phy_topo_destroy(base);
    }
    if (dmem) (void)hipFree(dmem);
  }
};

namespace {

template <typename real, int R>
int incr_run(phy_ctx* ctx, const phy_spr_plan* sp, const phy_model* m, const uint64_t* tips, int64_t stride,
             const uint32_t* weights, const double* d_blen, double* d_lnl, double* d_base, char* work) {
  const IncrTree& t = sp->t;
  const IncrPlan& pl = sp->pl;
  const int N = t.N, M = t.M, K = 4, S = pl.S, n_sub = sp->n_sub;
  const double* cst = m->d_consts;
  const double* d_pi = cst + 2 * K * K + K + R;
  const double* d_probs = d_pi + K;
  double* bl = reinterpret_cast<double*>(work + pl.off_bl);
  real* E = reinterpret_cast<real*>(work + pl.off_E);
  real* Eh = reinterpret_cast<real*>(work + pl.off_Eh);
  real* Ex = reinterpret_cast<real*>(work + pl.off_Ex);
  real* tab = reinterpret_cast<real*>(work + pl.off_tab);
  real* Pf = reinterpret_cast<real*>(work + pl.off_P);
  real* Ph = reinterpret_cast<real*>(work + pl.off_Ph);
  real* Px = reinterpret_cast<real*>(work + pl.off_Px);
  real* part = reinterpret_cast<real*>(work + pl.off_part);
  int32_t* ex = reinterpret_cast<int32_t*>(work + pl.off_ex);
  double* wsum = reinterpret_cast<double*>(work + pl.off_wsum);
  const char* dm = sp->dmem;
  const auto* d_prog = reinterpret_cast<const int4*>(dm + sp->o_prog);
  const auto* d_meta = reinterpret_cast<const int32_t*>(dm + sp->o_meta);
  const auto* d_ref = reinterpret_cast<const int32_t*>(dm + sp->o_ref);
  const auto* d_ops = reinterpret_cast<const int4*>(dm + sp->o_ops);
  const auto* d_xp = reinterpret_cast<const int32_t*>(dm + sp->o_xp);
  const auto* d_rows = reinterpret_cast<const int64_t*>(dm + sp->o_rows);
  const auto* d_idx = reinterpret_cast<const int64_t*>(dm + sp->o_idx);

  // lengths, eigen factors and P matrices of every branch and merged branch
  const int nl = std::max(M, n_sub);
  const auto* d_kx = sp->tbr ? reinterpret_cast<const int32_t*>(dm + sp->o_kx) : nullptr;
  hipLaunchKernelGGL(k_incr_lengths, dim3((unsigned)((nl + 255) / 256)), dim3(256), 0, ctx->stream, d_blen, M,
                     d_xp, n_sub, pl.slots, d_kx, bl);
  PHY_LAUNCHED(ctx);
  auto eig = [&](const double* lens, int nb, real* out) {
    const int64_t ne = std::max<int64_t>((int64_t)nb * R * K, 32 + 4 * R);
    hipLaunchKernelGGL((k_eigen_exp<real>), dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, ctx->stream,
                       cst, R, lens, nb, out, tab);
  };
  auto pmat = [&](const double* lens, int nb, real* out) {
    const int64_t n = (int64_t)nb * R * K * K;
    hipLaunchKernelGGL((k_pmats<real>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, cst,
                       K, R, lens, HostLens{}, nb, out, (real*)nullptr, (real*)nullptr);
  };
  eig(bl, M, E);
  PHY_LAUNCHED(ctx);
  eig(bl + M, M, Eh);
  PHY_LAUNCHED(ctx);
  pmat(bl + M, M, Ph);
  PHY_LAUNCHED(ctx);
  if (n_sub > 0) {
    eig(bl + 2 * M, n_sub, Ex);
    PHY_LAUNCHED(ctx);
    pmat(bl + 2 * M, n_sub, Px);
    PHY_LAUNCHED(ctx);
    if (sp->tbr) {
      pmat(bl + 2 * M + pl.slots, 2 * n_sub, reinterpret_cast<real*>(work + pl.off_Pk));
      PHY_LAUNCHED(ctx);
    }
  }
  // the base tree's partials: stored post-order pass into pool rows 0 .. N-2
  {
    Plan bp = pl.base;
    bp.off_P = pl.off_P;
    bp.off_part = pl.off_part;
    bp.off_ex = pl.off_ex;
    bp.off_bsum = pl.off_bsum;
    double* base_lnl = d_base ? d_base : reinterpret_cast<double*>(work + pl.off_lnl);
    int rc = launch_tree<real, R>(ctx, sp->base, m, false, false, sp->n_pat, tips, stride, weights, bl, base_lnl,
                                  nullptr, nullptr, nullptr, work, bp);
    if (rc) return rc;
  }
  const int64_t n_pat = sp->n_pat;
  for (int b0 = 0, bi = 0; b0 < n_sub; b0 += S, ++bi) {
    const int Sb = std::min(S, n_sub - b0);
    const int depth = std::max(sp->batch_depth[bi], 1);
    const int32_t* meta = d_meta + (size_t)b0 * kIncrMeta;
    const int32_t* ref = d_ref + (size_t)b0 * M;
    {
      KernelSpan span(ctx, "spr_repair");
      hipLaunchKernelGGL((k_spr_repair<real, R>), dim3((unsigned)(pl.pitch * R / 256), (unsigned)Sb), dim3(256),
                         0, ctx->stream, d_ops + (size_t)b0 * pl.lmax, meta, ref, tab, E,
                         Ex + (size_t)b0 * R * K, tips, stride, n_pat, pl.pitch, part, ex, N, M);
    }
    PHY_LAUNCHED(ctx);
    const int4* prog = d_prog + (size_t)b0 * (N - 1);
    const int64_t m0 = sp->mv_start[b0], m1 = sp->mv_start[b0 + Sb];
    if (sp->tbr) {
      const int32_t* tmeta = reinterpret_cast<const int32_t*>(dm + sp->o_tmeta) + (size_t)b0 * kIncrMeta;
      real* xv = reinterpret_cast<real*>(work + pl.off_xv);
      int32_t* xe = reinterpret_cast<int32_t*>(work + pl.off_xe);
      const int rdepth = std::max(sp->batch_rdepth[bi], 1);
      {
        KernelSpan span(ctx, "tbr_reroot");
        hipLaunchKernelGGL((k_tbr_reroot<real, R>), dim3((unsigned)pl.n_blk, (unsigned)Sb), dim3(64 * R),
                           tbr_reroot_lds(R, sizeof(real), rdepth), ctx->stream,
                           reinterpret_cast<const int4*>(dm + sp->o_tprog) + (size_t)b0 * (N - 1),
                           (int64_t)(N - 1), tmeta,
                           reinterpret_cast<const int32_t*>(dm + sp->o_wslot) + (size_t)b0 * M, Pf, Ph, tips,
                           stride, n_pat, pl.pitch, part, ex, xv, xe, N, M, rdepth);
      }
      PHY_LAUNCHED(ctx);
      {
        KernelSpan span(ctx, "tbr_insert");
        hipLaunchKernelGGL((k_tbr_insert_w<real, R>), dim3((unsigned)pl.n_blk, (unsigned)Sb), dim3(64 * R),
                           tbr_insert_lds(R, sizeof(real), depth), ctx->stream, prog, (int64_t)(N - 1), meta,
                           tmeta, ref, reinterpret_cast<const uint32_t*>(dm + sp->o_tmask) + (size_t)b0 * (M + 1),
                           reinterpret_cast<const int32_t*>(dm + sp->o_rbase) + (size_t)b0 * (M + 1), Pf, Ph,
                           Px + (size_t)b0 * R * 16,
                           reinterpret_cast<const real*>(work + pl.off_Pk) + (size_t)b0 * 2 * R * 16, xv, xe,
                           d_pi, d_probs, tips, stride, n_pat, pl.pitch, part, ex, weights, wsum,
                           pl.n_blk, N, M, depth);
      }
      PHY_LAUNCHED(ctx);
      if (m1 > m0) {
        hipLaunchKernelGGL(k_reduce_gather, dim3((unsigned)(m1 - m0)), dim3(256), 0, ctx->stream, wsum, pl.n_blk,
                           d_rows + m0, d_idx + m0, d_lnl);
        PHY_LAUNCHED(ctx);
      }
      continue;
    }
    const size_t wl = insert_w_lds(R, sizeof(real), depth);
    int64_t n_slot;
    if (wl <= kLdsMax) {
      n_slot = pl.n_blk;
      KernelSpan span(ctx, "spr_insert");
      hipLaunchKernelGGL((k_spr_insert_w<real, R>), dim3((unsigned)pl.n_blk, (unsigned)Sb), dim3(64 * R), wl,
                         ctx->stream, prog, (int64_t)(N - 1), meta, ref, Pf, Ph, Px + (size_t)b0 * R * 16, d_pi,
                         d_probs, tips, stride, n_pat, pl.pitch, part, ex, weights, wsum, pl.n_blk, N, M, depth);
    } else {
      // a stack too deep for one class per wave: the (pattern, rate) form
      const int dbs = fit_block(env_block("PHY_SPR_INCR_BLOCK", 256),
                                [&](int bs) { return deriv_lds(R, sizeof(real), bs, depth); });
      if (!dbs) return set_err(ctx, PHY_ERANGE, "spr incremental: LDS stack too deep");
      n_slot = pl.n_waves;
      KernelSpan span(ctx, "spr_insert");
      hipLaunchKernelGGL((k_spr_insert<real, R>), dim3((unsigned)(pl.pitch * R / dbs), (unsigned)Sb), dim3(dbs),
                         deriv_lds(R, sizeof(real), dbs, depth), ctx->stream, prog, (int64_t)(N - 1), meta, ref,
                         tab, E, Eh, Ex + (size_t)b0 * R * K, d_pi, d_probs, tips, stride, n_pat, pl.pitch, part,
                         ex, weights, wsum, pl.n_waves, N, M, depth);
    }
    PHY_LAUNCHED(ctx);
    if (m1 > m0) {
      hipLaunchKernelGGL(k_reduce_gather, dim3((unsigned)(m1 - m0)), dim3(256), 0, ctx->stream, wsum, n_slot,
                         d_rows + m0, d_idx + m0, d_lnl);
      PHY_LAUNCHED(ctx);
    }
  }
  return PHY_OK;
}

int incr_dispatch(phy_ctx* ctx, const phy_spr_plan* sp, const phy_model* m, const uint64_t* tips, int64_t stride,
                  const uint32_t* weights, const double* d_blen, double* d_lnl, double* d_base, char* work) {
  switch ((sp->form & PHY_F32 ? 10 : 0) + m->R) {
#define PHY_INCR(real, Rv) return incr_run<real, Rv>(ctx, sp, m, tips, stride, weights, d_blen, d_lnl, d_base, work)
    case 1: PHY_INCR(double, 1);
    case 2: PHY_INCR(double, 2);
    case 4: PHY_INCR(double, 4);
    case 8: PHY_INCR(double, 8);
    case 11: PHY_INCR(float, 1);
    case 12: PHY_INCR(float, 2);
    case 14: PHY_INCR(float, 4);
    case 18: PHY_INCR(float, 8);
#undef PHY_INCR
    default: return set_err(ctx, PHY_ERANGE, "spr incremental: n_rates must be 1, 2, 4 or 8");
  }
}

int plan_create(phy_ctx* ctx, const phy_model* m, int32_t form, int64_t n_pat, int32_t N, const int32_t* children,
                int32_t root, int64_t n_moves, const int32_t* moves, int32_t s_batch, phy_spr_plan** out) {
  std::unique_ptr<phy_spr_plan> sp(new phy_spr_plan());
  if (!sp->t.load(N, children, root)) return set_err(ctx, PHY_EINVAL, "spr incremental: not a rooted binary tree");
  const IncrTree& t = sp->t;
  const int M = t.M;
  sp->form = form;
  sp->R = m->R;
  sp->n_pat = n_pat;
  sp->n_moves = n_moves;
  sp->pl = incr_plan(t, m->R, n_pat, form, s_batch);
  if (phy_topo_create(nullptr, N, children, root, &sp->base) != PHY_OK)
    return set_err(ctx, PHY_EINVAL, "spr incremental: base topology");
  sp->pl.base = make_plan(sp->base, m, n_pat, form & ~PHY_FUSED, 0);
  std::vector<int32_t> subs;
  std::vector<int64_t> rows, idx;
  int64_t bad = -1;
  if (!incr_group(t, n_moves, moves, s_batch, &subs, &sp->mv_start, &rows, &idx, &bad))
    return set_err(ctx, PHY_EINVAL, "spr incremental: invalid move " + std::to_string(bad));
  const int n_sub = sp->n_sub = (int)subs.size();
  const int S = sp->pl.S, lmax = sp->pl.lmax;
  std::vector<IncrSub> sub(n_sub);
  {
    // each pruned tree's walk enters only the nodes above its moves' targets
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned w = 0; w < std::min<unsigned>(nt, std::max(n_sub, 1)); ++w)
      th.emplace_back([&, w] {
        std::vector<int32_t> targets;
        std::vector<char> enter, tgt(M, 0);
        for (int k = (int)w; k < n_sub; k += (int)nt) {
          targets.clear();
          for (int64_t j = sp->mv_start[k]; j < sp->mv_start[k + 1]; ++j) targets.push_back(moves[2 * idx[j] + 1]);
          for (int32_t c : targets) tgt[c] = 1;
          incr_enter(t, subs[k], targets, &enter);
          incr_prepare(t, subs[k], k % S, lmax, &sub[k], &enter, &tgt);
          for (int32_t c : targets) tgt[c] = 0;
        }
      });
    for (auto& x : th) x.join();
  }
  std::vector<int32_t> hprog((size_t)std::max(n_sub, 1) * (N - 1) * 4, 0), hmeta((size_t)std::max(n_sub, 1) * kIncrMeta, 0),
      href((size_t)std::max(n_sub, 1) * M, 0), hops((size_t)std::max(n_sub, 1) * lmax * 4, 0),
      hxp((size_t)std::max(n_sub, 1) * 2, 0);
  sp->batch_depth.assign((n_sub + S - 1) / S, 1);
  for (int k = 0; k < n_sub; ++k) {
    const IncrSub& q = sub[k];
    std::copy(q.prog.begin(), q.prog.end(), hprog.begin() + (size_t)k * (N - 1) * 4);
    std::copy(q.ref.begin(), q.ref.end(), href.begin() + (size_t)k * M);
    std::copy(q.ops.begin(), q.ops.end(), hops.begin() + (size_t)k * lmax * 4);
    int32_t* mk = &hmeta[(size_t)k * kIncrMeta];
    mk[0] = (k % S) * lmax;
    mk[1] = (int32_t)q.ops.size() / 4;
    mk[2] = q.x;
    mk[3] = q.s;
    mk[4] = q.n_ops;
    mk[5] = q.depth;
    hxp[2 * k] = q.x;
    hxp[2 * k + 1] = q.p;
    sp->batch_depth[k / S] = std::max(sp->batch_depth[k / S], q.depth);
  }
  std::vector<int32_t> fifo(phy_fifo_padded(N), sp->base->down_fifo[0]);
  std::copy(sp->base->down_fifo.begin(), sp->base->down_fifo.end(), fifo.begin());
  // one device allocation for every prepared array
  size_t o = 0;
  auto put = [&](size_t bytes) {
    const size_t at = o;
    o = align256(o + std::max<size_t>(bytes, 8));
    return at;
  };
  sp->o_dprog = put(sp->base->down_prog.size() * 4);
  sp->o_dfifo = put(fifo.size() * 4);
  sp->o_prog = put(hprog.size() * 4);
  sp->o_meta = put(hmeta.size() * 4);
  sp->o_ref = put(href.size() * 4);
  sp->o_ops = put(hops.size() * 4);
  sp->o_xp = put(hxp.size() * 4);
  sp->o_rows = put(rows.size() * 8);
  sp->o_idx = put(idx.size() * 8);
  PHY_HIP(ctx, hipMalloc((void**)&sp->dmem, o));
  auto up = [&](size_t off, const void* src, size_t bytes) {
    return bytes ? hipMemcpy(sp->dmem + off, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
  };
  PHY_HIP(ctx, up(sp->o_dprog, sp->base->down_prog.data(), sp->base->down_prog.size() * 4));
  PHY_HIP(ctx, up(sp->o_dfifo, fifo.data(), fifo.size() * 4));
  PHY_HIP(ctx, up(sp->o_prog, hprog.data(), hprog.size() * 4));
  PHY_HIP(ctx, up(sp->o_meta, hmeta.data(), hmeta.size() * 4));
  PHY_HIP(ctx, up(sp->o_ref, href.data(), href.size() * 4));
  PHY_HIP(ctx, up(sp->o_ops, hops.data(), hops.size() * 4));
  PHY_HIP(ctx, up(sp->o_xp, hxp.data(), hxp.size() * 4));
  PHY_HIP(ctx, up(sp->o_rows, rows.data(), rows.size() * 8));
  PHY_HIP(ctx, up(sp->o_idx, idx.data(), idx.size() * 8));
  sp->base->d_down_prog = reinterpret_cast<int4*>(sp->dmem + sp->o_dprog);
  sp->base->d_down_fifo = reinterpret_cast<int32_t*>(sp->dmem + sp->o_dfifo);
  *out = sp.release();
  return PHY_OK;
}

// TBR moves (s, w, v) grouped into slots: a pruned subtree with up to 32 of
// its rerootings (w = s: none), ascending s.  A slot scores every needed
// (target, rerooting) pair; its pairs take consecutive wsum rows of the
// batch, by target (0 .. M - 1, then M for keeping the attachment) and
// rerooting bit.  Validation follows phy_tbr_apply: w = s or w under s with
// s internal; v = -1 (keep) only with w != s; any other v a valid SPR target.
struct TbrGroups {
  int n_sub = 0;
  int64_t max_pairs = 1, max_xrows = 1;
  std::vector<int32_t> xbase;           // first X row of each slot in its batch
  std::vector<int32_t> slot_s;          // subtree of each slot
  std::vector<int64_t> slot_first;      // first (s, w) of each slot in uniq (n_sub + 1)
  std::vector<int64_t> uniq;            // distinct s * M + w, ascending
  std::vector<uint32_t> tmask;          // [slot][M + 1] rerooting bits of the needed pairs
  std::vector<int32_t> rbase;           // [slot][M + 1] wsum row of the first such pair
  std::vector<int64_t> mv_start, rows, idx;  // moves grouped by slot: wsum row, caller's index
};

// 0, or 1 with *bad the first invalid move, or 2 for a batch of more than
// 2^31 pairs
int tbr_group(const IncrTree& t, int64_t n_moves, const int32_t* moves, int S, TbrGroups* g, int64_t* bad) {
  const int32_t N = t.N, M = t.M;
  std::vector<int64_t> sw(n_moves);
  for (int64_t i = 0; i < n_moves; ++i) {
    const int32_t s = moves[3 * i], w = moves[3 * i + 1], v = moves[3 * i + 2];
    bool ok = s >= 0 && s < M && s != t.root && w >= 0 && w < M;
    if (ok && w != s) ok = s >= N && t.below(s, w);
    if (ok) {
      if (v == -1) {
        ok = w != s;
      } else {
        const int32_t p = t.parent[s], x = t.sibling(s);
        const int32_t new_root = (p == t.root) ? x : t.root;
        ok = v >= 0 && v < M && !t.below(s, v) && v != p && v != x && v != new_root;
      }
    }
    if (!ok) {
      *bad = i;
      return 1;
    }
    sw[i] = (int64_t)s * M + w;
  }
  g->uniq = sw;
  std::sort(g->uniq.begin(), g->uniq.end());
  g->uniq.erase(std::unique(g->uniq.begin(), g->uniq.end()), g->uniq.end());
  // slots: runs of at most 32 rerootings of one subtree
  g->slot_s.clear();
  g->slot_first.clear();
  for (size_t j = 0; j < g->uniq.size(); ++j) {
    const int32_t s = (int32_t)(g->uniq[j] / M);
    if (g->slot_s.empty() || g->slot_s.back() != s || (int64_t)j - g->slot_first.back() == 32) {
      g->slot_s.push_back(s);
      g->slot_first.push_back((int64_t)j);
    }
  }
  const int n_sub = g->n_sub = (int)g->slot_s.size();
  g->slot_first.push_back((int64_t)g->uniq.size());
  // X rows: a slot's rerootings take consecutive rows of its batch
  g->xbase.assign(std::max(n_sub, 1), 0);
  g->max_xrows = 1;
  for (int b0 = 0; b0 < n_sub; b0 += S) {
    int64_t cur = 0;
    for (int k = b0; k < std::min(n_sub, b0 + S); ++k) {
      g->xbase[k] = (int32_t)cur;
      cur += g->slot_first[k + 1] - g->slot_first[k];
    }
    g->max_xrows = std::max(g->max_xrows, cur);
  }
  std::vector<int32_t> pair_slot(g->uniq.size());
  for (int k = 0; k < n_sub; ++k)
    for (int64_t j = g->slot_first[k]; j < g->slot_first[k + 1]; ++j) pair_slot[j] = k;
  // needed pairs per slot and target; their rows
  const size_t M1 = (size_t)M + 1;
  g->tmask.assign((size_t)std::max(n_sub, 1) * M1, 0);
  std::vector<int32_t> mv_slot(n_moves), mv_bit(n_moves), mv_c(n_moves);
  for (int64_t i = 0; i < n_moves; ++i) {
    const int64_t j = std::lower_bound(g->uniq.begin(), g->uniq.end(), sw[i]) - g->uniq.begin();
    const int k = pair_slot[j];
    const int bit = (int)(j - g->slot_first[k]);
    const int32_t v = moves[3 * i + 2];
    const int c = v < 0 ? M : v;
    g->tmask[(size_t)k * M1 + c] |= 1u << bit;
    mv_slot[i] = k;
    mv_bit[i] = bit;
    mv_c[i] = c;
  }
  g->rbase.assign((size_t)std::max(n_sub, 1) * M1, 0);
  g->max_pairs = 1;
  for (int b0 = 0; b0 < n_sub; b0 += S) {
    int64_t cur = 0;
    for (int k = b0; k < std::min(n_sub, b0 + S); ++k)
      for (size_t c = 0; c < M1; ++c) {
        g->rbase[(size_t)k * M1 + c] = (int32_t)cur;
        cur += __builtin_popcount(g->tmask[(size_t)k * M1 + c]);
      }
    if (cur > 0x7fffffff) return 2;
    g->max_pairs = std::max(g->max_pairs, cur);
  }
  // moves grouped by slot, with their wsum rows
  g->mv_start.assign(n_sub + 1, 0);
  for (int64_t i = 0; i < n_moves; ++i) ++g->mv_start[mv_slot[i] + 1];
  for (int k = 0; k < n_sub; ++k) g->mv_start[k + 1] += g->mv_start[k];
  g->rows.assign(n_moves, 0);
  g->idx.assign(n_moves, 0);
  std::vector<int64_t> fill(g->mv_start.begin(), g->mv_start.end() - 1);
  for (int64_t i = 0; i < n_moves; ++i) {
    const int k = mv_slot[i];
    const uint32_t mk = g->tmask[(size_t)k * M1 + mv_c[i]];
    const int64_t j = fill[k]++;
    g->rows[j] = g->rbase[(size_t)k * M1 + mv_c[i]] + __builtin_popcount(mk & ((1u << mv_bit[i]) - 1u));
    g->idx[j] = i;
  }
  return 0;
}

int tbr_plan_create(phy_ctx* ctx, const phy_model* m, int32_t form, int64_t n_pat, int32_t N,
                    const int32_t* children, int32_t root, int64_t n_moves, const int32_t* moves, int32_t s_batch,
                    phy_spr_plan** out) {
  std::unique_ptr<phy_spr_plan> sp(new phy_spr_plan());
  if (!sp->t.load(N, children, root)) return set_err(ctx, PHY_EINVAL, "tbr incremental: not a rooted binary tree");
  const IncrTree& t = sp->t;
  const int M = t.M, S = s_batch, R = m->R;
  sp->tbr = true;
  sp->form = form;
  sp->R = R;
  sp->n_pat = n_pat;
  sp->n_moves = n_moves;
  TbrGroups grp;
  int64_t bad = -1;
  const int gr = tbr_group(t, n_moves, moves, S, &grp, &bad);
  if (gr == 1) return set_err(ctx, PHY_EINVAL, "tbr incremental: invalid move " + std::to_string(bad));
  if (gr == 2) return set_err(ctx, PHY_ERANGE, "tbr incremental: too many pairs in a batch");
  const int n_sub = sp->n_sub = grp.n_sub;
  const std::vector<int32_t>& slot_s = grp.slot_s;
  const std::vector<int64_t>& slot_first = grp.slot_first;
  const std::vector<int64_t>& uniq = grp.uniq;
  const std::vector<uint32_t>& tmask = grp.tmask;
  const std::vector<int32_t>& rbase = grp.rbase;
  const std::vector<int64_t>& rows = grp.rows;
  const std::vector<int64_t>& idx = grp.idx;
  sp->mv_start = grp.mv_start;
  const int64_t max_pairs = grp.max_pairs;
  sp->pl = incr_plan(t, R, n_pat, form, S, n_sub, grp.max_xrows, max_pairs);
  if (phy_topo_create(nullptr, N, children, root, &sp->base) != PHY_OK)
    return set_err(ctx, PHY_EINVAL, "tbr incremental: base topology");
  sp->pl.base = make_plan(sp->base, m, n_pat, form & ~PHY_FUSED, 0);
  const int lmax = sp->pl.lmax;
  // per slot: the repair and insertion of its pruned tree, the pre-order
  // program of its subtree (shared by the slots of one subtree)
  std::vector<IncrSub> sub(n_sub);
  std::vector<std::vector<int32_t>> tprog(n_sub);
  std::vector<int32_t> tdepth(n_sub, 0);
  {
    const unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> th;
    for (unsigned w = 0; w < std::min<unsigned>(nt, std::max(n_sub, 1)); ++w)
      th.emplace_back([&, w] {
        std::vector<int32_t> targets;
        std::vector<char> enter, tgt(M, 0);
        for (int k = (int)w; k < n_sub; k += (int)nt) {
          targets.clear();
          for (int c = 0; c < M; ++c)
            if (tmask[(size_t)k * (M + 1) + c]) targets.push_back(c);
          if (tmask[(size_t)k * (M + 1) + M]) targets.push_back(t.sibling(slot_s[k]));  // keeping
          for (int32_t c : targets) tgt[c] = 1;
          incr_enter(t, slot_s[k], targets, &enter);
          incr_prepare(t, slot_s[k], k % S, lmax, &sub[k], &enter, &tgt);
          for (int32_t c : targets) tgt[c] = 0;
          if (slot_s[k] >= N) phy_compile_up_program(N, t.ch, t.size, slot_s[k], &tprog[k], &tdepth[k]);
        }
      });
    for (auto& x : th) x.join();
  }
  const size_t ns = (size_t)std::max(n_sub, 1);
  std::vector<int32_t> hprog(ns * (N - 1) * 4, 0), hmeta(ns * kIncrMeta, 0), href(ns * M, 0),
      hops(ns * lmax * 4, 0), hxp(ns * 2, 0), htprog(ns * (N - 1) * 4, 0), htmeta(ns * kIncrMeta, 0),
      hwslot(ns * M, -1), hkx(ns, 0);
  sp->batch_depth.assign((n_sub + S - 1) / S, 1);
  sp->batch_rdepth.assign((n_sub + S - 1) / S, 1);
  for (int k = 0; k < n_sub; ++k) {
    const IncrSub& q = sub[k];
    const int32_t s = slot_s[k];
    std::copy(q.prog.begin(), q.prog.end(), hprog.begin() + (size_t)k * (N - 1) * 4);
    std::copy(q.ref.begin(), q.ref.end(), href.begin() + (size_t)k * M);
    std::copy(q.ops.begin(), q.ops.end(), hops.begin() + (size_t)k * lmax * 4);
    std::copy(tprog[k].begin(), tprog[k].end(), htprog.begin() + (size_t)k * (N - 1) * 4);
    int32_t* mk = &hmeta[(size_t)k * kIncrMeta];
    mk[0] = (k % S) * lmax;
    mk[1] = (int32_t)q.ops.size() / 4;
    mk[2] = q.x;
    mk[3] = s;
    mk[4] = q.n_ops;
    mk[5] = q.depth;
    int32_t* tk = &htmeta[(size_t)k * kIncrMeta];
    // a slot whose only rerooting is w = s needs no pass over its subtree
    const bool only_s = slot_first[k + 1] - slot_first[k] == 1 && uniq[slot_first[k]] % M == s;
    tk[0] = only_s ? 0 : (int32_t)tprog[k].size() / 4;
    tk[1] = s;
    tk[2] = grp.xbase[k];
    tk[3] = tdepth[k];
    tk[4] = t.sibling(s);
    tk[5] = t.parent[t.parent[s]] >= 0 ? 1 : 0;
    hxp[2 * k] = q.x;
    hxp[2 * k + 1] = q.p;
    hkx[k] = t.sibling(s);
    for (int64_t j = slot_first[k]; j < slot_first[k + 1]; ++j)
      hwslot[(size_t)k * M + (size_t)(uniq[j] % M)] = grp.xbase[k] + (int32_t)(j - slot_first[k]);
    sp->batch_depth[k / S] = std::max(sp->batch_depth[k / S], q.depth);
    if (!only_s) sp->batch_rdepth[k / S] = std::max(sp->batch_rdepth[k / S], tdepth[k]);
  }
  {
    int lds_max = 64 * 1024, dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
    const size_t esz = sp->pl.esz;
    for (size_t b = 0; b < sp->batch_depth.size(); ++b)
      if (tbr_insert_lds(R, esz, sp->batch_depth[b]) > (size_t)lds_max ||
          tbr_reroot_lds(R, esz, sp->batch_rdepth[b]) > (size_t)lds_max)
        return set_err(ctx, PHY_ERANGE, "tbr incremental: LDS stack too deep");
  }
  std::vector<int32_t> fifo(phy_fifo_padded(N), sp->base->down_fifo[0]);
  std::copy(sp->base->down_fifo.begin(), sp->base->down_fifo.end(), fifo.begin());
  size_t o = 0;
  auto put = [&](size_t bytes) {
    const size_t at = o;
    o = align256(o + std::max<size_t>(bytes, 8));
    return at;
  };
  sp->o_dprog = put(sp->base->down_prog.size() * 4);
  sp->o_dfifo = put(fifo.size() * 4);
  sp->o_prog = put(hprog.size() * 4);
  sp->o_meta = put(hmeta.size() * 4);
  sp->o_ref = put(href.size() * 4);
  sp->o_ops = put(hops.size() * 4);
  sp->o_xp = put(hxp.size() * 4);
  sp->o_rows = put(rows.size() * 8);
  sp->o_idx = put(idx.size() * 8);
  sp->o_tprog = put(htprog.size() * 4);
  sp->o_tmeta = put(htmeta.size() * 4);
  sp->o_wslot = put(hwslot.size() * 4);
  sp->o_tmask = put(tmask.size() * 4);
  sp->o_rbase = put(rbase.size() * 4);
  sp->o_kx = put(hkx.size() * 4);
  PHY_HIP(ctx, hipMalloc((void**)&sp->dmem, o));
  auto up = [&](size_t off, const void* src, size_t bytes) {
    return bytes ? hipMemcpy(sp->dmem + off, src, bytes, hipMemcpyHostToDevice) : hipSuccess;
  };
  PHY_HIP(ctx, up(sp->o_dprog, sp->base->down_prog.data(), sp->base->down_prog.size() * 4));
  PHY_HIP(ctx, up(sp->o_dfifo, fifo.data(), fifo.size() * 4));
  PHY_HIP(ctx, up(sp->o_prog, hprog.data(), hprog.size() * 4));
  PHY_HIP(ctx, up(sp->o_meta, hmeta.data(), hmeta.size() * 4));
  PHY_HIP(ctx, up(sp->o_ref, href.data(), href.size() * 4));
  PHY_HIP(ctx, up(sp->o_ops, hops.data(), hops.size() * 4));
  PHY_HIP(ctx, up(sp->o_xp, hxp.data(), hxp.size() * 4));
  PHY_HIP(ctx, up(sp->o_rows, rows.data(), rows.size() * 8));
  PHY_HIP(ctx, up(sp->o_idx, idx.data(), idx.size() * 8));
  PHY_HIP(ctx, up(sp->o_tprog, htprog.data(), htprog.size() * 4));
  PHY_HIP(ctx, up(sp->o_tmeta, htmeta.data(), htmeta.size() * 4));
  PHY_HIP(ctx, up(sp->o_wslot, hwslot.data(), hwslot.size() * 4));
  PHY_HIP(ctx, up(sp->o_tmask, tmask.data(), tmask.size() * 4));
  PHY_HIP(ctx, up(sp->o_rbase, rbase.data(), rbase.size() * 4));
  PHY_HIP(ctx, up(sp->o_kx, hkx.data(), hkx.size() * 4));
  sp->base->d_down_prog = reinterpret_cast<int4*>(sp->dmem + sp->o_dprog);
  sp->base->d_down_fifo = reinterpret_cast<int32_t*>(sp->dmem + sp->o_dfifo);
  *out = sp.release();
  return PHY_OK;
}

int check_plan_model(phy_ctx* ctx, const phy_model* m) {
  if (m->K != 4) return set_err(ctx, PHY_EINVAL, "spr incremental: 4-state models only");
  if (m->R != 1 && m->R != 2 && m->R != 4 && m->R != 8)
    return set_err(ctx, PHY_ERANGE, "spr incremental: n_rates must be 1, 2, 4 or 8");
  return PHY_OK;
}

}  // namespace

extern "C" {

int phy_lk_spr_incr_work_size(int32_t N, const int32_t* children, int32_t root, const phy_model* m,
                              int64_t n_pat, int32_t form, int32_t s_batch, size_t* bytes) {
  if (!m || !bytes || !children || N < 3 || n_pat < 1 || s_batch < 1) return PHY_EINVAL;
  try {
    IncrTree t;
    if (!t.load(N, children, root)) return PHY_EINVAL;
    *bytes = incr_plan(t, m->R, n_pat, form, s_batch).total;
    return PHY_OK;
  } catch (...) {
    return PHY_ENOMEM;
  }
}

int phy_spr_plan_create(phy_ctx* ctx, const phy_model* m, int32_t form, int64_t n_pat, int32_t N,
                        const int32_t* children, int32_t root, int64_t n_moves, const int32_t* moves,
                        int32_t s_batch, phy_spr_plan** out) {
  if (!ctx || !m || !out) return PHY_EINVAL;
  *out = nullptr;
  if (N < 3 || n_pat < 1 || s_batch < 1 || n_moves < 0 || !children || (n_moves > 0 && !moves))
    return set_err(ctx, PHY_EINVAL, "spr plan: bad argument");
  int rc = check_plan_model(ctx, m);
  if (rc) return rc;
  try {
    return plan_create(ctx, m, form, n_pat, N, children, root, n_moves, moves, s_batch, out);
  } catch (...) {
    return set_err(ctx, PHY_ENOMEM, "spr plan: host allocation or thread start failed");
  }
}

int phy_spr_plan_work_size(const phy_spr_plan* plan, size_t* bytes) {
  if (!plan || !bytes) return PHY_EINVAL;
  *bytes = plan->pl.total;
  return PHY_OK;
}

int phy_spr_plan_run(phy_ctx* ctx, const phy_spr_plan* plan, const phy_model* m, const uint64_t* tips,
                     int64_t stride, const uint32_t* weights, const double* d_blen, double* d_lnl,
                     double* d_base_lnl, void* d_work, size_t work_bytes) {
  if (!ctx || !plan || !m || !tips || !weights || !d_blen || !d_lnl || !d_work) return PHY_EINVAL;
  int rc = check_plan_model(ctx, m);
  if (rc) return rc;
  if (m->R != plan->R) return set_err(ctx, PHY_EINVAL, "spr plan: model has another number of rate classes");
  if (stride < phy_words(plan->n_pat)) return set_err(ctx, PHY_EINVAL, "spr plan: tip stride < words");
  if (work_bytes < plan->pl.total) return set_err(ctx, PHY_EINVAL, "spr plan: workspace too small");
  return incr_dispatch(ctx, plan, m, tips, stride, weights, d_blen, d_lnl, d_base_lnl, static_cast<char*>(d_work));
}

void phy_spr_plan_destroy(phy_spr_plan* plan) { delete plan; }

int phy_lk_spr_scores_incr(phy_ctx* ctx, const phy_model* m, int32_t form, int64_t n_pat,
                           const uint64_t* tips, int64_t stride, const uint32_t* weights, int32_t N,
                           const int32_t* children, int32_t root, const double* blen, int64_t n_moves,
                           const int32_t* moves, int32_t s_batch, double* d_lnl, double* d_base_lnl,
                           void* d_work, size_t work_bytes) {
  int rc = check_scores_args(ctx, m, n_pat, tips, stride, weights, N, children, blen, n_moves, moves,
                             s_batch, d_lnl, d_work, root, "spr incremental");
  if (rc) return rc;
  rc = check_plan_model(ctx, m);
  if (rc) return rc;
  try {
    IncrTree t;
    if (!t.load(N, children, root)) return set_err(ctx, PHY_EINVAL, "spr incremental: not a rooted binary tree");
    const IncrPlan lay = incr_plan(t, m->R, n_pat, form, s_batch);
    if (work_bytes < lay.total) return set_err(ctx, PHY_EINVAL, "spr incremental: workspace too small");
    phy_spr_plan* sp = nullptr;
    rc = plan_create(ctx, m, form, n_pat, N, children, root, n_moves, moves, s_batch, &sp);
    if (rc) return rc;
    std::unique_ptr<phy_spr_plan> hold(sp);
    char* work = static_cast<char*>(d_work);
    double* dbl = reinterpret_cast<double*>(work + lay.off_blh);
    PHY_HIP(ctx, hipMemcpyAsync(dbl, blen, (size_t)t.M * 8, hipMemcpyHostToDevice, ctx->stream));
    rc = incr_dispatch(ctx, sp, m, tips, stride, weights, dbl, d_lnl, d_base_lnl, work);
    // the plan's device arrays are freed on return
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (rc) return rc;
    if (e != hipSuccess) return set_err(ctx, PHY_EDEVICE, std::string("spr incremental: ") + hipGetErrorString(e));
    return PHY_OK;
  } catch (...) {
    return set_err(ctx, PHY_ENOMEM, "spr incremental: host allocation or thread start failed");
  }
}

int phy_lk_tbr_incr_work_size(int32_t N, const int32_t* children, int32_t root, const phy_model* m,
                              int64_t n_pat, int32_t form, int64_t n_moves, const int32_t* moves, int32_t s_batch,
                              size_t* bytes) {
  if (!m || !bytes || !children || N < 3 || n_pat < 1 || s_batch < 1 || n_moves < 0 || (n_moves > 0 && !moves))
    return PHY_EINVAL;
  try {
    IncrTree t;
    if (!t.load(N, children, root)) return PHY_EINVAL;
    TbrGroups g;
    int64_t bad = -1;
    const int gr = tbr_group(t, n_moves, moves, s_batch, &g, &bad);
    if (gr) return gr == 1 ? PHY_EINVAL : PHY_ERANGE;
    *bytes = incr_plan(t, m->R, n_pat, form, s_batch, g.n_sub, g.max_xrows, g.max_pairs).total;
    return PHY_OK;
  } catch (...) {
    return PHY_ENOMEM;
  }
}

int phy_tbr_plan_create(phy_ctx* ctx, const phy_model* m, int32_t form, int64_t n_pat, int32_t N,
                        const int32_t* children, int32_t root, int64_t n_moves, const int32_t* moves,
                        int32_t s_batch, phy_spr_plan** out) {
  if (!ctx || !m || !out) return PHY_EINVAL;
  *out = nullptr;
  if (N < 3 || n_pat < 1 || s_batch < 1 || n_moves < 0 || !children || (n_moves > 0 && !moves))
    return set_err(ctx, PHY_EINVAL, "tbr plan: bad argument");
  int rc = check_plan_model(ctx, m);
  if (rc) return rc;
  try {
    return tbr_plan_create(ctx, m, form, n_pat, N, children, root, n_moves, moves, s_batch, out);
  } catch (...) {
    return set_err(ctx, PHY_ENOMEM, "tbr plan: host allocation or thread start failed");
  }
}

int phy_lk_tbr_scores_incr(phy_ctx* ctx, const phy_model* m, int32_t form, int64_t n_pat,
                           const uint64_t* tips, int64_t stride, const uint32_t* weights, int32_t N,
                           const int32_t* children, int32_t root, const double* blen, int64_t n_moves,
                           const int32_t* moves, int32_t s_batch, double* d_lnl, double* d_base_lnl,
                           void* d_work, size_t work_bytes) {
  int rc = check_scores_args(ctx, m, n_pat, tips, stride, weights, N, children, blen, n_moves, moves,
                             s_batch, d_lnl, d_work, root, "tbr incremental");
  if (rc) return rc;
  rc = check_plan_model(ctx, m);
  if (rc) return rc;
  try {
    phy_spr_plan* sp = nullptr;
    rc = tbr_plan_create(ctx, m, form, n_pat, N, children, root, n_moves, moves, s_batch, &sp);
    if (rc) return rc;
    std::unique_ptr<phy_spr_plan> hold(sp);
    if (work_bytes < sp->pl.total) return set_err(ctx, PHY_EINVAL, "tbr incremental: workspace too small");
    char* work = static_cast<char*>(d_work);
    double* dbl = reinterpret_cast<double*>(work + sp->pl.off_blh);
    PHY_HIP(ctx, hipMemcpyAsync(dbl, blen, (size_t)sp->t.M * 8, hipMemcpyHostToDevice, ctx->stream));
    rc = incr_dispatch(ctx, sp, m, tips, stride, weights, dbl, d_lnl, d_base_lnl, work);
    // the plan's device arrays are freed on return
    const hipError_t e = hipStreamSynchronize(ctx->stream);
    if (rc) return rc;
    if (e != hipSuccess) return set_err(ctx, PHY_EDEVICE, std::string("tbr incremental: ") + hipGetErrorString(e));
    return PHY_OK;
  } catch (...) {
    return set_err(ctx, PHY_ENOMEM, "tbr incremental: host allocation or thread start failed");
  }
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

// model constants: pi and the class probabilities
inline const double* model_pi(const phy_model* m) { return m->d_consts + 2 * 4 * 4 + 4 + m->R; }

// P (and P', P'' if with_derivs) of nb <= 3 branches into the model scratch,
// followed by `extra` bytes for the caller; returns the scratch base.
template <typename real, int R>
static int node_pmats(phy_ctx* ctx, phy_model* m, HostLens hl, int nb, bool with_derivs,
                      size_t extra, real** P, unsigned char** rest) {
  const size_t mats = align256((size_t)(with_derivs ? 3 : 1) * nb * R * 16 * sizeof(real));
  int rc = model_scratch(m, std::max<size_t>(mats + extra, 4096));
  if (rc) return rc;
  *P = (real*)m->d_scratch;
  *rest = (unsigned char*)m->d_scratch + mats;
  const int64_t nm = (int64_t)nb * R * 16;
  real* P1 = with_derivs ? *P + nm : nullptr;
  real* P2 = with_derivs ? *P + 2 * nm : nullptr;
  hipLaunchKernelGGL((k_pmats<real>), dim3((unsigned)((nm + 255) / 256)), dim3(256), 0, ctx->stream,
                     m->d_consts, 4, R, (const double*)nullptr, hl, nb, *P, P1, P2);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

template <typename real, int R>
static int node_median(phy_ctx* ctx, phy_model* m, int64_t n, const void* x, const int32_t* ex,
                       double tx, const void* y, const int32_t* ey, double ty, const void* z,
                       const int32_t* ez, double tz, void* out, int32_t* eout, int64_t pitch,
                       const uint32_t* w, double* lnl) {
  const int64_t nb = (n + 255) / 256;
  real* Pm;
  unsigned char* rest;
  int rc = node_pmats<real, R>(ctx, m, HostLens{{tx, ty, tz}}, z ? 3 : 2, false,
                               lnl ? (size_t)nb * 4 * 8 : 0, &Pm, &rest);
  if (rc) return rc;
  double* bs = lnl ? (double*)rest : nullptr;
  const double* d_pi = model_pi(m);
  hipLaunchKernelGGL((k_lk_median<real, R>), dim3((unsigned)nb), dim3(256), 0, ctx->stream, Pm,
                     (const real*)x, ex, (const real*)y, ey, (const real*)z, ez, n, pitch,
                     (real*)out, eout, d_pi, d_pi + 4, w, bs);
  PHY_LAUNCHED(ctx);
  if (lnl) {
    hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, ctx->stream, bs, nb * 4, (int64_t)1, lnl);
    PHY_LAUNCHED(ctx);
  }
  return PHY_OK;
}

template <typename real, int R>
static int node_root(phy_ctx* ctx, phy_model* m, int64_t n, const void* L, const int32_t* e,
                     int64_t pitch, const uint32_t* w, double* lnl) {
  const int64_t nb = (n + 255) / 256;
  int rc = model_scratch(m, std::max<size_t>((size_t)nb * 4 * 8, 4096));
  if (rc) return rc;
  double* bs = (double*)m->d_scratch;  // block sums
  const double* d_pi = model_pi(m);
  hipLaunchKernelGGL((k_lk_root<real, R>), dim3((unsigned)nb), dim3(256), 0, ctx->stream,
                     (const real*)L, e, n, pitch, d_pi, d_pi + 4, w, bs);
  PHY_LAUNCHED(ctx);
  hipLaunchKernelGGL(k_reduce_sum, dim3(1), dim3(256), 0, ctx->stream, bs, nb * 4, (int64_t)1, lnl);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

// lnL, dlnL/dt, d2lnL/dt2 of the edge of length t between x and y, into the
// device array out[3]
template <typename real, int R>
static int node_edge(phy_ctx* ctx, phy_model* m, int64_t n, const void* x, const int32_t* ex,
                     const void* y, const int32_t* ey, int64_t pitch, double t, const uint32_t* w,
                     double* out) {
  const int64_t nb = (n + 255) / 256, slots = nb * 4;
  real* Pm;
  unsigned char* rest;
  int rc = node_pmats<real, R>(ctx, m, HostLens{{t, t, t}}, 1, true, (size_t)3 * slots * 8, &Pm, &rest);
  if (rc) return rc;
  double* bs = (double*)rest;
  const double* d_pi = model_pi(m);
  hipLaunchKernelGGL((k_lk_edge<real, R>), dim3((unsigned)nb), dim3(256), 0, ctx->stream, Pm,
                     Pm + R * 16, Pm + 2 * R * 16, (const real*)x, ex, (const real*)y, ey, n, pitch,
                     d_pi, d_pi + 4, w, bs, slots);
  PHY_LAUNCHED(ctx);
  hipLaunchKernelGGL(k_reduce_rows, dim3(3), dim3(256), 0, ctx->stream, bs, slots, out);
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
  const void* z = nullptr;
  const int32_t* ez = nullptr;
  const uint32_t* w = nullptr;
  double* lnl = nullptr;
  if (form & PHY_F32) {
    PHY_DISPATCH_R(node_median, float, ctx, m, n, x, ex, tx, y, ey, ty, z, ez, 0.0, out, eout, pitch, w, lnl);
  }
  PHY_DISPATCH_R(node_median, double, ctx, m, n, x, ex, tx, y, ey, ty, z, ez, 0.0, out, eout, pitch, w, lnl);
}

int phy_lk_median3(phy_ctx* ctx, const phy_model* mc, int32_t form, int64_t n, const void* x,
                   const int32_t* ex, double tx, const void* y, const int32_t* ey, double ty,
                   const void* z, const int32_t* ez, double tz, void* out, int32_t* eout,
                   int64_t pitch, const uint32_t* w, double* lnl) {
  if (!ctx || !mc) return PHY_EINVAL;
  phy_model* m = const_cast<phy_model*>(mc);
  if (m->K != 4 || n < 1 || pitch < n || !x || !ex || !y || !ey || !z || !ez ||
      (out == nullptr) != (eout == nullptr) || (!out && !lnl))
    return set_err(ctx, PHY_EINVAL, "lk_median3: bad argument");
  if (form & PHY_F32) {
    PHY_DISPATCH_R(node_median, float, ctx, m, n, x, ex, tx, y, ey, ty, z, ez, tz, out, eout, pitch, w, lnl);
  }
  PHY_DISPATCH_R(node_median, double, ctx, m, n, x, ex, tx, y, ey, ty, z, ez, tz, out, eout, pitch, w, lnl);
}

int phy_lk_edge(phy_ctx* ctx, const phy_model* mc, int32_t form, int64_t n, const void* x,
                const int32_t* ex, const void* y, const int32_t* ey, int64_t pitch, double t,
                const uint32_t* w, double* d_out) {
  if (!ctx || !mc) return PHY_EINVAL;
  phy_model* m = const_cast<phy_model*>(mc);
  if (m->K != 4 || n < 1 || pitch < n || !x || !ex || !y || !ey || !d_out || !(t >= 0.0) ||
      !std::isfinite(t))
    return set_err(ctx, PHY_EINVAL, "lk_edge: bad argument");
  if (form & PHY_F32) { PHY_DISPATCH_R(node_edge, float, ctx, m, n, x, ex, y, ey, pitch, t, w, d_out); }
  PHY_DISPATCH_R(node_edge, double, ctx, m, n, x, ex, y, ey, pitch, t, w, d_out);
}

int phy_lk_adjust3(phy_ctx* ctx, const phy_model* mc, int32_t form, int64_t n, const void* x,
                   const int32_t* ex, const void* y, const int32_t* ey, int64_t pitch,
                   const uint32_t* w, double t0, double t_lo, double t_hi, int32_t max_iter,
                   double tol, double* t_out, double* lnl_out, int32_t* n_evals) {
  if (!ctx || !mc) return PHY_EINVAL;
  if (!t_out || !(t_lo >= 0.0) || !(t_hi >= t_lo) || !std::isfinite(t_hi) || max_iter < 0 ||
      !(tol >= 0.0) || !std::isfinite(t0))
    return set_err(ctx, PHY_EINVAL, "lk_adjust3: bad argument");
  // the edge's lnL and derivatives at t, back on the host
  double* d_res = nullptr;
  PHY_HIP(ctx, hipMallocAsync((void**)&d_res, 3 * sizeof(double), ctx->stream));
  int evals = 0;
  auto eval = [&](double t, double* f) -> int {
    int rc = phy_lk_edge(ctx, mc, form, n, x, ex, y, ey, pitch, t, w, d_res);
    if (rc) return rc;
    PHY_HIP(ctx, hipMemcpyAsync(f, d_res, 3 * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
    PHY_HIP(ctx, hipStreamSynchronize(ctx->stream));
    ++evals;
    return PHY_OK;
  };
  auto clamp = [&](double t) { return std::min(std::max(t, t_lo), t_hi); };
  double t = clamp(t0), f[3];
  int rc = eval(t, f);
  // Newton-Raphson on lnL(t) with step halving; where d2 >= 0 (not a
  // local maximum) a gradient step of size t (or 0.1 at t = 0) is taken
  for (int it = 0; rc == PHY_OK && it < max_iter; ++it) {
    double step = f[2] < 0.0 ? -f[1] / f[2] : (f[1] > 0.0 ? std::max(t, 0.1) : -0.5 * t);
    double tn = clamp(t + step), fn[3] = {0, 0, 0};
    bool better = false;
    for (int h = 0; h < 30 && rc == PHY_OK; ++h) {
      if (tn == t) break;
      rc = eval(tn, fn);
      if (rc) break;
      if (fn[0] >= f[0]) {
        better = true;
        break;
      }
      step *= 0.5;
      tn = clamp(t + step);
    }
    if (!better) break;
    const double dt = std::fabs(tn - t);
    t = tn;
    f[0] = fn[0], f[1] = fn[1], f[2] = fn[2];
    if (dt <= tol * std::max(t, 1e-8)) break;
  }
  hipError_t fe = hipFreeAsync(d_res, ctx->stream);
  if (rc) return rc;
  if (fe != hipSuccess) return set_err(ctx, PHY_EDEVICE, "lk_adjust3: hipFreeAsync failed");
  *t_out = t;
  if (lnl_out) *lnl_out = f[0];
  if (n_evals) *n_evals = evals;
  return PHY_OK;
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
