// Batched 20-state (amino-acid) Felsenstein pruning on BF16 MFMA.
//
// Many independent trees of uneven taxon and pattern counts, each with its own
// GTR-20 model and Gamma rates, are evaluated in one persistent launch:
// workgroups take (tree, pattern tile) items from an atomic queue until it is
// empty (every workgroup reaches the exit).  A workgroup of WAVES waves owns
// 16*WAVES patterns of one tree and walks that tree's post-order program
// (phy_topo_create's Sethi-Ullman order); the waves stay in lock step, so the
// transition-matrix operands of the node being computed are staged once into
// LDS and shared.
//
// Per node, child partial x (16 patterns x 20 states, per rate) times P^T:
//   D[i][p] = sum_j P_r[i][j] x[j][p]     v_mfma_f32_16x16x32_bf16
// with A = P (16 output states x 32 source states, zero padded) and B = x (32
// source states x 16 patterns).  The output rows are a permutation rho of the
// states chosen so that lane l (pattern l%16, quarter q = l/16) receives
// states 8q..8q+7 -- exactly the k-slice that lane supplies as B operand of
// the next product:  rho(4q+v) = 8q+v (block 0), rho(16+4q+v) = 8q+4+v
// (block 1).  Partials therefore never move between lanes; combining the two
// children, per-pattern power-of-two rescaling and the root sum are FP32 /
// FP64 VALU work on 32 values per lane.  Spilled partials are stored as the
// BF16 operand they will be used as (no extra rounding).
//
// Precision: MFMA inputs are BF16 (8-bit mantissa), accumulation, partials and
// exponents FP32 / int32, the root log-sum FP64.  P(t) = U diag(exp(lambda t))
// Ui is formed in FP64 from the host eigen-decomposition of each tree's Q.
#include <algorithm>
#include <cstdlib>
#include <cmath>
#include <numeric>

#include "common.hpp"
#include "model_host.hpp"
#include "stream.hpp"

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

namespace {

constexpr int KAA = 20;
constexpr int OPB = 64;  // uint4 (16 B) per lane-ordered operand block (1 KiB)

struct AADev {
  const int4* prog;
  const int64_t* prog_off;  // entries
  const int32_t* fifo;
  const int64_t* fifo_off;
  const int32_t* n_leaves;
  const uint8_t* tips;
  const int64_t* tips_off;  // bytes
  const int64_t* pitch;     // padded patterns of each tree
  const uint32_t* wts;
  const int64_t* w_off;
  const uint4* pops;
  const int64_t* pop_off;   // uint4 units
  const double* pi;         // [T][20]
  const double* probs;      // [T][R]
  const int2* items;        // (tree, tile)
  int n_items;
  double* part;
  const int64_t* part_off;
  unsigned* counter;
  int depth;                // LDS stack entries per wave
};

// source state of output row i of block b
__host__ __device__ __forceinline__ int rho(int b, int i) {
  const int q = i >> 2, v = i & 3;
  return 8 * q + 4 * b + v;
}

// --- P operands: one wave per (tree, branch, rate), four to a workgroup ---
constexpr int kPopsPerBlock = 4;

__global__ __launch_bounds__(64 * kPopsPerBlock) void k_aa_pops(
    const double* __restrict__ eig,  // [T][U 400|lam 20|Ui 400]
    const double* __restrict__ rates,  // [T][R]
    const double* __restrict__ blen,   // flat
    const int64_t* __restrict__ blen_off,
    const int32_t* __restrict__ br_tree,  // flat branch -> tree
    int R, int64_t n_pairs, const int64_t* __restrict__ pop_off, uint4* __restrict__ pops) {
  const int wv = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int64_t pair = (int64_t)blockIdx.x * kPopsPerBlock + wv;
  const bool valid = pair < n_pairs;
  const int64_t pc = valid ? pair : n_pairs - 1;  // an idle wave reads in bounds, writes nothing
  const int64_t gb = pc / R;  // flat branch index
  const int r = (int)(pc % R);
  const int t = br_tree[gb];
  const int64_t b = gb - blen_off[t];
  const double* U = eig + (size_t)t * (2 * KAA * KAA + KAA);
  const double* lam = U + KAA * KAA;
  const double* Ui = lam + KAA;
  const double rt = rates[(size_t)t * R + r] * blen[gb];
  __shared__ double ex[kPopsPerBlock][KAA];
  __shared__ double P[kPopsPerBlock][KAA * KAA];
  if (l < KAA) ex[wv][l] = exp(lam[l] * rt);
  __syncthreads();
  for (int e = l; e < KAA * KAA; e += 64) {
    const int i = e / KAA, j = e % KAA;
    double s = 0.0;
    for (int m = 0; m < KAA; ++m) s += U[i * KAA + m] * ex[wv][m] * Ui[m * KAA + j];
    if (!(rt >= 1e-10)) s = (i == j) ? 1.0 : 0.0;  // compose rule, lib/mlmodel.c:339-341
    P[wv][e] = s;
  }
  __syncthreads();
  if (!valid) return;
  uint4* dst = pops + pop_off[t] + ((b * R + r) * 2) * OPB;
  for (int blk = 0; blk < 2; ++blk) {
    // lane l: output row l%16 -> state rho(blk, l%16); k = 8*(l/16) + v
    const int srow = rho(blk, l & 15);
    bf16x8 a;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const int k = 8 * (l >> 4) + v;
      const float val = (srow < KAA && k < KAA) ? (float)P[wv][srow * KAA + k] : 0.0f;
      a[v] = (__bf16)val;
    }
    dst[blk * OPB + l] = *reinterpret_cast<uint4*>(&a);
  }
}

__device__ __forceinline__ bf16x8 to_bf16x8(const float* x) {
  bf16x8 r;
#pragma unroll
  for (int v = 0; v < 8; ++v) r[v] = (__bf16)x[v];
  return r;
}

__device__ __forceinline__ bf16x8 tip_operand(uint32_t code, int q) {
  bf16x8 r;
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const int s = 8 * q + v;
    const bool on = (code < (uint32_t)KAA) ? (s == (int)code) : (s < KAA);
    r[v] = (__bf16)(on ? 1.0f : 0.0f);
  }
  return r;
}

template <int WAVES, int R, int LV>
__global__ __launch_bounds__(64 * WAVES) void k_aa_prune(AADev a) {
  constexpr int NT = 64 * WAVES;
  constexpr int STAGE = 2 * R * 2 * OPB;  // uint4 per node: [c][r][blk][lane]
  constexpr int NBLK = 2 * R * 2;         // 1 KiB operand blocks per node
  __shared__ uint4 stage[2][STAGE];       // double buffer: node i and node i+1
  __shared__ uint4 tipop[(KAA + 1) * 4];  // B operand of a tip: [code 0..20][quarter]
  __shared__ int s_item;
  extern __shared__ uint4 dyn[];
  // the wave index is uniform: readfirstlane keeps the operand DMA loop scalar
  const int tid = threadIdx.x, wave = __builtin_amdgcn_readfirstlane(tid >> 6), l = tid & 63,
            q = l >> 4, pl = l & 15;
  // per-wave stack: the bottom LV levels in registers (selected by the
  // uniform stack pointer with value selects), the rest in LDS:
  // (depth - LV) x (R operands of 64 lanes) + exponents
  const int rows = a.depth > LV ? a.depth - LV : 0;
  uint4* stk = dyn + (size_t)wave * rows * (R * 64 + 16);
  uint4 rs[LV > 0 ? LV : 1][R];
  int rse[LV > 0 ? LV : 1];
  // the 21 tip operands (a state 0..19, or missing = all 20 states) for each
  // quarter of the lanes, built once; a tip child is then one LDS read
  for (int e = tid; e < (KAA + 1) * 4; e += NT) {
    const bf16x8 b = tip_operand((uint32_t)(e >> 2), e & 3);
    tipop[e] = *reinterpret_cast<const uint4*>(&b);
  }

  for (;;) {
    __syncthreads();
    if (tid == 0) s_item = (int)atomicAdd(a.counter, 1u);
    __syncthreads();
    // uniform in fact; readfirstlane tells the compiler, so everything
    // derived from the item (tree offsets, fifo ids, operand addresses) is
    // scalar
    const int item = __builtin_amdgcn_readfirstlane(s_item);
    if (item >= a.n_items) break;
    const int2 it = a.items[item];
    const int t = __builtin_amdgcn_readfirstlane(it.x), tile = __builtin_amdgcn_readfirstlane(it.y);
    const int64_t pitch = a.pitch[t];
    const int64_t p = (int64_t)tile * (16 * WAVES) + wave * 16 + pl;  // < pitch by construction
    const int N = a.n_leaves[t];
    const int4* prog = a.prog + a.prog_off[t];
    const int32_t* fifo = a.fifo + a.fifo_off[t];
    const uint8_t* tips = a.tips + a.tips_off[t];
    const uint4* pops = a.pops + a.pop_off[t];
    int sp = 0;
    float acc[R][8];
    int acc_e = 0;
#pragma unroll
    for (int r = 0; r < R; ++r)
#pragma unroll
      for (int v = 0; v < 8; ++v) acc[r][v] = 0.f;

    // Every global load issued while node i computes (the operand DMA of node
    // i+1, the tip codes of node i+1, fifo id refills) is consumed at node i+1
    // or later, after the node's closing vmcnt(0) and barrier.  Nothing loaded
    // from global memory is waited on in the middle of a node, so the DMA
    // latency overlaps a whole node of MFMA and VALU work.
    //
    // operands of a node's two children straight into LDS (global_load_lds,
    // 16 B per lane: one 1 KiB block per wave instruction), no registers
    auto dma_stage = [&](const int4 o, int buf) {
#pragma unroll
      for (int k = 0; k < (NBLK + WAVES - 1) / WAVES; ++k) {
        const int bk = wave + k * WAVES;  // uniform
        if (NBLK % WAVES != 0 && bk >= NBLK) break;
        const int c = bk / (R * 2), rem = bk % (R * 2);
        const int child = c ? o.z : o.y;
        const uint4* src = pops + ((size_t)child * R * 2 + rem) * OPB + l;
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)src,
                                         (__attribute__((address_space(3))) void*)(&stage[buf][bk * OPB]),
                                         16, 0, 0);
      }
    };
    // leaf ids of the stream, 128 at a time in two lane vectors (entries
    // fbase .. fbase+127; the tree's fifo is padded by 128 ids)
    int fbase = 0, fpos = 0;
    int fcur = fifo[l], fnx = fifo[64 + l];
    auto fid = [&](int pos) -> int {
      const int k = pos - fbase;  // uniform, 0 <= k < 128
      return k < 64 ? __builtin_amdgcn_readlane(fcur, k) : __builtin_amdgcn_readlane(fnx, k - 64);
    };
    auto tip_at = [&](int pos) -> uint32_t { return tips[(size_t)fid(pos) * pitch + p]; };
    // tip codes of a node's leaf children: the first leaf of the stream goes
    // to child a when a is a leaf, else to child b
    auto leaf_codes = [&](const int4 o, int pos, uint32_t& ca, uint32_t& cb) {
      ca = tip_at(pos);
      cb = tip_at(pos + ((o.w & 3) == SRC_LEAF ? 1 : 0));
    };
    uint32_t code_a, code_b;
    {
      const int4 o0 = prog[0];
      leaf_codes(o0, 0, code_a, code_b);
      dma_stage(o0, 0);
    }
    __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): the DMA has landed
    __syncthreads();

    for_each_entry_ahead(prog, N - 1, [&](int i, const int4 op, const int4 on) {
      const int buf = i & 1;
      const bool more = i + 1 < N - 1;
      const int fl = op.w;
      const int nl = ((fl & 3) == SRC_LEAF ? 1 : 0) + (((fl >> 2) & 3) == SRC_LEAF ? 1 : 0);
      if (fpos - fbase >= 64) {  // fnx landed at least one node ago
        fbase += 64;
        fcur = fnx;
        fnx = fifo[fbase + 64 + l];
      }
      if (more) dma_stage(on, buf ^ 1);  // in flight while node i computes
      const uint32_t ca = code_a, cb = code_b;
      leaf_codes(on, fpos + nl, code_a, code_b);  // node i+1's tips, used there
      fpos += nl;
      const uint4* cur = stage[buf];
      if (fl & DOWN_SPILL) {
        if (sp < LV) {
          uint4 u[R];
#pragma unroll
          for (int r = 0; r < R; ++r) {
            bf16x8 b = to_bf16x8(acc[r]);
            u[r] = *reinterpret_cast<uint4*>(&b);
          }
          static_for<0, LV>([&](auto lv) {
            const bool hit = sp == lv;
#pragma unroll
            for (int r = 0; r < R; ++r) {
              rs[lv][r].x = hit ? u[r].x : rs[lv][r].x;
              rs[lv][r].y = hit ? u[r].y : rs[lv][r].y;
              rs[lv][r].z = hit ? u[r].z : rs[lv][r].z;
              rs[lv][r].w = hit ? u[r].w : rs[lv][r].w;
            }
            rse[lv] = hit ? acc_e : rse[lv];
          });
        } else {
          uint4* s = stk + (size_t)(sp - LV) * (R * 64 + 16);
#pragma unroll
          for (int r = 0; r < R; ++r) {
            bf16x8 b = to_bf16x8(acc[r]);
            s[r * 64 + l] = *reinterpret_cast<uint4*>(&b);
          }
          if (q == 0) reinterpret_cast<int*>(s + R * 64)[pl] = acc_e;
        }
        ++sp;
      }
      float xa[R][8];
      int ea = 0, eb = 0;
#pragma unroll
      for (int c = 0; c < 2; ++c) {
        const int src = c ? ((fl >> 2) & 3) : (fl & 3);
        bf16x8 bop[R];
        int e = 0;
        if (src == SRC_LEAF) {
          const uint4 tu = tipop[min(c ? cb : ca, (uint32_t)KAA) * 4 + q];
          const bf16x8 tb = *reinterpret_cast<const bf16x8*>(&tu);
#pragma unroll
          for (int r = 0; r < R; ++r) bop[r] = tb;
        } else if (src == SRC_ACC) {
#pragma unroll
          for (int r = 0; r < R; ++r) bop[r] = to_bf16x8(acc[r]);
          e = acc_e;
        } else {
          --sp;
          if (sp < LV) {
            uint4 u[R];
#pragma unroll
            for (int r = 0; r < R; ++r) u[r] = rs[0][r];
            e = rse[0];
            static_for<1, LV>([&](auto lv) {
              const bool hit = sp == lv;
#pragma unroll
              for (int r = 0; r < R; ++r) {
                u[r].x = hit ? rs[lv][r].x : u[r].x;
                u[r].y = hit ? rs[lv][r].y : u[r].y;
                u[r].z = hit ? rs[lv][r].z : u[r].z;
                u[r].w = hit ? rs[lv][r].w : u[r].w;
              }
              e = hit ? rse[lv] : e;
            });
#pragma unroll
            for (int r = 0; r < R; ++r) bop[r] = *reinterpret_cast<const bf16x8*>(&u[r]);
          } else {
            const uint4* s = stk + (size_t)(sp - LV) * (R * 64 + 16);
#pragma unroll
            for (int r = 0; r < R; ++r) {
              const uint4 u = s[r * 64 + l];
              bop[r] = *reinterpret_cast<const bf16x8*>(&u);
            }
            e = reinterpret_cast<const int*>(s + R * 64)[pl];
          }
        }
#pragma unroll
        for (int r = 0; r < R; ++r) {
          f32x4 d0 = {0.f, 0.f, 0.f, 0.f}, d1 = {0.f, 0.f, 0.f, 0.f};
          const uint4 a0 = cur[((c * R + r) * 2 + 0) * OPB + l];
          const uint4 a1 = cur[((c * R + r) * 2 + 1) * OPB + l];
          d0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8*>(&a0), bop[r], d0, 0, 0, 0);
          d1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8*>(&a1), bop[r], d1, 0, 0, 0);
          if (c == 0) {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              xa[r][v] = d0[v];
              xa[r][4 + v] = d1[v];
            }
          } else {
#pragma unroll
            for (int v = 0; v < 4; ++v) {
              acc[r][v] = xa[r][v] * d0[v];
              acc[r][4 + v] = xa[r][4 + v] * d1[v];
            }
          }
        }
        if (c == 0) ea = e; else eb = e;
      }
      acc_e = ea + eb;
      // per-pattern rescale: lanes l, l^16, l^32, l^48 share pattern pl
      float mx = 0.f;
#pragma unroll
      for (int r = 0; r < R; ++r)
#pragma unroll
        for (int v = 0; v < 8; ++v) mx = fmaxf(mx, acc[r][v]);
      // with lanes l ^ 16 and l ^ 32 (v_permlane16/32_swap: no LDS round trip)
      {
        const unsigned u = __float_as_uint(mx);
        const auto s16 = __builtin_amdgcn_permlane16_swap(u, u, false, false);
        mx = fmaxf(__uint_as_float(s16[0]), __uint_as_float(s16[1]));
        const unsigned u2 = __float_as_uint(mx);
        const auto s32 = __builtin_amdgcn_permlane32_swap(u2, u2, false, false);
        mx = fmaxf(__uint_as_float(s32[0]), __uint_as_float(s32[1]));
      }
      if (mx < 0x1p-24f && mx > 0.f) {
        int ex;
        (void)frexpf(mx, &ex);
#pragma unroll
        for (int r = 0; r < R; ++r)
#pragma unroll
          for (int v = 0; v < 8; ++v) acc[r][v] = ldexpf(acc[r][v], -ex);
        acc_e += ex;
      }
      __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): node i+1 operands and tips landed
      __syncthreads();
    });
    // root: L_p = sum_r prob_r sum_i pi_i L[r][i]
    const double* pi = a.pi + (size_t)t * KAA;
    const double* pr = a.probs + (size_t)t * R;
    double s = 0.0;
#pragma unroll
    for (int r = 0; r < R; ++r) {
      double sr = 0.0;
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int st = 8 * q + v;
        if (st < KAA) sr += pi[st] * (double)acc[r][v];
      }
      s += pr[r] * sr;
    }
    s += __shfl_xor(s, 16, 64);
    s += __shfl_xor(s, 32, 64);
    const uint32_t w = a.wts[a.w_off[t] + p];
    double contrib = (q == 0 && w) ? (double)w * (log(s) + (double)acc_e * 0.6931471805599453) : 0.0;
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) contrib += __shfl_xor(contrib, off, 64);
    if (l == 0) a.part[a.part_off[t] + (int64_t)tile * WAVES + wave] = contrib;
  }
}

// one block per tree: Neumaier sum of its partials
__global__ void k_aa_reduce(const double* __restrict__ part, const int64_t* __restrict__ off,
                            double* __restrict__ out) {
  const int t = blockIdx.x;
  const int64_t b = off[t], e = off[t + 1];
  __shared__ double ss[256], cc[256];
  double s = 0.0, c = 0.0;
  for (int64_t i = b + threadIdx.x; i < e; i += blockDim.x) {
    const double v = part[i], y = s + v;
    if (fabs(s) >= fabs(v)) c += (s - y) + v; else c += (v - y) + s;
    s = y;
  }
  ss[threadIdx.x] = s;
  cc[threadIdx.x] = c;
  __syncthreads();
  if (threadIdx.x == 0) {
    double S = 0.0, C = 0.0;
    for (int k = 0; k < (int)blockDim.x; ++k) {
      const double v = ss[k], y = S + v;
      if (fabs(S) >= fabs(v)) C += (S - y) + v; else C += (v - y) + S;
      S = y;
      C += cc[k];
    }
    out[t] = S + C;
  }
}

__global__ void k_zero_u32(unsigned* p) { *p = 0u; }

}  // namespace

// ---------------------------------------------------------------------------
struct phy_aa_batch {
  phy_ctx* ctx = nullptr;
  int32_t T = 0, R = 4, waves = 4, depth = 0;
  std::vector<int32_t> n_leaves;
  std::vector<int64_t> n_pat, pitch;
  int64_t n_items = 0;
  int64_t n_branches = 0, pattern_nodes = 0;
  std::vector<void*> bufs;  // owned device allocations
  // device arrays
  int4* prog = nullptr;
  int64_t *prog_off = nullptr, *fifo_off = nullptr, *tips_off = nullptr, *pitch_d = nullptr,
          *w_off = nullptr, *pop_off = nullptr, *part_off = nullptr, *blen_off = nullptr;
  int32_t *fifo = nullptr, *nl = nullptr, *br_tree = nullptr;
  uint8_t* tips = nullptr;
  uint32_t* wts = nullptr;
  uint4* pops = nullptr;
  double *pi = nullptr, *probs = nullptr, *rates = nullptr, *eig = nullptr, *blen = nullptr,
         *part = nullptr;
  int2* items = nullptr;
  unsigned* counter = nullptr;
};

namespace {

template <class T>
int upload(phy_aa_batch* b, T** dst, const std::vector<T>& src) {
  const size_t bytes = std::max<size_t>(src.size() * sizeof(T), 16);
  if (hipMalloc((void**)dst, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return set_err(b->ctx, PHY_ENOMEM, "aa_batch: device allocation failed");
  }
  b->bufs.push_back(*dst);
  if (!src.empty()) PHY_HIP(b->ctx, hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice));
  return PHY_OK;
}

size_t lds_bytes(int waves, int R, int depth, int lv) {
  return (size_t)waves * (depth > lv ? depth - lv : 0) * (R * 64 + 16) * sizeof(uint4);
}

// stack levels kept in registers: 3 for stacks of 3 or more levels (frees
// LDS for a second workgroup per CU), else none (fewer VGPRs).
// PHY_AA_REG=0|3 overrides.
int aa_register_levels(int depth) {
  const char* s = getenv("PHY_AA_REG");
  if (s && (s[0] == '0' || s[0] == '3') && s[1] == 0) return s[0] - '0';
  return depth >= 3 ? 3 : 0;
}

template <int WAVES, int R, int LV>
int launch(phy_ctx* ctx, phy_aa_batch* b) {
  AADev a;
  a.prog = b->prog;
  a.prog_off = b->prog_off;
  a.fifo = b->fifo;
  a.fifo_off = b->fifo_off;
  a.n_leaves = b->nl;
  a.tips = b->tips;
  a.tips_off = b->tips_off;
  a.pitch = b->pitch_d;
  a.wts = b->wts;
  a.w_off = b->w_off;
  a.pops = b->pops;
  a.pop_off = b->pop_off;
  a.pi = b->pi;
  a.probs = b->probs;
  a.items = b->items;
  a.n_items = (int)b->n_items;
  a.part = b->part;
  a.part_off = b->part_off;
  a.counter = b->counter;
  a.depth = std::max(1, b->depth);
  const size_t lds = lds_bytes(WAVES, R, a.depth, LV);
  int dev = 0, n_cu = 256;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev);
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_aa_prune<WAVES, R, LV>, 64 * WAVES, lds) != hipSuccess || per_cu < 1)
    per_cu = 1;
  const int64_t grid = std::min<int64_t>((int64_t)n_cu * per_cu, b->n_items);
  KernelSpan span(ctx, "aa_prune");
  hipLaunchKernelGGL((k_aa_prune<WAVES, R, LV>), dim3((unsigned)std::max<int64_t>(grid, 1)), dim3(64 * WAVES), lds,
                     ctx->stream, a);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

template <int R>
int launch_w(phy_ctx* ctx, phy_aa_batch* b) {
  const bool reg = aa_register_levels(std::max(1, b->depth)) == 3;
  switch (b->waves) {
    case 1: return reg ? launch<1, R, 3>(ctx, b) : launch<1, R, 0>(ctx, b);
    case 2: return reg ? launch<2, R, 3>(ctx, b) : launch<2, R, 0>(ctx, b);
    case 4: return reg ? launch<4, R, 3>(ctx, b) : launch<4, R, 0>(ctx, b);
    case 8: return reg ? launch<8, R, 3>(ctx, b) : launch<8, R, 0>(ctx, b);
    default: return set_err(ctx, PHY_ERANGE, "aa_batch: waves per workgroup must be 1, 2, 4 or 8");
  }
}

}  // namespace

extern "C" {

int phy_aa_batch_create(phy_ctx* ctx, int32_t T, const int32_t* n_leaves, const int64_t* n_patterns,
                        const int32_t* children, const int32_t* roots, const double* blen,
                        const double* pi, const double* Q, int32_t R, const double* rates,
                        const double* probs, const uint8_t* tips, const uint32_t* weights,
                        int32_t waves, phy_aa_batch** out) {
  if (!ctx || !out) return PHY_EINVAL;
  *out = nullptr;
  if (T < 1 || !n_leaves || !n_patterns || !children || !roots || !blen || !pi || !Q || !rates ||
      !probs || !tips || !weights)
    return set_err(ctx, PHY_EINVAL, "aa_batch: null argument");
  if (!(R == 1 || R == 2 || R == 4)) return set_err(ctx, PHY_ERANGE, "aa_batch: n_rates must be 1, 2 or 4");
  if (!(waves == 1 || waves == 2 || waves == 4 || waves == 8))
    return set_err(ctx, PHY_ERANGE, "aa_batch: waves must be 1, 2, 4 or 8");
  phy_aa_batch* b = new phy_aa_batch();
  b->ctx = ctx;
  b->T = T;
  b->R = R;
  b->waves = waves;
  const int tile = 16 * waves;
  std::vector<int4> prog;
  std::vector<int64_t> prog_off, fifo_off, tips_off, pitch, w_off, pop_off, part_off, blen_off;
  std::vector<int32_t> fifo, nl, br_tree;
  std::vector<uint8_t> tipv;
  std::vector<uint32_t> wv;
  std::vector<double> eig, blv;
  std::vector<int2> items;
  int64_t ch_pos = 0, tip_pos = 0, pat_pos = 0, pop_pos = 0, part_pos = 0;
  for (int32_t t = 0; t < T; ++t) {
    const int32_t N = n_leaves[t];
    const int64_t P = n_patterns[t];
    if (N < 2 || P < 1) {
      delete b;
      return set_err(ctx, PHY_EINVAL, "aa_batch: every tree needs N >= 2 and P >= 1");
    }
    phy_topo* topo = nullptr;
    if (phy_topo_create(nullptr, N, children + ch_pos, roots[t], &topo) != PHY_OK) {
      delete b;
      return set_err(ctx, PHY_EINVAL, "aa_batch: malformed tree " + std::to_string(t));
    }
    b->depth = std::max(b->depth, topo->down_depth);
    prog_off.push_back((int64_t)prog.size());
    for (size_t k = 0; k < topo->down_prog.size(); k += 4)
      prog.push_back(make_int4(topo->down_prog[k], topo->down_prog[k + 1], topo->down_prog[k + 2],
                               topo->down_prog[k + 3]));
    fifo_off.push_back((int64_t)fifo.size());
    fifo.insert(fifo.end(), topo->down_fifo.begin(), topo->down_fifo.end());
    fifo.insert(fifo.end(), 128, 0);  // the kernel holds ids fbase .. fbase+127, fbase <= N
    phy_topo_destroy(topo);
    nl.push_back(N);
    const int64_t pp = (P + tile - 1) / tile * tile;
    pitch.push_back(pp);
    tips_off.push_back((int64_t)tipv.size());
    for (int32_t leaf = 0; leaf < N; ++leaf) {
      const uint8_t* row = tips + tip_pos + (size_t)leaf * P;
      tipv.insert(tipv.end(), row, row + P);
      tipv.insert(tipv.end(), (size_t)(pp - P), (uint8_t)255);  // padding: missing data
    }
    tip_pos += (int64_t)N * P;
    w_off.push_back((int64_t)wv.size());
    wv.insert(wv.end(), weights + pat_pos, weights + pat_pos + P);
    wv.insert(wv.end(), (size_t)(pp - P), 0u);
    pat_pos += P;
    // eigen-decomposition of Q_t on the host
    std::vector<double> U(KAA * KAA), lam(KAA), Ui(KAA * KAA);
    std::string err;
    if (!phy::eigen_reversible(KAA, Q + (size_t)t * KAA * KAA, U.data(), lam.data(), Ui.data(), &err)) {
      delete b;
      return set_err(ctx, PHY_ENUMERIC, "aa_batch: tree " + std::to_string(t) + ": " + err);
    }
    eig.insert(eig.end(), U.begin(), U.end());
    eig.insert(eig.end(), lam.begin(), lam.end());
    eig.insert(eig.end(), Ui.begin(), Ui.end());
    const int32_t M = 2 * N - 1;
    blen_off.push_back((int64_t)blv.size());
    blv.insert(blv.end(), blen + b->n_branches, blen + b->n_branches + M);
    for (int32_t v = 0; v < M; ++v) br_tree.push_back(t);
    b->n_branches += M;
    pop_off.push_back(pop_pos);
    pop_pos += (int64_t)M * R * 2 * OPB;
    part_off.push_back(part_pos);
    const int64_t tiles = pp / tile;
    part_pos += tiles * waves;
    for (int64_t k = 0; k < tiles; ++k) items.push_back(make_int2(t, (int)k));
    ch_pos += 2 * (int64_t)(N - 1);
    b->n_leaves.push_back(N);
    b->n_pat.push_back(P);
    b->pitch.push_back(pp);
    b->pattern_nodes += P * (N - 1);
  }
  part_off.push_back(part_pos);
  // largest trees first (longest items first), a tree's tiles kept together
  std::stable_sort(items.begin(), items.end(), [&](const int2& x, const int2& y) {
    const int nx = nl[x.x], ny = nl[y.x];
    if (nx != ny) return nx > ny;
    return x.x < y.x;
  });
  b->n_items = (int64_t)items.size();
  if (b->n_items > 0x7fffffff) {
    delete b;
    return set_err(ctx, PHY_ERANGE, "aa_batch: too many work items");
  }
  // static operand double buffer (2 x 2R x 2 KiB) plus the dynamic stack,
  // within the device's LDS per workgroup
  int lds_max = 64 * 1024, dev = 0;
  (void)hipGetDevice(&dev);
  (void)hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
  const int lv = aa_register_levels(std::max(1, b->depth));
  if (lds_bytes(waves, R, std::max(1, b->depth), lv) + (size_t)2 * 2 * R * 2 * OPB * sizeof(uint4) > (size_t)lds_max) {
    delete b;
    return set_err(ctx, PHY_ERANGE, "aa_batch: LDS stack too deep for this many waves per workgroup");
  }
  std::vector<double> pis(pi, pi + (size_t)T * KAA), prv(probs, probs + (size_t)T * R),
      rtv(rates, rates + (size_t)T * R);
  int rc = PHY_OK;
  if (!rc) rc = upload(b, &b->prog, prog);
  if (!rc) rc = upload(b, &b->prog_off, prog_off);
  if (!rc) rc = upload(b, &b->fifo, fifo);
  if (!rc) rc = upload(b, &b->fifo_off, fifo_off);
  if (!rc) rc = upload(b, &b->nl, nl);
  if (!rc) rc = upload(b, &b->tips, tipv);
  if (!rc) rc = upload(b, &b->tips_off, tips_off);
  if (!rc) rc = upload(b, &b->pitch_d, pitch);
  if (!rc) rc = upload(b, &b->wts, wv);
  if (!rc) rc = upload(b, &b->w_off, w_off);
  if (!rc) rc = upload(b, &b->pop_off, pop_off);
  if (!rc) rc = upload(b, &b->part_off, part_off);
  if (!rc) rc = upload(b, &b->pi, pis);
  if (!rc) rc = upload(b, &b->probs, prv);
  if (!rc) rc = upload(b, &b->rates, rtv);
  if (!rc) rc = upload(b, &b->eig, eig);
  if (!rc) rc = upload(b, &b->blen, blv);
  if (!rc) rc = upload(b, &b->blen_off, blen_off);
  if (!rc) rc = upload(b, &b->br_tree, br_tree);
  if (!rc) rc = upload(b, &b->items, items);
  if (!rc) rc = upload(b, &b->pops, std::vector<uint4>((size_t)pop_pos));
  if (!rc) rc = upload(b, &b->part, std::vector<double>((size_t)part_pos));
  if (!rc) rc = upload(b, &b->counter, std::vector<unsigned>(4, 0u));
  if (rc) {
    phy_aa_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PHY_OK;
}

void phy_aa_batch_destroy(phy_aa_batch* b) {
  if (!b) return;
  for (void* p : b->bufs) (void)hipFree(p);
  delete b;
}

int phy_aa_batch_info(const phy_aa_batch* b, int64_t* pattern_nodes, int64_t* n_items,
                      int32_t* stack_depth) {
  if (!b) return PHY_EINVAL;
  if (pattern_nodes) *pattern_nodes = b->pattern_nodes;
  if (n_items) *n_items = b->n_items;
  if (stack_depth) *stack_depth = b->depth;
  return PHY_OK;
}

int phy_aa_batch_run(phy_ctx* ctx, phy_aa_batch* b, double* d_lnl) {
  if (!ctx || !b || !d_lnl) return PHY_EINVAL;
  if (b->ctx != ctx) return set_err(ctx, PHY_EINVAL, "aa_batch: built on another context");
  // 1. P operands for every (tree, branch, rate)
  const int64_t n_pairs = b->n_branches * b->R;
  const int64_t nwg = (n_pairs + kPopsPerBlock - 1) / kPopsPerBlock;
  hipLaunchKernelGGL(k_aa_pops, dim3((unsigned)nwg), dim3(64 * kPopsPerBlock), 0, ctx->stream, b->eig,
                     b->rates, b->blen, b->blen_off, b->br_tree, b->R, n_pairs, b->pop_off, b->pops);
  PHY_LAUNCHED(ctx);
  // 2. the persistent pruning launch
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(1), 0, ctx->stream, b->counter);
  PHY_LAUNCHED(ctx);
  int rc;
  switch (b->R) {
    case 1: rc = launch_w<1>(ctx, b); break;
    case 2: rc = launch_w<2>(ctx, b); break;
    default: rc = launch_w<4>(ctx, b); break;
  }
  if (rc) return rc;
  // 3. per-tree compensated sums
  hipLaunchKernelGGL(k_aa_reduce, dim3((unsigned)b->T), dim3(256), 0, ctx->stream, b->part, b->part_off, d_lnl);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

}  // extern "C"
