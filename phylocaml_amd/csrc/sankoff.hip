// Sankoff (1975) parsimony under an integer transformation cost matrix, the
// whole tree in one launch.
//
// The reference has the cost-matrix type (PTcm.X, lib/pTcm.ml:5-50) and the
// median search over it (lib/costMatrix.ml:52-65, child state j under
// median k costs tcm[j][k]) but no tree-level Sankoff pass; this is the
// standard dynamic programme over that cost convention:
//   C_leaf[k] = 0 for k in the leaf's state set, else infinite
//   C_v[k]    = sum over children c of min_j (C_c[j] + tcm[j][k])
//   site cost = min_k C_root[k].
//
// Layout: lane = site (one byte per leaf and site holds the state set, K <= 8
// states).  A wave walks the post-order program of the topology (the
// Sethi-Ullman order of the pruning pass, accumulator plus LDS stack); the
// K x K matrix is wave-uniform (SGPRs), so an interior operand costs K^2 adds
// and K^2 - K mins on the lane's own K-vector.  A leaf operand is one LDS
// read of a table built once per block: tab[code][k] = min over j in code of
// tcm[j][k].  Leaf bytes stream in batches of 8 leaves, issued 16-24 leaves
// ahead.  All arithmetic is int32 (the host bounds the costs so no sum can
// overflow) and exact.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdlib>

#include "common.hpp"
#include "stream.hpp"

namespace {

constexpr int32_t kInf = 1 << 28;  // larger than any finite cost of a site

// One byte per leaf for this lane's site, 8 leaves a batch.
struct TipBytes {
  const uint8_t* tips;
  int64_t stride, col;
  const int32_t* ids;
  uint32_t pre[8];
  uint32_t cur0, cur1, nxt0, nxt1;
  int pos, next_start;

  __device__ __forceinline__ void issue(int start) {
    const int4* iv = reinterpret_cast<const int4*>(ids + start);
    const int4 a = iv[0], b = iv[1];
    const int id[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
    for (int j = 0; j < 8; ++j) pre[j] = tips[(int64_t)id[j] * stride + col];
  }
  __device__ __forceinline__ void pack(uint32_t& lo, uint32_t& hi) const {
    lo = pre[0] | (pre[1] << 8) | (pre[2] << 16) | (pre[3] << 24);
    hi = pre[4] | (pre[5] << 8) | (pre[6] << 16) | (pre[7] << 24);
  }
  __device__ __forceinline__ void init(const uint8_t* t, int64_t s, int64_t c, const int32_t* id) {
    tips = t;
    stride = s;
    col = c;
    ids = id;
    issue(0);
    pack(cur0, cur1);
    issue(8);
    pack(nxt0, nxt1);
    issue(16);
    next_start = 16;
    pos = 0;
  }
  __device__ __forceinline__ uint32_t get() {
    const uint32_t v = ((pos < 4 ? cur0 : cur1) >> (8 * (pos & 3))) & 255u;
    if (++pos == 8) {
      cur0 = nxt0;
      cur1 = nxt1;
      pack(nxt0, nxt1);
      next_start += 8;
      issue(next_start);
      pos = 0;
    }
    return v;
  }
};
static_assert(24 + 8 <= (int)kFifoSlack, "leaf stream look-ahead vs id padding");

__device__ __forceinline__ uint64_t wave_sum_u64(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

template <int K>
__global__ __launch_bounds__(256) void k_sankoff_down(
    const int4* __restrict__ prog, int n_ops, const int32_t* __restrict__ fifo,
    const int32_t* __restrict__ tcm, const uint8_t* __restrict__ tips, int64_t stride,
    int64_t n_sites, const uint32_t* __restrict__ wt, uint32_t* __restrict__ site_cost,
    unsigned long long* __restrict__ total) {
  constexpr int NC = 1 << K;
  extern __shared__ int32_t lds[];
  int32_t* tab = lds;            // [NC][K]
  int32_t* stk = lds + NC * K;   // [depth][K][blockDim]
  const int tid = threadIdx.x, bd = blockDim.x;
  for (int e = tid; e < NC * K; e += bd) {
    const int c = e / K ? e / K : NC - 1, k = e % K;  // code 0: every state
    int32_t m = kInf;
    for (int j = 0; j < K; ++j)
      if ((c >> j) & 1) m = min(m, tcm[j * K + k]);
    tab[e] = m;
  }
  int32_t T[K][K];
#pragma unroll
  for (int j = 0; j < K; ++j)
#pragma unroll
    for (int k = 0; k < K; ++k) T[j][k] = tcm[j * K + k];
  __syncthreads();

  const int64_t s = (int64_t)blockIdx.x * bd + tid;
  const bool act = s < n_sites;
  TipBytes ts;
  ts.init(tips, stride, act ? s : 0, fifo);
  int32_t acc[K];
#pragma unroll
  for (int k = 0; k < K; ++k) acc[k] = 0;
  int sp = 0;

  auto leaf = [&](int32_t* m) {
    const int c = (int)(ts.get() & (NC - 1));
#pragma unroll
    for (int k = 0; k < K; ++k) m[k] = tab[c * K + k];
  };
  auto through = [&](const int32_t* x, int32_t* m) {  // m[k] = min_j x[j] + T[j][k]
#pragma unroll
    for (int k = 0; k < K; ++k) {
      int32_t v = x[0] + T[0][k];
#pragma unroll
      for (int j = 1; j < K; ++j) v = min(v, x[j] + T[j][k]);
      m[k] = v;
    }
  };

  for_each_entry(prog, n_ops, [&](int, const int4 op) {
    const int fl = op.w;
    if (fl & DOWN_SPILL) {
#pragma unroll
      for (int k = 0; k < K; ++k) stk[(sp * K + k) * bd + tid] = acc[k];
      ++sp;
    }
    int32_t ma[K], mb[K];
    const int sa = fl & 3, sb = (fl >> 2) & 3;
    if (sa == SRC_LEAF) {
      leaf(ma);
    } else if (sa == SRC_ACC) {
      through(acc, ma);
    } else {
      --sp;
      int32_t x[K];
#pragma unroll
      for (int k = 0; k < K; ++k) x[k] = stk[(sp * K + k) * bd + tid];
      through(x, ma);
    }
    if (sb == SRC_LEAF)
      leaf(mb);
    else
      through(acc, mb);
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = ma[k] + mb[k];
  });

  int32_t best = acc[0];
#pragma unroll
  for (int k = 1; k < K; ++k) best = min(best, acc[k]);
  uint64_t contrib = 0;
  if (act) {
    site_cost[s] = (uint32_t)best;
    contrib = (uint64_t)(wt ? wt[s] : 1u) * (uint64_t)best;
  }
  contrib = wave_sum_u64(contrib);
  if ((tid & 63) == 0 && contrib) atomicAdd(total, (unsigned long long)contrib);
}

size_t sankoff_lds(int K, int depth, int bs) {
  return ((size_t)(1 << K) * K + (size_t)depth * K * bs) * 4;
}

template <int K>
int sankoff_launch(phy_ctx* ctx, const phy_topo* topo, const int32_t* tcm, const uint8_t* tips,
                   int64_t stride, int64_t n_sites, const uint32_t* wt, uint32_t* site_cost,
                   unsigned long long* total) {
  int bs = 256;
  while (bs > 64 && sankoff_lds(K, topo->down_depth, bs) > 64 * 1024) bs /= 2;
  const size_t lds = sankoff_lds(K, topo->down_depth, bs);
  if (lds > 64 * 1024) return set_err(ctx, PHY_ERANGE, "sankoff: LDS stack too deep");
  const int64_t nb = (n_sites + bs - 1) / bs;
  KernelSpan span(ctx, "sankoff_down");
  hipLaunchKernelGGL((k_sankoff_down<K>), dim3((unsigned)nb), dim3(bs), lds, ctx->stream,
                     topo->d_down_prog, topo->n_leaves - 1, topo->d_down_fifo, tcm, tips, stride,
                     n_sites, wt, site_cost, total);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

}  // namespace

extern "C" {

int phy_sankoff_downpass(phy_ctx* ctx, const phy_topo* topo, int32_t n_states,
                         const int32_t* d_tcm, const uint8_t* d_leaves, int64_t stride,
                         int64_t n_sites, const uint32_t* d_weights, uint32_t* d_site_cost,
                         uint64_t* d_total) {
  if (!ctx) return PHY_EINVAL;
  if (!topo || !topo->d_down_prog) return set_err(ctx, PHY_EINVAL, "sankoff: topology not on a device");
  if (n_states < 2 || n_states > 8) return set_err(ctx, PHY_EINVAL, "sankoff: n_states must be 2..8");
  if (n_sites < 0 || stride < n_sites) return set_err(ctx, PHY_EINVAL, "sankoff: stride < n_sites");
  if (!d_tcm || !d_site_cost || !d_total || (n_sites > 0 && !d_leaves))
    return set_err(ctx, PHY_EINVAL, "sankoff: null buffer");
  int rc;
  if ((rc = phy_zero_u64(ctx, d_total, 1))) return rc;
  if (n_sites == 0) return PHY_OK;
  auto* tt = reinterpret_cast<unsigned long long*>(d_total);
  switch (n_states) {
    case 2: return sankoff_launch<2>(ctx, topo, d_tcm, d_leaves, stride, n_sites, d_weights, d_site_cost, tt);
    case 3: return sankoff_launch<3>(ctx, topo, d_tcm, d_leaves, stride, n_sites, d_weights, d_site_cost, tt);
    case 4: return sankoff_launch<4>(ctx, topo, d_tcm, d_leaves, stride, n_sites, d_weights, d_site_cost, tt);
    case 5: return sankoff_launch<5>(ctx, topo, d_tcm, d_leaves, stride, n_sites, d_weights, d_site_cost, tt);
    case 6: return sankoff_launch<6>(ctx, topo, d_tcm, d_leaves, stride, n_sites, d_weights, d_site_cost, tt);
    case 7: return sankoff_launch<7>(ctx, topo, d_tcm, d_leaves, stride, n_sites, d_weights, d_site_cost, tt);
    default: return sankoff_launch<8>(ctx, topo, d_tcm, d_leaves, stride, n_sites, d_weights, d_site_cost, tt);
  }
}

}  // extern "C"
