// Branch-length Newton-Raphson with its step control on the device.
//
// The host optimiser (optimize_branches in phylocaml_amd/optimize.py, an
// implementation of the branch-length hooks the reference declares in
// lib/mlModel.ml:193-196) proposes a simultaneous Newton step on every
// branch, evaluates lnL, d1 and d2 at the proposal, and accepts it or halves
// it towards the current lengths.  Here the propose and accept steps are
// kernels over device state, so an iteration -- propose, the derivative
// pass at the proposal (phy_lk_tree), accept -- issues kernels only: no copy
// to the host, no synchronisation, and K iterations can be captured in one
// HIP graph and replayed.
//
// State (4 doubles on the device): [0] lnL at the current lengths t, [1] 1
// when the next proposal is a Newton step from t and 0 when it halves the
// last proposal towards t, [2] accepted steps, [3] halvings.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>

#include "../../include/phylo_hip.h"
#include "common.hpp"

namespace {

constexpr double kBlMin = 1e-8, kBlMax = 10.0;

// d2 < 0: the Newton step; otherwise half the length along the gradient
__device__ __forceinline__ double newton_step(double x, double g, double h) {
  return h < 0.0 ? x - g / h : x + (g > 0.0 ? 0.5 * x : -0.5 * x);
}

// tn from (t, d1, d2), or halved towards t; one thread per branch.  The two
// root branches are one unrooted edge: one step on their sum, their ratio
// kept.
__global__ void k_newton_propose(int M, int root, int ra, int rb, const double* __restrict__ st,
                                 const double* __restrict__ t, const double* __restrict__ d1,
                                 const double* __restrict__ d2, double* __restrict__ tn) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= M) return;
  if (st[1] == 0.0) {
    tn[i] = 0.5 * (t[i] + tn[i]);
    return;
  }
  if (i == root) {
    tn[i] = 0.0;
  } else if (i == ra || i == rb) {
    const double T = t[ra] + t[rb];
    const double Tn = fmin(fmax(newton_step(T, d1[ra], d2[ra]), 2.0 * kBlMin), 2.0 * kBlMax);
    const double fa = T > 0.0 ? t[ra] / T : 0.5;
    tn[i] = fmax(i == ra ? Tn * fa : Tn * (1.0 - fa), kBlMin);
  } else {
    tn[i] = fmin(fmax(newton_step(t[i], d1[i], d2[i]), kBlMin), kBlMax);
  }
}

// Accept the proposal when it does not lower lnL (t, d1, d2 and lnL take
// the proposal's), else ask for a halving.  One block: every thread reads
// the old lnL before thread 0 replaces it.
__global__ void k_newton_accept(int M, double* __restrict__ st, const double* __restrict__ lnl_n,
                                double* __restrict__ t, const double* __restrict__ tn, double* __restrict__ d1,
                                const double* __restrict__ d1n, double* __restrict__ d2,
                                const double* __restrict__ d2n) {
  const double ln = lnl_n[0];
  const bool ok = ln >= st[0];  // false for NaN
  if (ok)
    for (int i = threadIdx.x; i < M; i += blockDim.x) {
      t[i] = tn[i];
      d1[i] = d1n[i];
      d2[i] = d2n[i];
    }
  __syncthreads();
  if (threadIdx.x == 0) {
    if (ok) {
      st[0] = ln;
      st[2] += 1.0;
    } else {
      st[3] += 1.0;
    }
    st[1] = ok ? 1.0 : 0.0;
  }
}

// state <- [lnL, 1, 0, 0]
__global__ void k_newton_init(double* __restrict__ st, const double* __restrict__ lnl) {
  if (threadIdx.x == 0) {
    st[0] = lnl[0];
    st[1] = 1.0;
    st[2] = 0.0;
    st[3] = 0.0;
  }
}

bool root_edge(const phy_topo* t, int32_t* ra, int32_t* rb) {
  const int N = t->n_leaves;
  if (t->root < N) return false;
  *ra = t->children[2 * (t->root - N)];
  *rb = t->children[2 * (t->root - N) + 1];
  return true;
}

}  // namespace

extern "C" {

int phy_newton_init(phy_ctx* ctx, double* d_state, const double* d_lnl) {
  if (!ctx || !d_state || !d_lnl) return PHY_EINVAL;
  hipLaunchKernelGGL(k_newton_init, dim3(1), dim3(64), 0, ctx->stream, d_state, d_lnl);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

int phy_newton_propose(phy_ctx* ctx, const phy_topo* topo, const double* d_state, const double* d_blen,
                       const double* d_d1, const double* d_d2, double* d_blen_new) {
  if (!ctx || !topo || !d_state || !d_blen || !d_d1 || !d_d2 || !d_blen_new) return PHY_EINVAL;
  int32_t ra, rb;
  if (!root_edge(topo, &ra, &rb)) return set_err(ctx, PHY_EINVAL, "newton: bad topology");
  const int M = 2 * topo->n_leaves - 1;
  hipLaunchKernelGGL(k_newton_propose, dim3((unsigned)((M + 255) / 256)), dim3(256), 0, ctx->stream, M,
                     topo->root, ra, rb, d_state, d_blen, d_d1, d_d2, d_blen_new);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

int phy_newton_accept(phy_ctx* ctx, const phy_topo* topo, double* d_state, const double* d_lnl_new,
                      double* d_blen, const double* d_blen_new, double* d_d1, const double* d_d1_new,
                      double* d_d2, const double* d_d2_new) {
  if (!ctx || !topo || !d_state || !d_lnl_new || !d_blen || !d_blen_new || !d_d1 || !d_d1_new || !d_d2 ||
      !d_d2_new)
    return PHY_EINVAL;
  const int M = 2 * topo->n_leaves - 1;
  hipLaunchKernelGGL(k_newton_accept, dim3(1), dim3(1024), 0, ctx->stream, M, d_state, d_lnl_new, d_blen,
                     d_blen_new, d_d1, d_d1_new, d_d2, d_d2_new);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

}  // extern "C"
