// Per-wave streaming helpers of the whole-tree passes.
//
// A tree pass walks a uniform program (one int4 entry per internal node) and
// consumes an operand stream (leaf ids, or node ids) in a fixed order.  Both
// are read off the per-node latency chain:
//  * for_each_entry: the wave holds 64 program entries in VGPRs (lane l holds
//    entry base + l) and takes entry j with v_readlane, so no scalar-memory
//    round trip sits between two nodes; the next 64 entries are loaded one
//    chunk ahead.
//  * OperandRing: operand loads are issued B ahead into registers, parked in
//    an LDS ring of 2B slots per thread, and the next two operands are read
//    from LDS one step ahead of use.  The id stream must be padded
//    (phy_fifo_padded) because up to 3B ids past the last consumed one are
//    read.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "common.hpp"

// compile-time loop: f(integral_constant<int, I>) for I in [B, E)
template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) {
    f(std::integral_constant<int, B>{});
    static_for<B + 1, E>(f);
  }
}


template <class F>
__device__ __forceinline__ void for_each_entry(const int4* __restrict__ prog, int n_ops, F&& f) {
  if (n_ops <= 0) return;
  const int lane = threadIdx.x & 63;
  int4 cur = prog[min(lane, n_ops - 1)];
  int4 nxt = prog[min(64 + lane, n_ops - 1)];
  for (int base = 0; base < n_ops; base += 64) {
    const int cnt = __builtin_amdgcn_readfirstlane(min(64, n_ops - base));  // wave-uniform
    for (int j = 0; j < cnt; ++j) {
      int4 op;
      op.x = __builtin_amdgcn_readlane(cur.x, j);
      op.y = __builtin_amdgcn_readlane(cur.y, j);
      op.z = __builtin_amdgcn_readlane(cur.z, j);
      op.w = __builtin_amdgcn_readlane(cur.w, j);
      f(base + j, op);
    }
    cur = nxt;
    nxt = prog[min(base + 128 + lane, n_ops - 1)];
  }
}

// Same, passing the following entry too (equal to the last entry at the end),
// for passes that prefetch operands one node ahead.
template <class F>
__device__ __forceinline__ void for_each_entry_ahead(const int4* __restrict__ prog, int n_ops, F&& f) {
  if (n_ops <= 0) return;
  const int lane = threadIdx.x & 63;
  int4 cur = prog[min(lane, n_ops - 1)];
  int4 nxt = prog[min(64 + lane, n_ops - 1)];
  for (int base = 0; base < n_ops; base += 64) {
    const int cnt = __builtin_amdgcn_readfirstlane(min(64, n_ops - base));  // wave-uniform
    for (int j = 0; j < cnt; ++j) {
      int4 op, on;
      op.x = __builtin_amdgcn_readlane(cur.x, j);
      op.y = __builtin_amdgcn_readlane(cur.y, j);
      op.z = __builtin_amdgcn_readlane(cur.z, j);
      op.w = __builtin_amdgcn_readlane(cur.w, j);
      if (j + 1 < 64) {
        on.x = __builtin_amdgcn_readlane(cur.x, j + 1);
        on.y = __builtin_amdgcn_readlane(cur.y, j + 1);
        on.z = __builtin_amdgcn_readlane(cur.z, j + 1);
        on.w = __builtin_amdgcn_readlane(cur.w, j + 1);
      } else {
        on.x = __builtin_amdgcn_readlane(nxt.x, 0);
        on.y = __builtin_amdgcn_readlane(nxt.y, 0);
        on.z = __builtin_amdgcn_readlane(nxt.z, 0);
        on.w = __builtin_amdgcn_readlane(nxt.w, 0);
      }
      f(base + j, op, on);
    }
    cur = nxt;
    nxt = prog[min(base + 128 + lane, n_ops - 1)];
  }
}

// Loader: T operator()(int id) const, an in-bounds load for every id of the
// padded stream.
template <int B, class Loader, class T = uint32_t>
struct OperandRing {
  static_assert(B % 4 == 0 && 3 * B <= (int)kFifoSlack, "ring batch vs id padding");
  T pre[B];
  T* ring;  // [2B][bd]
  const int32_t* ids;
  Loader load;
  int bd, tid;
  int c, filled;
  T la0, la1;

  __device__ __forceinline__ void issue(int start) {
    const int4* iv = reinterpret_cast<const int4*>(ids + start);
#pragma unroll
    for (int q = 0; q < B / 4; ++q) {
      const int4 id = iv[q];
      pre[4 * q + 0] = load(id.x);
      pre[4 * q + 1] = load(id.y);
      pre[4 * q + 2] = load(id.z);
      pre[4 * q + 3] = load(id.w);
    }
  }
  __device__ __forceinline__ void park(int first) {  // pre -> slots of ids first..first+B-1
    const int h = (first / B) & 1;
#pragma unroll
    for (int j = 0; j < B; ++j) ring[(h * B + j) * bd + tid] = pre[j];
  }
  __device__ __forceinline__ T slot(int k) const { return ring[(k % (2 * B)) * bd + tid]; }

  __device__ __forceinline__ void init(T* lds, const int32_t* id_stream, Loader ld,
                                       int block_dim, int thread) {
    ring = lds;
    ids = id_stream;
    load = ld;
    bd = block_dim;
    tid = thread;
    issue(0);
    park(0);
    issue(B);
    park(B);
    issue(2 * B);
    filled = 2 * B;
    c = 0;
    la0 = slot(0);
    la1 = slot(1);
  }
  __device__ __forceinline__ T get() {
    const T v = la0;
    la0 = la1;
    ++c;
    if (c == filled - B) {  // the older half is consumed: refill it
      park(filled);
      filled += B;
      issue(filled);
    }
    la1 = slot(c + 1);
    return v;
  }
};
