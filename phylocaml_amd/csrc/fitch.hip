// Fitch parsimony on packed 4-bit state sets (16 sites per uint64 word).
//
// Per-site rule (NonAdditive_c.median_2, lib/nonAdditive_c.ml:19-35, and
// bv_fitch, lib/bitvector/bv.c:147-161): m = x & y; if m is empty the median
// is x | y at cost 1, else m at cost 0.  Sixteen sites are handled at once
// with SWAR arithmetic on each 32-bit half of a word:
//   nonempty flag in bit 3 of each nibble: ((t & 0x7..7) + 0x7..7 | t) & 0x8..8
//   (the masked add cannot carry across nibbles), a full-nibble mask from it
//   by  m8 | (m8 - (m8 >> 3)),  and the median by one bit-field insert.
//
// Whole-tree passes: one thread owns V 32-bit words (8V sites) of every node
// and walks the uniform program compiled by phy_topo_create.  Intermediate
// sets live in an accumulator register and a per-thread LDS stack, so the
// algorithmic HBM traffic is the minimum: leaves read once and every
// preliminary set written once (down), every set read once and every final
// set written once (up).  The operand stream (leaves for the down-pass, node
// sets for the up-pass) is consumed in a fixed order and is prefetched a
// batch ahead through registers into an LDS ring.  Trees of >= 128 leaves
// run split: parts under cut nodes side by side (grid.y = part), plus a top
// program, which multiplies the waves in flight by the number of parts.
#include <algorithm>
#include <cstdlib>

#include "common.hpp"
#include "stream.hpp"

namespace {

constexpr uint32_t K1 = 0x11111111u, K7 = 0x77777777u, K8 = 0x88888888u;

__device__ __forceinline__ uint32_t nz8(uint32_t t) { return (((t & K7) + K7) | t) & K8; }
// nibble mask: 0xF where the nibble of t is non-empty
__device__ __forceinline__ uint32_t nzmask(uint32_t t) {
  const uint32_t m = nz8(t);
  return m | (m - (m >> 3));
}
__device__ __forceinline__ uint32_t bfi(uint32_t mask, uint32_t a, uint32_t b) {
  return (mask & a) | (~mask & b);
}

// Fitch median of one 32-bit half; returns the empty-intersection flags in
// bit 0 of each nibble (the per-site costs).
__device__ __forceinline__ uint32_t fitch32(uint32_t a, uint32_t b, uint32_t* out) {
  const uint32_t t = a & b, u = a | b;
  const uint32_t e8 = nz8(t) ^ K8;
  const uint32_t e0 = e8 >> 3;
  const uint32_t mask = e8 | (e8 - e0);
  *out = bfi(mask, u, t);
  return e0;
}

struct W2 {
  uint32_t lo, hi;
};
__device__ __forceinline__ W2 split(uint64_t x) { return {(uint32_t)x, (uint32_t)(x >> 32)}; }
__device__ __forceinline__ uint64_t join(uint32_t lo, uint32_t hi) {
  return (uint64_t)lo | ((uint64_t)hi << 32);
}

__device__ __forceinline__ uint64_t fitch64(uint64_t a, uint64_t b, uint32_t* elo, uint32_t* ehi) {
  const W2 x = split(a), y = split(b);
  uint32_t ol, oh;
  *elo = fitch32(x.lo, y.lo, &ol);
  *ehi = fitch32(x.hi, y.hi, &oh);
  return join(ol, oh);
}

// Fitch (1971) final set of an internal node, per 32-bit half.
__device__ __forceinline__ uint32_t final32(uint32_t pv, uint32_t pa, uint32_t pb, uint32_t fp) {
  const uint32_t sub_ok = ~nzmask(fp & ~pv);       // F_p subset of P_v
  const uint32_t inter = nzmask(pa & pb);          // P_v was an intersection
  const uint32_t r = bfi(inter, pv | (fp & (pa | pb)), pv | fp);
  return bfi(sub_ok, fp, r);
}
__device__ __forceinline__ uint32_t leaf32(uint32_t fp, uint32_t l) {
  const uint32_t i = fp & l;
  return bfi(nzmask(i), i, l);
}

__device__ __forceinline__ uint64_t final64(uint64_t pv, uint64_t pa, uint64_t pb, uint64_t fp) {
  return join(final32((uint32_t)pv, (uint32_t)pa, (uint32_t)pb, (uint32_t)fp),
              final32((uint32_t)(pv >> 32), (uint32_t)(pa >> 32), (uint32_t)(pb >> 32),
                      (uint32_t)(fp >> 32)));
}
__device__ __forceinline__ uint64_t leaf64(uint64_t fp, uint64_t l) {
  return join(leaf32((uint32_t)fp, (uint32_t)l), leaf32((uint32_t)(fp >> 32), (uint32_t)(l >> 32)));
}

// weighted count of the flagged sites of word w (flags in bit 0 of nibbles)
__device__ __forceinline__ uint64_t weigh(uint32_t elo, uint32_t ehi, const uint32_t* __restrict__ wt,
                                          int64_t w) {
  uint64_t c = 0;
  uint64_t e = join(elo, ehi);
  while (e) {
    const int k = __builtin_ctzll(e) >> 2;
    c += wt[w * 16 + k];
    e &= e - 1;
  }
  return c;
}

// 0xF in every nibble of word w that holds a real site (the last word of a
// block may end in padding nibbles, which must not count as empty sites)
__device__ __forceinline__ uint64_t valid_nibbles(int64_t w, int64_t n_sites) {
  const int64_t r = n_sites - w * 16;
  if (r <= 0) return 0ull;
  return r >= 16 ? ~0ull : ((1ull << (4 * r)) - 1ull);
}

__device__ __forceinline__ uint64_t wave_sum(uint64_t v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// block-wide sum of one uint64 per thread, added atomically to *dst
__device__ __forceinline__ void block_add(uint64_t v, unsigned long long* dst) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd(dst, (unsigned long long)v);
}

// ---------------------------------------------------------------------------
// Elementwise NodeData operations
// ---------------------------------------------------------------------------
enum Op { OP_MEDIAN2 = 0, OP_DIST1 = 1, OP_DIST2 = 2, OP_MEDIAN3 = 3, OP_ADJUST3 = 4 };

template <int OP>
__global__ __launch_bounds__(256) void k_fitch_elem(const uint64_t* __restrict__ a,
                                                    const uint64_t* __restrict__ b,
                                                    const uint64_t* __restrict__ c,
                                                    const uint64_t* __restrict__ d,
                                                    uint64_t* __restrict__ out, int64_t W,
                                                    int64_t n_sites,
                                                    const uint32_t* __restrict__ wt,
                                                    unsigned long long* __restrict__ cost) {
  uint64_t acc = 0;
  for (int64_t w = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; w < W;
       w += (int64_t)gridDim.x * blockDim.x) {
    uint32_t elo = 0, ehi = 0;
    const uint64_t vm = valid_nibbles(w, n_sites);
    if (OP == OP_MEDIAN2 || OP == OP_DIST1) {
      const uint64_t r = fitch64(a[w], b[w], &elo, &ehi);
      if (OP == OP_MEDIAN2) out[w] = r;
    } else if (OP == OP_DIST2) {
      uint32_t t0, t1;
      const uint64_t m = fitch64(a[w], b[w], &t0, &t1);
      (void)fitch64(m, c[w], &elo, &ehi);
    } else if (OP == OP_MEDIAN3) {
      // states present in all three / at least two / at least one
      const uint64_t x = a[w], y = b[w], z = c[w];
      const uint64_t all3 = x & y & z;
      const uint64_t two = (x & y) | (x & z) | (y & z);
      const uint64_t any = x | y | z;
      const W2 s3 = split(all3), s2 = split(two), s1 = split(any);
      const uint32_t m3l = nzmask(s3.lo), m3h = nzmask(s3.hi);
      const uint32_t m2l = nzmask(s2.lo), m2h = nzmask(s2.hi);
      const uint32_t rl = bfi(m3l, s3.lo, bfi(m2l, s2.lo, s1.lo));
      const uint32_t rh = bfi(m3h, s3.hi, bfi(m2h, s2.hi, s1.hi));
      out[w] = join(rl, rh);
      // cost 0 / 1 / 2: flag "not all three" and "not even two" separately
      const uint32_t vl = (uint32_t)vm & nzmask(s1.lo), vh = (uint32_t)(vm >> 32) & nzmask(s1.hi);
      const uint32_t n3l = (~m3l) & K1 & vl, n3h = (~m3h) & K1 & vh;
      const uint32_t n2l = (~m2l) & K1 & vl, n2h = (~m2h) & K1 & vh;
      if (wt) {
        acc += weigh(n3l, n3h, wt, w) + weigh(n2l, n2h, wt, w);
      } else {
        acc += __popc(n3l) + __popc(n3h) + __popc(n2l) + __popc(n2h);
      }
      continue;
    } else {  // OP_ADJUST3: a = P_v, b = P_a, c = P_b, d = F_p
      const uint64_t r = final64(a[w], b[w], c[w], d[w]);
      out[w] = r;
      const uint64_t diff = r ^ a[w];
      const W2 s = split(diff);
      acc += __popc(nzmask(s.lo) & K1) + __popc(nzmask(s.hi) & K1);
      continue;
    }
    elo &= (uint32_t)vm;
    ehi &= (uint32_t)(vm >> 32);
    if (wt) acc += weigh(elo, ehi, wt, w);
    else acc += __popc(elo) + __popc(ehi);
  }
  if (cost) block_add(acc, cost);
}

// ---------------------------------------------------------------------------
// Whole-tree down-pass.  The passes work on 32-bit lanes (8 sites per thread;
// the uint64 layout read as uint32 pairs, little endian), which gives twice
// as many waves as 64-bit lanes to hide the per-node latency chain.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t valid_nibbles32(int64_t w, int64_t n_sites) {
  const int64_t r = n_sites - w * 8;
  if (r <= 0) return 0u;
  return r >= 8 ? ~0u : (uint32_t)((1ull << (4 * r)) - 1ull);
}

__device__ __forceinline__ uint64_t weigh32(uint32_t e, const uint32_t* __restrict__ wt, int64_t w) {
  uint64_t c = 0;
  while (e) {
    const int k = __builtin_ctz(e) >> 2;
    c += wt[w * 8 + k];
    e &= e - 1;
  }
  return c;
}

// V 32-bit words (8 sites each) per thread, loaded and stored as one
// 4V-byte access
template <int V>
struct alignas(4 * V) Words {
  uint32_t w[V];
};

// Row accesses, optionally non-temporal (streamed past the caches: every
// row is read once and written once per pass)
template <int V> struct RawVec;
template <> struct RawVec<1> { typedef uint32_t T; };
template <> struct RawVec<2> { typedef uint32_t T __attribute__((ext_vector_type(2))); };
template <> struct RawVec<4> { typedef uint32_t T __attribute__((ext_vector_type(4))); };

template <int V, bool NT>
__device__ __forceinline__ Words<V> ld_row(const Words<V>* p) {
  if constexpr (NT) {
    const typename RawVec<V>::T r = __builtin_nontemporal_load(reinterpret_cast<const typename RawVec<V>::T*>(p));
    return __builtin_bit_cast(Words<V>, r);
  } else {
    return *p;
  }
}
template <int V, bool NT>
__device__ __forceinline__ void st_row(Words<V>* p, const Words<V>& v) {
  if constexpr (NT)
    __builtin_nontemporal_store(__builtin_bit_cast(typename RawVec<V>::T, v),
                                reinterpret_cast<typename RawVec<V>::T*>(p));
  else
    *p = v;
}

template <int V, bool NT>
struct SetLoad {
  const Words<V>* sets;
  int64_t stride, wc;  // in units of V words
  __device__ __forceinline__ Words<V> operator()(int id) const {
    return ld_row<V, NT>(sets + (int64_t)id * stride + wc);
  }
};

// Pass modes: the whole tree in one program, or a tree split at cut nodes:
// the parts side by side (grid.y = part; per-site counts of a part go to
// scratch[part][lane][8V]) and then the top, which adds the parts' counts.
// A part's count at a site is at most its number of internal nodes, so when
// every part has at most 255 of them (narrow) the counts are stored as bytes
// (a quarter of the scratch traffic).
enum : int { MODE_WHOLE = 0, MODE_PART = 1, MODE_TOP = 2 };

// program k of a split program: meta = prog_off | n_ops | fifo_off, each
// n_parts + 1 long
#define SPLIT_SELECT(k)                                 \
  do {                                                  \
    prog += meta[(k)];                                  \
    n_ops = meta[(n_parts + 1) + (k)];                  \
    fifo_ids += meta[2 * (n_parts + 1) + (k)];          \
  } while (0)

// W = number of 32-bit words per node row; lane w owns words V*w .. V*w+V-1
template <int BATCH, bool NODE_COST, int MODE, int V, bool NT>
__global__ __launch_bounds__(256) void k_fitch_down(
    const int4* __restrict__ prog, int n_ops, const int32_t* __restrict__ fifo_ids,
    const int32_t* __restrict__ meta, int n_parts, uint4* __restrict__ scratch, int narrow,
    uint32_t* __restrict__ sets32, int64_t stride, int64_t W, int64_t n_sites, int N,
    const uint32_t* __restrict__ wt, uint32_t* __restrict__ site_cost,
    unsigned long long* __restrict__ node_cost, unsigned long long* __restrict__ total) {
  using T = Words<V>;
  extern __shared__ uint4 lds_dyn[];
  const int tid = threadIdx.x, bd = blockDim.x;
  const int part = MODE == MODE_PART ? (int)blockIdx.y : n_parts;
  if (MODE != MODE_WHOLE) SPLIT_SELECT(part);
  const int64_t L = W / V;                           // lanes (W is a multiple of V)
  const int64_t w = (int64_t)blockIdx.x * bd + tid;  // lane index
  const bool act = w < L;
  const int64_t wc = act ? w : 0;
  T* sets = reinterpret_cast<T*>(sets32);
  const int64_t strideT = stride / V;
  T* ring = reinterpret_cast<T*>(lds_dyn);
  T* stk = ring + (size_t)2 * BATCH * bd;
  OperandRing<BATCH, SetLoad<V, NT>, T> leaves;
  leaves.init(ring, fifo_ids, SetLoad<V, NT>{sets, strideT, wc}, bd, tid);

  uint32_t cnt[V];  // nibble counters, flushed every 15 nodes
  uint32_t site[8 * V];
  uint32_t vm[V];
#pragma unroll
  for (int i = 0; i < V; ++i) {
    cnt[i] = 0;
    vm[i] = valid_nibbles32(wc * V + i, n_sites);
  }
#pragma unroll
  for (int k = 0; k < 8 * V; ++k) site[k] = 0;
  int since = 0;
  T acc = {};
  int sp = 0;
  for_each_entry(prog, n_ops, [&](int, const int4 op) {
    const int fl = op.w;
    if (fl & DOWN_SPILL) {
      stk[sp * bd + tid] = acc;
      ++sp;
    }
    const int sa = fl & 3, sb = (fl >> 2) & 3;
    T x, y;
    if (sa == SRC_LEAF) x = leaves.get();
    else if (sa == SRC_ACC) x = acc;
    else { --sp; x = stk[sp * bd + tid]; }
    if (sb == SRC_LEAF) y = leaves.get();
    else y = acc;  // SRC_ACC (B is never taken from the stack)
    uint32_t e[V];
#pragma unroll
    for (int i = 0; i < V; ++i) {
      uint32_t r;
      e[i] = fitch32(x.w[i], y.w[i], &r);
      acc.w[i] = r;
      cnt[i] += e[i];
    }
    if (act) st_row<V, NT>(sets + (int64_t)op.x * strideT + w, acc);
    if (NODE_COST) {
      uint64_t c = 0;
      if (act) {
#pragma unroll
        for (int i = 0; i < V; ++i) {
          const uint32_t m = e[i] & vm[i];
          c += wt ? weigh32(m, wt, w * V + i) : (uint64_t)__popc(m);
        }
      }
      c = wave_sum(c);
      if ((tid & 63) == 0 && c) atomicAdd(&node_cost[op.x - N], (unsigned long long)c);
    }
    if (++since == 15) {
#pragma unroll
      for (int i = 0; i < V; ++i) {
#pragma unroll
        for (int k = 0; k < 8; ++k) site[8 * i + k] += (cnt[i] >> (4 * k)) & 15u;
        cnt[i] = 0;
      }
      since = 0;
    }
  });
#pragma unroll
  for (int i = 0; i < V; ++i)
#pragma unroll
    for (int k = 0; k < 8; ++k) site[8 * i + k] += (cnt[i] >> (4 * k)) & 15u;
  if (MODE == MODE_PART) {
    if (act) {
      if (narrow) {  // 4 counts per word, 8V bytes per lane
        uint32_t* dst = reinterpret_cast<uint32_t*>(scratch) + ((size_t)part * L + w) * 2 * V;
#pragma unroll
        for (int q = 0; q < 2 * V; ++q)
          dst[q] = site[4 * q] | site[4 * q + 1] << 8 | site[4 * q + 2] << 16 | site[4 * q + 3] << 24;
      } else {
        uint4* dst = scratch + ((size_t)part * L + w) * 2 * V;
#pragma unroll
        for (int q = 0; q < 2 * V; ++q)
          dst[q] = make_uint4(site[4 * q], site[4 * q + 1], site[4 * q + 2], site[4 * q + 3]);
      }
    }
    return;
  }
  if (MODE == MODE_TOP && act) {
    if (narrow) {
#pragma unroll 4
      for (int pq = 0; pq < n_parts; ++pq) {
        const uint32_t* src = reinterpret_cast<const uint32_t*>(scratch) + ((size_t)pq * L + w) * 2 * V;
#pragma unroll
        for (int q = 0; q < 2 * V; ++q) {
          const uint32_t x = src[q];
#pragma unroll
          for (int b = 0; b < 4; ++b) site[4 * q + b] += (x >> (8 * b)) & 255u;
        }
      }
    } else {
#pragma unroll 4
      for (int pq = 0; pq < n_parts; ++pq) {
        const uint4* src = scratch + ((size_t)pq * L + w) * 2 * V;
#pragma unroll
        for (int q = 0; q < 2 * V; ++q) {
          const uint4 x = src[q];
          site[4 * q] += x.x;
          site[4 * q + 1] += x.y;
          site[4 * q + 2] += x.z;
          site[4 * q + 3] += x.w;
        }
      }
    }
  }
  uint64_t tot = 0;
  if (act) {
    const int64_t s0 = w * 8 * V;
    if (s0 + 8 * V <= n_sites) {
      uint4* dst = reinterpret_cast<uint4*>(site_cost + s0);
#pragma unroll
      for (int q = 0; q < 2 * V; ++q)
        dst[q] = make_uint4(site[4 * q], site[4 * q + 1], site[4 * q + 2], site[4 * q + 3]);
      if (wt) {
        const uint4* ws = reinterpret_cast<const uint4*>(wt + s0);
#pragma unroll
        for (int q = 0; q < 2 * V; ++q) {
          const uint4 xw = ws[q];
          tot += (uint64_t)site[4 * q] * xw.x + (uint64_t)site[4 * q + 1] * xw.y +
                 (uint64_t)site[4 * q + 2] * xw.z + (uint64_t)site[4 * q + 3] * xw.w;
        }
      } else {
#pragma unroll
        for (int k = 0; k < 8 * V; ++k) tot += site[k];
      }
    } else {
#pragma unroll
      for (int k = 0; k < 8 * V; ++k)
        if (s0 + k < n_sites) {
          site_cost[s0 + k] = site[k];
          tot += (uint64_t)site[k] * (wt ? wt[s0 + k] : 1u);
        }
    }
  }
  block_add(tot, total);
}

// ---------------------------------------------------------------------------
// Whole-tree up-pass
// ---------------------------------------------------------------------------
template <int BATCH, int MODE, int V, bool NT>
__global__ __launch_bounds__(256) void k_fitch_up(const int4* __restrict__ prog, int n_ops,
                                                  const int32_t* __restrict__ fifo_ids,
                                                  const int32_t* __restrict__ meta, int n_parts,
                                                  const int32_t* __restrict__ part_parent,
                                                  const uint32_t* __restrict__ sets32,
                                                  uint32_t* __restrict__ fin32, int64_t stride,
                                                  int64_t W, int N) {
  using T = Words<V>;
  extern __shared__ uint4 lds_dyn[];
  const int tid = threadIdx.x, bd = blockDim.x;
  // up: the top is program 0, part k is program 1 + k
  const int part = MODE == MODE_PART ? (int)blockIdx.y : -1;
  if (MODE != MODE_WHOLE) SPLIT_SELECT(part + 1);
  const int64_t L = W / V;
  const int64_t w = (int64_t)blockIdx.x * bd + tid;
  const bool act = w < L;
  const int64_t wc = act ? w : 0;
  const T* sets = reinterpret_cast<const T*>(sets32);
  T* fin = reinterpret_cast<T*>(fin32);
  const int64_t strideT = stride / V;
  T* ring = reinterpret_cast<T*>(lds_dyn);
  T* stk = ring + (size_t)2 * BATCH * bd;  // pairs (P, F)
  OperandRing<BATCH, SetLoad<V, NT>, T> nodes;
  nodes.init(ring, fifo_ids, SetLoad<V, NT>{sets, strideT, wc}, bd, tid);

  T accP = {}, accF = {};
  int sp = 0;
  for_each_entry(prog, n_ops, [&](int, const int4 op) {
    const int fl = op.w;
    const int src = fl & 3;
    T pv, fp = {};
    if (src == SRC_LEAF) pv = nodes.get();
    else if (src == SRC_ACC) { pv = accP; fp = accF; }
    else if (MODE == MODE_PART && src == SRC_EXT) {  // a part's root: parent final from the top pass
      pv = nodes.get();
      fp = fin[(int64_t)part_parent[part] * strideT + wc];
    }
    else {
      --sp;
      pv = stk[(2 * sp) * bd + tid];
      fp = stk[(2 * sp + 1) * bd + tid];
    }
    const T pa = nodes.get();
    const T pb = nodes.get();
    T fv, la, lb;
#pragma unroll
    for (int i = 0; i < V; ++i) {
      fv.w[i] = (src == SRC_LEAF) ? pv.w[i] : final32(pv.w[i], pa.w[i], pb.w[i], fp.w[i]);
      la.w[i] = leaf32(fv.w[i], pa.w[i]);
      lb.w[i] = leaf32(fv.w[i], pb.w[i]);
    }
    if (act) st_row<V, NT>(fin + (int64_t)op.x * strideT + w, fv);
    if (fl & UP_A_LEAF) { if (act) st_row<V, NT>(fin + (int64_t)op.y * strideT + w, la); }
    if (fl & UP_B_LEAF) { if (act) st_row<V, NT>(fin + (int64_t)op.z * strideT + w, lb); }
    if (fl & UP_PUSH) {
      stk[(2 * sp) * bd + tid] = (fl & UP_CONT_B) ? pa : pb;
      stk[(2 * sp + 1) * bd + tid] = fv;
      ++sp;
    }
    accP = (fl & UP_CONT_B) ? pb : pa;
    accF = fv;
  });
}

// Insertion costs of many (c, a, b) queries: block q sums, over the words of
// the sites, w_s [P_c & med(F_a, F_b) = 0] with med the Fitch median.
__global__ __launch_bounds__(256) void k_fitch_insert(const uint64_t* __restrict__ pre,
                                                      const uint64_t* __restrict__ fin,
                                                      int64_t stride, int64_t W, int64_t n_sites,
                                                      const uint32_t* __restrict__ wt,
                                                      const int32_t* __restrict__ qs,
                                                      unsigned long long* __restrict__ cost) {
  const int64_t q = blockIdx.x;
  const int32_t c = qs[3 * q], a = qs[3 * q + 1], b = qs[3 * q + 2];
  const uint64_t* pc = pre + (int64_t)c * stride;
  const uint64_t* fa = fin + (int64_t)a * stride;
  const uint64_t* fb = fin + (int64_t)b * stride;
  uint64_t acc = 0;
  for (int64_t w = threadIdx.x; w < W; w += blockDim.x) {
    uint32_t t0, t1, elo, ehi;
    const uint64_t m = fitch64(fa[w], fb[w], &t0, &t1);
    (void)fitch64(m, pc[w], &elo, &ehi);
    const uint64_t vm = valid_nibbles(w, n_sites);
    elo &= (uint32_t)vm;
    ehi &= (uint32_t)(vm >> 32);
    acc += wt ? weigh(elo, ehi, wt, w) : (uint64_t)(__popc(elo) + __popc(ehi));
  }
  acc = wave_sum(acc);
  __shared__ unsigned long long sh[4];
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    unsigned long long t = 0;
    for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    cost[q] = t;
  }
}

int block_size_env(const char* name, int dflt) {
  const char* s = getenv(name);
  if (!s) return dflt;
  const int v = atoi(s);
  return (v == 64 || v == 128 || v == 256) ? v : dflt;
}


}  // namespace

static int elem_launch_checks(phy_ctx* ctx, int64_t n_sites) {
  if (!ctx) return PHY_EINVAL;
  if (n_sites < 0) return set_err(ctx, PHY_EINVAL, "n_sites < 0");
  return PHY_OK;
}

template <int OP>
static int run_elem(phy_ctx* ctx, const uint64_t* a, const uint64_t* b, const uint64_t* c,
                    const uint64_t* d, uint64_t* out, int64_t n_sites, const uint32_t* wt,
                    uint64_t* cost) {
  int rc = elem_launch_checks(ctx, n_sites);
  if (rc) return rc;
  const int64_t W = phy_words(n_sites);
  if (cost && (rc = phy_zero_u64(ctx, cost, 1))) return rc;
  if (W == 0) return PHY_OK;
  const int bs = 256;
  int64_t g = (W + bs - 1) / bs;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(k_fitch_elem<OP>, dim3((unsigned)g), dim3(bs), 0, ctx->stream, a, b, c, d, out,
                     W, n_sites, wt, (unsigned long long*)cost);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

extern "C" {

int phy_fitch_median2(phy_ctx* ctx, const uint64_t* x, const uint64_t* y, uint64_t* out,
                      int64_t n_sites, const uint32_t* wt, uint64_t* cost) {
  if (ctx && n_sites > 0 && (!x || !y || !out)) return set_err(ctx, PHY_EINVAL, "median2: null block");
  return run_elem<OP_MEDIAN2>(ctx, x, y, nullptr, nullptr, out, n_sites, wt, cost);
}

int phy_fitch_median3(phy_ctx* ctx, const uint64_t* a, const uint64_t* b, const uint64_t* c,
                      uint64_t* out, int64_t n_sites, const uint32_t* wt, uint64_t* cost) {
  if (ctx && n_sites > 0 && (!a || !b || !c || !out)) return set_err(ctx, PHY_EINVAL, "median3: null block");
  return run_elem<OP_MEDIAN3>(ctx, a, b, c, nullptr, out, n_sites, wt, cost);
}

int phy_fitch_distance1(phy_ctx* ctx, const uint64_t* a, const uint64_t* b, int64_t n_sites,
                        const uint32_t* wt, uint64_t* cost) {
  if (ctx && (!cost || (n_sites > 0 && (!a || !b)))) return set_err(ctx, PHY_EINVAL, "distance1: null argument");
  return run_elem<OP_DIST1>(ctx, a, b, nullptr, nullptr, nullptr, n_sites, wt, cost);
}

int phy_fitch_distance2(phy_ctx* ctx, const uint64_t* a, const uint64_t* b, const uint64_t* c,
                        int64_t n_sites, const uint32_t* wt, uint64_t* cost) {
  if (ctx && (!cost || (n_sites > 0 && (!a || !b || !c)))) return set_err(ctx, PHY_EINVAL, "distance2: null argument");
  return run_elem<OP_DIST2>(ctx, a, b, c, nullptr, nullptr, n_sites, wt, cost);
}

int phy_fitch_adjust3(phy_ctx* ctx, const uint64_t* pv, const uint64_t* pa, const uint64_t* pb,
                      const uint64_t* fp, uint64_t* out, int64_t n_sites, uint64_t* changed) {
  if (ctx && n_sites > 0 && (!pv || !pa || !pb || !fp || !out)) return set_err(ctx, PHY_EINVAL, "adjust3: null block");
  return run_elem<OP_ADJUST3>(ctx, pv, pa, pb, fp, out, n_sites, nullptr, changed);
}

}  // extern "C"

namespace {

// one thread per output word: the 16 selected sites' nibbles (and weights)
__global__ __launch_bounds__(256) void k_fitch_select(const uint64_t* __restrict__ src,
                                                      const uint32_t* __restrict__ w_src,
                                                      const int64_t* __restrict__ idx, int64_t n_idx,
                                                      uint64_t* __restrict__ dst,
                                                      uint32_t* __restrict__ w_dst) {
  const int64_t word = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (word >= (n_idx + 15) / 16) return;
  uint64_t out = 0;
  for (int j = 0; j < 16; ++j) {
    const int64_t s = word * 16 + j;
    if (s >= n_idx) break;
    const int64_t k = idx[s];
    out |= ((src[k >> 4] >> (4 * (k & 15))) & 15ull) << (4 * j);
    if (w_dst) w_dst[s] = w_src[k];
  }
  dst[word] = out;
}

}  // namespace

extern "C" {

int phy_fitch_select(phy_ctx* ctx, const uint64_t* d_states, int64_t n_sites,
                     const uint32_t* d_weights, const int64_t* h_idx, int64_t n_idx,
                     uint64_t* d_states_out, uint32_t* d_weights_out) {
  if (!ctx) return PHY_EINVAL;
  if (n_sites < 0 || n_idx < 0) return set_err(ctx, PHY_EINVAL, "select: negative size");
  if (n_idx == 0) return PHY_OK;
  if (!d_states || !h_idx || !d_states_out) return set_err(ctx, PHY_EINVAL, "select: null argument");
  if ((d_weights == nullptr) != (d_weights_out == nullptr))
    return set_err(ctx, PHY_EINVAL, "select: weights in and out must both be given or both be null");
  for (int64_t s = 0; s < n_idx; ++s)
    if (h_idx[s] < 0 || h_idx[s] >= n_sites) return set_err(ctx, PHY_EINVAL, "select: site index out of range");
  int64_t* d_idx = nullptr;
  if (hipMalloc((void**)&d_idx, (size_t)n_idx * sizeof(int64_t)) != hipSuccess) {
    (void)hipGetLastError();
    return set_err(ctx, PHY_ENOMEM, "select: index buffer");
  }
  int rc = PHY_OK;
  if (hipMemcpyAsync(d_idx, h_idx, (size_t)n_idx * sizeof(int64_t), hipMemcpyHostToDevice, ctx->stream) != hipSuccess) {
    rc = set_err(ctx, PHY_EDEVICE, "select: index upload");
  } else {
    const int64_t words = (n_idx + 15) / 16;
    hipLaunchKernelGGL(k_fitch_select, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, ctx->stream,
                       d_states, d_weights, d_idx, n_idx, d_states_out, d_weights_out);
    if (hipGetLastError() != hipSuccess) rc = set_err(ctx, PHY_EDEVICE, "select: launch");
  }
  // the index buffer is released once the stream has used it
  if (hipStreamSynchronize(ctx->stream) != hipSuccess && rc == PHY_OK) rc = set_err(ctx, PHY_EDEVICE, "select: sync");
  (void)hipFree(d_idx);
  return rc;
}

}  // extern "C"

namespace {

// Split passes (parts side by side, then the top) unless disabled; the
// whole-tree program otherwise.
bool use_split(const phy_topo* t) {
  const char* e = getenv("PHY_FITCH_SPLIT");
  return t->parts.size() >= 2 && t->split_down.d_prog && !(e && e[0] == '0');
}

template <int BATCH, int V, bool NT>
int downpass_launch(phy_ctx* ctx, const phy_topo* topo, uint32_t* s32, int64_t stride32,
                    int64_t W32, int64_t n_sites, const uint32_t* wt, uint32_t* site_cost,
                    unsigned long long* node_cost, unsigned long long* total, int bs) {
  const int N = topo->n_leaves;
  const bool split = use_split(topo);
  const int depth = split ? topo->split_down.depth : topo->down_depth;
  // LDS per thread: the operand ring (2 BATCH entries) and the stack, V words
  // each; the block shrinks (down to one wave) until they fit
  auto lds_of = [&](int b) { return (size_t)(2 * BATCH + depth) * b * 4 * V; };
  while (bs > 64 && lds_of(bs) > 64 * 1024) bs /= 2;
  const size_t lds = lds_of(bs);
  if (lds > 64 * 1024) return set_err(ctx, PHY_ERANGE, "downpass: LDS stack too deep for this tree");
  const unsigned grid = (unsigned)((W32 / V + bs - 1) / bs);
  if (!split) {
    KernelSpan span(ctx, "fitch_down");
    if (node_cost)
      hipLaunchKernelGGL((k_fitch_down<BATCH, true, MODE_WHOLE, V, NT>), dim3(grid), dim3(bs), lds, ctx->stream,
                         topo->d_down_prog, N - 1, topo->d_down_fifo, nullptr, 0, nullptr, 0, s32, stride32, W32,
                         n_sites, N, wt, site_cost, node_cost, total);
    else
      hipLaunchKernelGGL((k_fitch_down<BATCH, false, MODE_WHOLE, V, NT>), dim3(grid), dim3(bs), lds, ctx->stream,
                         topo->d_down_prog, N - 1, topo->d_down_fifo, nullptr, 0, nullptr, 0, s32, stride32, W32,
                         n_sites, N, wt, site_cost, node_cost, total);
    PHY_LAUNCHED(ctx);
    return PHY_OK;
  }
  const SplitProg& sp = topo->split_down;
  const int P = (int)topo->parts.size();
  const size_t need = (size_t)P * W32 * 8 * sizeof(uint32_t);  // 8 counts per 32-bit word
  if (topo->scratch_bytes < need) {
    if (topo->d_scratch) PHY_HIP(ctx, hipFree(topo->d_scratch));
    topo->d_scratch = nullptr;
    topo->scratch_bytes = 0;
    if (hipMalloc((void**)&topo->d_scratch, need) != hipSuccess) {
      (void)hipGetLastError();
      return set_err(ctx, PHY_ENOMEM, "downpass: scratch allocation failed");
    }
    topo->scratch_bytes = need;
  }
  uint4* scr = reinterpret_cast<uint4*>(topo->d_scratch);
  int max_ops = 0;  // internal nodes of the largest part
  for (int k = 0; k < P; ++k) max_ops = std::max(max_ops, (int)sp.n_ops[k]);
  // PHY_FITCH_NARROW=0 keeps 32-bit counts
  const char* ne = getenv("PHY_FITCH_NARROW");
  const int narrow = max_ops <= 255 && !(ne && ne[0] == '0');
  KernelSpan span(ctx, "fitch_down");
  if (node_cost) {
    hipLaunchKernelGGL((k_fitch_down<BATCH, true, MODE_PART, V, NT>), dim3(grid, (unsigned)P), dim3(bs), lds,
                       ctx->stream, sp.d_prog, 0, sp.d_fifo, sp.d_meta, P, scr, narrow, s32, stride32, W32, n_sites,
                       N, wt, site_cost, node_cost, total);
    PHY_LAUNCHED(ctx);
    hipLaunchKernelGGL((k_fitch_down<BATCH, true, MODE_TOP, V, NT>), dim3(grid), dim3(bs), lds, ctx->stream,
                       sp.d_prog, 0, sp.d_fifo, sp.d_meta, P, scr, narrow, s32, stride32, W32, n_sites, N, wt,
                       site_cost, node_cost, total);
  } else {
    hipLaunchKernelGGL((k_fitch_down<BATCH, false, MODE_PART, V, NT>), dim3(grid, (unsigned)P), dim3(bs), lds,
                       ctx->stream, sp.d_prog, 0, sp.d_fifo, sp.d_meta, P, scr, narrow, s32, stride32, W32, n_sites,
                       N, wt, site_cost, node_cost, total);
    PHY_LAUNCHED(ctx);
    hipLaunchKernelGGL((k_fitch_down<BATCH, false, MODE_TOP, V, NT>), dim3(grid), dim3(bs), lds, ctx->stream,
                       sp.d_prog, 0, sp.d_fifo, sp.d_meta, P, scr, narrow, s32, stride32, W32, n_sites, N, wt,
                       site_cost, node_cost, total);
  }
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

template <int BATCH, int V, bool NT>
int uppass_launch(phy_ctx* ctx, const phy_topo* topo, const uint32_t* s32, uint32_t* f32,
                  int64_t stride32, int64_t W32, int bs) {
  const int N = topo->n_leaves;
  const bool split = use_split(topo);
  const int depth = split ? topo->split_up.depth : topo->up_depth;
  auto lds_of = [&](int b) { return (size_t)(2 * BATCH + 2 * depth) * b * 4 * V; };
  while (bs > 64 && lds_of(bs) > 64 * 1024) bs /= 2;
  const size_t lds = lds_of(bs);
  if (lds > 64 * 1024) return set_err(ctx, PHY_ERANGE, "uppass: LDS stack too deep for this tree");
  const unsigned grid = (unsigned)((W32 / V + bs - 1) / bs);
  KernelSpan span(ctx, "fitch_up");
  if (!split) {
    hipLaunchKernelGGL((k_fitch_up<BATCH, MODE_WHOLE, V, NT>), dim3(grid), dim3(bs), lds, ctx->stream,
                       topo->d_up_prog, N - 1, topo->d_up_fifo, nullptr, 0, nullptr, s32, f32, stride32,
                       W32, N);
    PHY_LAUNCHED(ctx);
    return PHY_OK;
  }
  const SplitProg& sp = topo->split_up;
  const int P = (int)topo->parts.size();
  hipLaunchKernelGGL((k_fitch_up<BATCH, MODE_TOP, V, NT>), dim3(grid), dim3(bs), lds, ctx->stream, sp.d_prog,
                     0, sp.d_fifo, sp.d_meta, P, topo->d_part_parent, s32, f32, stride32, W32, N);
  PHY_LAUNCHED(ctx);
  hipLaunchKernelGGL((k_fitch_up<BATCH, MODE_PART, V, NT>), dim3(grid, (unsigned)P), dim3(bs), lds,
                     ctx->stream, sp.d_prog, 0, sp.d_fifo, sp.d_meta, P, topo->d_part_parent, s32, f32,
                     stride32, W32, N);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

// 32-bit words per thread: 1, 2 or 4 (PHY_FITCH_WORDS overrides the pass's
// default)
int words_env(int dflt) {
  const char* s = getenv("PHY_FITCH_WORDS");
  const int v = s ? atoi(s) : dflt;
  return (v == 1 || v == 2 || v == 4) ? v : dflt;
}

// operand prefetch depth (leaves / node sets per batch), 4 for both passes
// (a short ring keeps more waves resident); PHY_FITCH_BATCH=4|8|16|32
// overrides both
int batch_env(int dflt) {
  const char* s = getenv("PHY_FITCH_BATCH");
  const int v = s ? atoi(s) : dflt;
  return (v == 4 || v == 8 || v == 16 || v == 32) ? v : dflt;
}

// non-temporal row loads and stores (PHY_FITCH_NT=0|1 overrides the pass's
// default)
bool nt_env(bool dflt) {
  const char* s = getenv("PHY_FITCH_NT");
  return s ? s[0] == '1' : dflt;
}

// the widest lane (<= want words) whose accesses stay aligned and whose
// lanes tile the rows exactly
int lane_words(int want, const void* base, int64_t stride32, int64_t W32) {
  for (int v = want; v > 1; v /= 2)
    if (stride32 % v == 0 && W32 % v == 0 && reinterpret_cast<uintptr_t>(base) % (4 * v) == 0) return v;
  return 1;
}

}  // namespace

extern "C" {

int phy_fitch_downpass(phy_ctx* ctx, const phy_topo* topo, uint64_t* sets, int64_t stride,
                       int64_t n_sites, const uint32_t* wt, uint32_t* site_cost,
                       uint64_t* node_cost, uint64_t* total) {
  if (!ctx) return PHY_EINVAL;
  if (!topo || !topo->d_down_prog) return set_err(ctx, PHY_EINVAL, "downpass: topology not on a device");
  const int64_t W = phy_words(n_sites);
  if (n_sites < 0 || stride < W) return set_err(ctx, PHY_EINVAL, "downpass: stride < words");
  if (!sets || !site_cost || !total) return set_err(ctx, PHY_EINVAL, "downpass: null output");
  const int N = topo->n_leaves;
  int rc;
  if ((rc = phy_zero_u64(ctx, total, 1))) return rc;
  if (node_cost && (rc = phy_zero_u64(ctx, node_cost, N - 1))) return rc;
  if (W == 0) return PHY_OK;
  const int bs = block_size_env("PHY_FITCH_BLOCK", 256);
  const int64_t W32 = 2 * W;  // 32-bit lanes
  uint32_t* s32 = reinterpret_cast<uint32_t*>(sets);
  auto* nc = reinterpret_cast<unsigned long long*>(node_cost);
  auto* tt = reinterpret_cast<unsigned long long*>(total);
  // 16-byte non-temporal lanes and a 4-deep ring: at 1,024 taxa x 10^6
  // sites 0.243 ms against 0.266 ms for 8-byte cached lanes and an 8-deep
  // ring (sweep over words 1 / 2 / 4, ring 4 / 8 / 16, non-temporal or not,
  // blocks of 128 / 256, 8 / 16 / 32 parts)
  const int v = lane_words(words_env(4), s32, 2 * stride, W32);
#define PHY_DOWN(B, VV) (nt_env(true) ? downpass_launch<B, VV, true>(ctx, topo, s32, 2 * stride, W32, n_sites, wt, site_cost, nc, tt, bs) \
                              : downpass_launch<B, VV, false>(ctx, topo, s32, 2 * stride, W32, n_sites, wt, site_cost, nc, tt, bs))
  switch (batch_env(4)) {
    case 4: return v == 4 ? PHY_DOWN(4, 4) : v == 2 ? PHY_DOWN(4, 2) : PHY_DOWN(4, 1);
    case 8: return v == 4 ? PHY_DOWN(8, 4) : v == 2 ? PHY_DOWN(8, 2) : PHY_DOWN(8, 1);
    case 32: return v == 4 ? PHY_DOWN(32, 4) : v == 2 ? PHY_DOWN(32, 2) : PHY_DOWN(32, 1);
    default: return v == 4 ? PHY_DOWN(16, 4) : v == 2 ? PHY_DOWN(16, 2) : PHY_DOWN(16, 1);
  }
#undef PHY_DOWN
}

int phy_fitch_uppass(phy_ctx* ctx, const phy_topo* topo, const uint64_t* sets, uint64_t* fin,
                     int64_t stride, int64_t n_sites) {
  if (!ctx) return PHY_EINVAL;
  if (!topo || !topo->d_up_prog) return set_err(ctx, PHY_EINVAL, "uppass: topology not on a device");
  const int64_t W = phy_words(n_sites);
  if (n_sites < 0 || stride < W) return set_err(ctx, PHY_EINVAL, "uppass: stride < words");
  if (!sets || !fin) return set_err(ctx, PHY_EINVAL, "uppass: null buffer");
  if (W == 0) return PHY_OK;
  const int bs = block_size_env("PHY_FITCH_BLOCK", 256);
  const int64_t W32 = 2 * W;
  const uint32_t* s32 = reinterpret_cast<const uint32_t*>(sets);
  uint32_t* f32 = reinterpret_cast<uint32_t*>(fin);
  // 8-byte cached lanes, 4-deep ring: 0.485 ms at 1,024 x 10^6 (16-byte
  // non-temporal lanes 0.491 ms)
  const int v = std::min(lane_words(words_env(2), s32, 2 * stride, W32),
                         lane_words(words_env(2), f32, 2 * stride, W32));
#define PHY_UP(B, VV) (nt_env(false) ? uppass_launch<B, VV, true>(ctx, topo, s32, f32, 2 * stride, W32, bs) \
                            : uppass_launch<B, VV, false>(ctx, topo, s32, f32, 2 * stride, W32, bs))
  switch (batch_env(4)) {
    case 4: return v == 4 ? PHY_UP(4, 4) : v == 2 ? PHY_UP(4, 2) : PHY_UP(4, 1);
    case 8: return v == 4 ? PHY_UP(8, 4) : v == 2 ? PHY_UP(8, 2) : PHY_UP(8, 1);
    case 32: return v == 4 ? PHY_UP(32, 4) : v == 2 ? PHY_UP(32, 2) : PHY_UP(32, 1);
    default: return v == 4 ? PHY_UP(16, 4) : v == 2 ? PHY_UP(16, 2) : PHY_UP(16, 1);
  }
#undef PHY_UP
}

int phy_fitch_insertion_costs(phy_ctx* ctx, const uint64_t* pre, const uint64_t* fin,
                              int64_t stride, int64_t n_sites, const uint32_t* wt, int64_t n_q,
                              const int32_t* d_q, uint64_t* d_costs) {
  if (!ctx) return PHY_EINVAL;
  const int64_t W = phy_words(n_sites);
  if (n_sites < 0 || stride < W || n_q < 0 || (n_q > 0 && (!pre || !fin || !d_q || !d_costs)))
    return set_err(ctx, PHY_EINVAL, "insertion costs: bad argument");
  if (n_q == 0) return PHY_OK;
  if (n_q > 0x7fffffff) return set_err(ctx, PHY_ERANGE, "insertion costs: too many queries");
  hipLaunchKernelGGL(k_fitch_insert, dim3((unsigned)n_q), dim3(256), 0, ctx->stream, pre, fin, stride, W,
                     n_sites, wt, d_q, (unsigned long long*)d_costs);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

}  // extern "C"
