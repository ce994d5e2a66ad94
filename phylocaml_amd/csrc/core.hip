// Context, device memory helpers and topology compilation.
//
// A topology is compiled once into two uniform programs that every thread of
// a kernel walks for its own column of sites (the tree is shared by all
// sites, so control flow is wave-uniform and the program is read through the
// scalar cache):
//
//  * post-order: every internal node once, children before parents
//    (Tree.post_order_edges, lib/tree.ml:171-187).  Children are visited
//    heavier-first (Sethi & Ullman 1970 register need), the last result lives
//    in an accumulator register and earlier ones in a per-thread LDS stack, so
//    the stack depth is at most ceil(log2 N).  Leaves are consumed in a fixed
//    order, which lets the kernels prefetch them far ahead.
//  * pre-order: every internal node once, parents before children
//    (Tree.pre_order_nodes, lib/tree.ml:123-147), descending into the smaller
//    child first and stacking the larger; each node's preliminary set is
//    loaded exactly once, in a fixed order.
#include <algorithm>
#include <cstdlib>
#include <functional>

#include "common.hpp"

extern "C" int64_t phy_words(int64_t n_sites) { return n_sites <= 0 ? 0 : (n_sites + 15) / 16; }

__global__ void k_zero_u64(uint64_t* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = 0;
}

int phy_zero_u64(phy_ctx* ctx, uint64_t* d, int64_t n) {
  if (n <= 0) return PHY_OK;
  const int bs = 256;
  const int64_t g = (n + bs - 1) / bs;
  hipLaunchKernelGGL(k_zero_u64, dim3((unsigned)g), dim3(bs), 0, ctx->stream, d, n);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

extern "C" {

int phy_ctx_create(int device, void* stream, phy_ctx** out) {
  if (!out) return PHY_EINVAL;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return PHY_EDEVICE;
  if (device < 0 || device >= n) return PHY_EINVAL;
  if (hipSetDevice(device) != hipSuccess) return PHY_EDEVICE;
  phy_ctx* c = new phy_ctx();
  c->device = device;
  c->stream = (hipStream_t)stream;  // null = the default stream
  // a failed lookup leaves the allocation-free entry point unusable (it then
  // reports PHY_EDEVICE) and every other entry point as it is
  if (phy_noalloc_scratch_init(device, &c->noalloc_scratch) != PHY_OK) c->noalloc_scratch = nullptr;
  *out = c;
  return PHY_OK;
}

static void drop_spans(phy_ctx* ctx) {
  for (auto& s : ctx->spans) {
    (void)hipEventDestroy(s.start);
    (void)hipEventDestroy(s.stop);
  }
  ctx->spans.clear();
}

void phy_ctx_destroy(phy_ctx* ctx) {
  if (ctx) drop_spans(ctx);
  delete ctx;
}

int phy_ctx_kernel_timer(phy_ctx* ctx, int32_t enable) {
  if (!ctx) return PHY_EINVAL;
  drop_spans(ctx);
  ctx->totals.clear();
  ctx->timing = enable != 0;
  return PHY_OK;
}

int phy_ctx_kernel_time(phy_ctx* ctx, const char* label, double* total_ms, int64_t* count) {
  if (!ctx || !label || !total_ms || !count) return PHY_EINVAL;
  for (auto& s : ctx->spans) {
    float ms = 0.f;
    PHY_HIP(ctx, hipEventSynchronize(s.stop));
    PHY_HIP(ctx, hipEventElapsedTime(&ms, s.start, s.stop));
    auto it = std::find_if(ctx->totals.begin(), ctx->totals.end(),
                           [&](const auto& t) { return t.first == s.label; });
    if (it == ctx->totals.end()) {
      ctx->totals.push_back({s.label, {0.0, 0}});
      it = ctx->totals.end() - 1;
    }
    it->second.first += ms;
    it->second.second += 1;
  }
  drop_spans(ctx);
  *total_ms = 0.0;
  *count = 0;
  for (const auto& t : ctx->totals)
    if (t.first == label) {
      *total_ms = t.second.first;
      *count = t.second.second;
    }
  return PHY_OK;
}

// STREAM-style copy: one 16-byte vector per lane, non-temporal load and
// store, a grid that covers the buffer (the loop runs once unless the grid
// is capped), the tail byte by byte.  Measured on a 2 GiB copy: 6.46 TB/s
// read + write, against 5.58 TB/s for a grid-stride loop over 4 workgroups
// per CU and 5.89 TB/s for 16 KiB chunks per workgroup (waves dispatched in
// address order keep the DRAM pages they open busy).
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ void k_copy_stream(const u32x4* __restrict__ src, u32x4* __restrict__ dst, int64_t n16,
                              const unsigned char* __restrict__ src_tail,
                              unsigned char* __restrict__ dst_tail, int64_t n_tail) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n16; i += stride) {
    const u32x4 v = __builtin_nontemporal_load(src + i);
    __builtin_nontemporal_store(v, dst + i);
  }
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n_tail) dst_tail[t] = src_tail[t];
}

int phy_copy_stream(phy_ctx* ctx, void* d_dst, const void* d_src, size_t bytes) {
  if (!ctx || (bytes && (!d_dst || !d_src))) return PHY_EINVAL;
  if (((uintptr_t)d_dst | (uintptr_t)d_src) & 15)
    return set_err(ctx, PHY_EINVAL, "copy_stream: pointers must be 16-byte aligned");
  if (bytes == 0) return PHY_OK;
  const int64_t n16 = (int64_t)(bytes / 16), n_tail = (int64_t)(bytes % 16);
  const int bs = 256;
  const int64_t grid = std::max<int64_t>(1, std::min<int64_t>((n16 + bs - 1) / bs, 0x7fffffff));
  KernelSpan span(ctx, "copy_stream");
  hipLaunchKernelGGL(k_copy_stream, dim3((unsigned)grid), dim3(bs), 0, ctx->stream,
                     (const u32x4*)d_src, (u32x4*)d_dst, n16,
                     (const unsigned char*)d_src + n16 * 16, (unsigned char*)d_dst + n16 * 16,
                     n_tail);
  PHY_LAUNCHED(ctx);
  return PHY_OK;
}

const char* phy_last_error(const phy_ctx* ctx) { return ctx ? ctx->err.c_str() : "no context"; }

int phy_ctx_sync(phy_ctx* ctx) {
  if (!ctx) return PHY_EINVAL;
  PHY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PHY_OK;
}

int phy_ctx_device(const phy_ctx* ctx) { return ctx ? ctx->device : -1; }
void* phy_ctx_stream(const phy_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int phy_dev_alloc(phy_ctx* ctx, size_t bytes, void** d_out) {
  if (!ctx || !d_out) return PHY_EINVAL;
  *d_out = nullptr;
  if (bytes == 0) return PHY_OK;
  if (hipMalloc(d_out, bytes) != hipSuccess) {
    (void)hipGetLastError();
    return set_err(ctx, PHY_ENOMEM, "hipMalloc failed");
  }
  return PHY_OK;
}

int phy_dev_free(phy_ctx* ctx, void* d) {
  if (!ctx) return PHY_EINVAL;
  if (d) PHY_HIP(ctx, hipFree(d));
  return PHY_OK;
}

int phy_copy_h2d(phy_ctx* ctx, void* d_dst, const void* h_src, size_t bytes) {
  if (!ctx || (!d_dst && bytes) || (!h_src && bytes)) return PHY_EINVAL;
  if (bytes) PHY_HIP(ctx, hipMemcpyAsync(d_dst, h_src, bytes, hipMemcpyHostToDevice, ctx->stream));
  PHY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PHY_OK;
}

int phy_copy_d2h(phy_ctx* ctx, void* h_dst, const void* d_src, size_t bytes) {
  if (!ctx || (!d_src && bytes) || (!h_dst && bytes)) return PHY_EINVAL;
  if (bytes) PHY_HIP(ctx, hipMemcpyAsync(h_dst, d_src, bytes, hipMemcpyDeviceToHost, ctx->stream));
  PHY_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return PHY_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Program compilation (host)
// ---------------------------------------------------------------------------
namespace {

using Stop = std::function<bool(int32_t)>;  // nodes compiled as leaves

// post-order of the subtree under `top`, stop nodes as leaves
void post_order(int32_t N, int32_t top, const std::vector<int32_t>& ch, const Stop& stop,
                std::vector<int32_t>* post) {
  post->clear();
  std::vector<std::pair<int32_t, int>> st = {{top, 0}};
  while (!st.empty()) {
    auto& f = st.back();
    const int32_t v = f.first;
    if (stop(v)) { post->push_back(v); st.pop_back(); continue; }
    if (f.second < 2) {
      const int32_t c = ch[2 * (v - N) + f.second];
      f.second++;
      st.push_back({c, 0});
    } else {
      post->push_back(v);
      st.pop_back();
    }
  }
}

// Post-order program of the subtree under `top` with an accumulator and a
// stack (Sethi-Ullman order); returns false on an internal inconsistency.
bool compile_down(int32_t N, const std::vector<int32_t>& ch, int32_t top, const Stop& stop,
                  std::vector<int32_t>* prog, std::vector<int32_t>* fifo, int32_t* depth_out) {
  std::vector<int32_t> post;
  post_order(N, top, ch, stop, &post);
  std::vector<int32_t> need(2 * N - 1, 0);
  for (int32_t v : post) {
    if (stop(v)) continue;
    const int32_t a = ch[2 * (v - N)], b = ch[2 * (v - N) + 1];
    const int32_t na = stop(a) ? 0 : need[a], nb = stop(b) ? 0 : need[b];
    need[v] = (na == nb) ? na + 1 : std::max(na, nb);
  }
  prog->clear();
  fifo->clear();
  if (stop(top)) { *depth_out = 0; return true; }
  bool acc_live = false;
  int32_t depth = 0, max_depth = 0, n_int = 0;
  struct Frame { int32_t v; int stage; int32_t A, B; };
  std::vector<Frame> fs = {{top, 0, -1, -1}};
  while (!fs.empty()) {
    Frame& f = fs.back();
    if (f.stage == 0) {
      int32_t a = ch[2 * (f.v - N)], b = ch[2 * (f.v - N) + 1];
      const bool la = stop(a), lb = stop(b);
      if (la && !lb) std::swap(a, b);                              // internal child first
      else if (!la && !lb && need[b] > need[a]) std::swap(a, b);   // heavier first
      f.A = a;
      f.B = b;
      f.stage = 1;
      if (!stop(f.A)) { fs.push_back({f.A, 0, -1, -1}); continue; }
    }
    if (f.stage == 1) {
      f.stage = 2;
      if (!stop(f.B)) { fs.push_back({f.B, 0, -1, -1}); continue; }
    }
    const bool la = stop(f.A), lb = stop(f.B);
    int flags = 0, srcA, srcB;
    if (la && lb) {
      if (acc_live) {
        flags |= DOWN_SPILL;
        max_depth = std::max(max_depth, ++depth);
      }
      srcA = SRC_LEAF;
      srcB = SRC_LEAF;
    } else if (!la && lb) {
      srcA = SRC_ACC;
      srcB = SRC_LEAF;
    } else {  // both internal (A leaf & B internal never happens)
      srcA = SRC_STACK;
      srcB = SRC_ACC;
      depth--;
    }
    if (srcA == SRC_LEAF) fifo->push_back(f.A);
    if (srcB == SRC_LEAF) fifo->push_back(f.B);
    flags |= srcA | (srcB << 2);
    prog->insert(prog->end(), {f.v, f.A, f.B, flags});
    acc_live = true;
    ++n_int;
    fs.pop_back();
  }
  *depth_out = max_depth;
  return depth == 0 && (int32_t)post.size() == 2 * n_int + 1;
}

// Pre-order program of the subtree under `top`.  Stop nodes are not entered:
// a stop child below N is a leaf (its final set is written), above N a cut
// node (left to the program of its own part).  The first entry reads the
// subtree root's set from the fifo with source `first_src` (SRC_LEAF at the
// tree root: F = P; SRC_EXT at a cut node: the parent's final from memory).
bool compile_up(int32_t N, const std::vector<int32_t>& ch, const std::vector<int32_t>& size,
                int32_t top, const Stop& stop, int first_src, std::vector<int32_t>* prog,
                std::vector<int32_t>* fifo, int32_t* depth_out) {
  prog->clear();
  fifo->clear();
  *depth_out = 0;
  if (top < N) return true;
  fifo->push_back(top);
  std::vector<int32_t> stack;
  int32_t cur = top, depth = 0, max_depth = 0;
  int src = first_src;
  for (;;) {
    const int32_t a = ch[2 * (cur - N)], b = ch[2 * (cur - N) + 1];
    fifo->push_back(a);
    fifo->push_back(b);
    const bool sa = stop(a), sb = stop(b);
    int flags = src;
    if (sa) flags |= a < N ? UP_A_LEAF : UP_A_CUT;
    if (sb) flags |= b < N ? UP_B_LEAF : UP_B_CUT;
    int32_t next = -1;
    if (!sa && !sb) {
      // continue with the smaller subtree, push the larger
      const bool contb = size[b] < size[a];
      if (contb) flags |= UP_CONT_B;
      flags |= UP_PUSH;
      stack.push_back(contb ? a : b);
      max_depth = std::max(max_depth, ++depth);
      next = contb ? b : a;
    } else if (!sa) {
      next = a;
    } else if (!sb) {
      flags |= UP_CONT_B;
      next = b;
    }
    prog->insert(prog->end(), {cur, a, b, flags});
    if (next >= 0) {
      cur = next;
      src = SRC_ACC;
    } else if (!stack.empty()) {
      cur = stack.back();
      stack.pop_back();
      depth--;
      src = SRC_STACK;
    } else {
      break;
    }
  }
  *depth_out = max_depth;
  return true;
}

// Cut nodes for side-by-side parts: split the largest internal subtree of
// the frontier until it holds `want` internal nodes or the largest one has
// fewer than `min_leaves` leaves.
std::vector<int32_t> choose_cuts(int32_t N, const std::vector<int32_t>& ch,
                                 const std::vector<int32_t>& size, int32_t root, int want,
                                 int32_t min_leaves) {
  std::vector<int32_t> front = {root};
  for (;;) {
    int best = -1, n_int = 0;
    for (size_t k = 0; k < front.size(); ++k) {
      if (front[k] < N) continue;
      ++n_int;
      if (best < 0 || size[front[k]] > size[front[best]]) best = (int)k;
    }
    if (best < 0 || n_int >= want || size[front[best]] < 2 * min_leaves) break;
    const int32_t v = front[best];
    front.erase(front.begin() + best);
    front.push_back(ch[2 * (v - N)]);
    front.push_back(ch[2 * (v - N) + 1]);
  }
  std::vector<int32_t> cuts;
  for (int32_t v : front)
    if (v >= N) cuts.push_back(v);
  std::sort(cuts.begin(), cuts.end(), [&](int32_t x, int32_t y) { return size[x] > size[y]; });
  return cuts;
}

}  // namespace

void phy_compile_up_program(int32_t N, const std::vector<int32_t>& ch, const std::vector<int32_t>& size,
                            int32_t top, std::vector<int32_t>* prog, int32_t* depth,
                            const std::vector<char>* enter) {
  std::vector<int32_t> fifo;
  if (enter) {
    const Stop cut = [N, enter](int32_t v) { return v < N || !(*enter)[v]; };
    compile_up(N, ch, size, top, cut, SRC_LEAF, prog, &fifo, depth);
  } else {
    const Stop is_leaf = [N](int32_t v) { return v < N; };
    compile_up(N, ch, size, top, is_leaf, SRC_LEAF, prog, &fifo, depth);
  }
}

int phy_compile_topo(phy_topo* t) {
  const int32_t N = t->n_leaves, M = 2 * N - 1;
  const std::vector<int32_t>& ch = t->children;
  const Stop is_leaf = [N](int32_t v) { return v < N; };
  std::vector<int32_t> post;
  post_order(N, t->root, ch, is_leaf, &post);
  if ((int32_t)post.size() != M) return 1;
  t->subtree.assign(M, 0);
  for (int32_t v : post)
    t->subtree[v] = v < N ? 1 : t->subtree[ch[2 * (v - N)]] + t->subtree[ch[2 * (v - N) + 1]];
  if (!compile_down(N, ch, t->root, is_leaf, &t->down_prog, &t->down_fifo, &t->down_depth) ||
      (int32_t)t->down_prog.size() != 4 * (N - 1) || (int32_t)t->down_fifo.size() != N)
    return 2;
  compile_up(N, ch, t->subtree, t->root, is_leaf, SRC_LEAF, &t->up_prog, &t->up_fifo, &t->up_depth);
  if ((int32_t)t->up_prog.size() != 4 * (N - 1) || (int32_t)t->up_fifo.size() != M) return 3;

  // ---- split programs: parts under cut nodes run side by side ----
  t->parts.clear();
  // default: up to 16 parts of >= 64 leaves on average; PHY_FITCH_PARTS
  // asks for a count (parts of >= 16 leaves)
  const char* env_parts = getenv("PHY_FITCH_PARTS");
  const int want = env_parts ? (int)std::min<int32_t>(std::max(1, std::min(kMaxParts, atoi(env_parts))), N / 16)
                             : (int)std::min<int32_t>(16, N / 64);
  if (want >= 2) {
    const std::vector<int32_t> cuts = choose_cuts(N, ch, t->subtree, t->root, want, 16);
    if (cuts.size() >= 2) {
      std::vector<char> is_cut(M, 0);
      for (int32_t c : cuts) is_cut[c] = 1;
      const Stop stop_top = [N, &is_cut](int32_t v) { return v < N || is_cut[v]; };
      SplitProg& sd = t->split_down;
      SplitProg& su = t->split_up;
      sd = SplitProg();
      su = SplitProg();
      std::vector<int32_t> pg, ff;
      int32_t d = 0, n_ops = 0;
      auto add = [&](SplitProg& sp) {
        sp.prog_off.push_back((int32_t)sp.prog.size() / 4);
        sp.n_ops.push_back((int32_t)pg.size() / 4);
        sp.fifo_off.push_back((int32_t)sp.fifo.size());
        sp.prog.insert(sp.prog.end(), pg.begin(), pg.end());
        std::vector<int32_t> pad(phy_fifo_padded(ff.size()), ff.empty() ? 0 : ff[0]);
        std::copy(ff.begin(), ff.end(), pad.begin());
        sp.fifo.insert(sp.fifo.end(), pad.begin(), pad.end());
        sp.depth = std::max(sp.depth, d);
        n_ops += (int32_t)pg.size() / 4;
      };
      // down: parts first, then the top (cut nodes read as leaves)
      for (int32_t c : cuts) {
        if (!compile_down(N, ch, c, is_leaf, &pg, &ff, &d)) return 4;
        add(sd);
      }
      if (!compile_down(N, ch, t->root, stop_top, &pg, &ff, &d)) return 4;
      add(sd);
      if (n_ops != N - 1) return 5;
      // up: the top first, then the parts (entered from the parent's final)
      n_ops = 0;
      compile_up(N, ch, t->subtree, t->root, stop_top, SRC_LEAF, &pg, &ff, &d);
      add(su);
      for (int32_t c : cuts) {
        compile_up(N, ch, t->subtree, c, is_leaf, SRC_EXT, &pg, &ff, &d);
        add(su);
      }
      if (n_ops != N - 1) return 6;
      for (int32_t c : cuts) t->parts.push_back(c);
    }
  }

  // ---- split form of the fused pruning pass ----
  // default: up to 16 parts of >= 128 leaves on average; PHY_PRUNE_PARTS
  // asks for a count (parts of >= 16 leaves)
  t->pparts.clear();
  t->psplit = SplitProg();
  t->ptop_prog.clear();
  t->ptop_fifo.clear();
  const char* env_pp = getenv("PHY_PRUNE_PARTS");
  const int want_p = env_pp ? (int)std::min<int32_t>(std::max(1, std::min(kMaxParts, atoi(env_pp))), N / 16)
                            : (int)std::min<int32_t>(16, N / 128);
  if (want_p >= 2) {
    const std::vector<int32_t> cuts = choose_cuts(N, ch, t->subtree, t->root, want_p, 16);
    if (cuts.size() >= 2) {
      const int P = (int)cuts.size();
      std::vector<int32_t> slot(M, -1);
      for (int k = 0; k < P; ++k) slot[cuts[k]] = k;
      SplitProg& ps = t->psplit;
      std::vector<int32_t> pg, ff;
      int32_t d = 0, n_ops = 0;
      for (int32_t c : cuts) {
        if (!compile_down(N, ch, c, is_leaf, &pg, &ff, &d)) return 8;
        ps.prog_off.push_back((int32_t)ps.prog.size() / 4);
        ps.n_ops.push_back((int32_t)pg.size() / 4);
        ps.fifo_off.push_back((int32_t)ps.fifo.size());
        ps.prog.insert(ps.prog.end(), pg.begin(), pg.end());
        std::vector<int32_t> pad(phy_fifo_padded(ff.size()), ff[0]);
        std::copy(ff.begin(), ff.end(), pad.begin());
        ps.fifo.insert(ps.fifo.end(), pad.begin(), pad.end());
        ps.depth = std::max(ps.depth, d);
        n_ops += (int32_t)pg.size() / 4;
      }
      // the top: cut nodes stop the walk; their "leaf" operands become
      // stored operands and leave the leaf stream
      const Stop stop_top = [N, &slot](int32_t v) { return v < N || slot[v] >= 0; };
      if (!compile_down(N, ch, t->root, stop_top, &pg, &ff, &d)) return 8;
      t->ptop_depth = d;
      n_ops += (int32_t)pg.size() / 4;
      if (n_ops != N - 1) return 9;
      std::vector<int32_t> leaves;
      for (size_t i = 0; i < pg.size() / 4; ++i) {
        const int32_t A = pg[4 * i + 1], B = pg[4 * i + 2], fl = pg[4 * i + 3];
        int sa = fl & 3, sb = (fl >> 2) & 3, ka = 0, kb = 0;
        if (sa == SRC_LEAF && slot[A] >= 0) sa = SRC_EXT, ka = slot[A];
        if (sb == SRC_LEAF && slot[B] >= 0) sb = SRC_EXT, kb = slot[B];
        if (sa == SRC_LEAF) leaves.push_back(A);
        if (sb == SRC_LEAF) leaves.push_back(B);
        const int32_t v = pg[4 * i];
        t->ptop_prog.insert(t->ptop_prog.end(), {v, A, B, (fl & ~15) | sa | (sb << 2) | (ka << 8) | (kb << 16)});
      }
      // the leaf stream must be the walk's own with the cut nodes left out
      size_t j = 0;
      for (int32_t v : ff) {
        if (slot[v] >= 0) continue;
        if (j >= leaves.size() || v != leaves[j]) return 10;
        ++j;
      }
      if (j != leaves.size() || ff.size() != leaves.size() + cuts.size()) return 10;
      t->ptop_fifo.assign(phy_fifo_padded(leaves.size()), leaves.empty() ? 0 : leaves[0]);
      std::copy(leaves.begin(), leaves.end(), t->ptop_fifo.begin());
      t->pparts = cuts;
    }
  }
  return 0;
}

// device copies of a split program: programs, leaf streams and the offsets
// [prog_off | n_ops | fifo_off], one per program each
static hipError_t upload_split(SplitProg* sp) {
  std::vector<int32_t> meta;
  meta.insert(meta.end(), sp->prog_off.begin(), sp->prog_off.end());
  meta.insert(meta.end(), sp->n_ops.begin(), sp->n_ops.end());
  meta.insert(meta.end(), sp->fifo_off.begin(), sp->fifo_off.end());
  hipError_t e = hipMalloc((void**)&sp->d_prog, sp->prog.size() * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&sp->d_fifo, sp->fifo.size() * 4);
  if (e == hipSuccess) e = hipMalloc((void**)&sp->d_meta, meta.size() * 4);
  if (e == hipSuccess) e = hipMemcpy(sp->d_prog, sp->prog.data(), sp->prog.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(sp->d_fifo, sp->fifo.data(), sp->fifo.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(sp->d_meta, meta.data(), meta.size() * 4, hipMemcpyHostToDevice);
  return e;
}

extern "C" {

int phy_topo_create(phy_ctx* ctx, int32_t N, const int32_t* children, int32_t root,
                    phy_topo** out) {
  // ctx == NULL compiles the programs on the host only (no device upload):
  // enough for phy_topo_info / phy_topo_programs.
  if (!out) return PHY_EINVAL;
  *out = nullptr;
  if (N < 2 || !children) return set_err(ctx, PHY_EINVAL, "topology: need N >= 2 and children");
  if (N > (1 << 24)) return set_err(ctx, PHY_ERANGE, "topology: too many leaves");
  const int32_t M = 2 * N - 1;
  if (root < N || root >= M) return set_err(ctx, PHY_EINVAL, "topology: root must be an internal node");
  std::vector<int32_t> parent(M, -1);
  for (int32_t k = 0; k < N - 1; ++k)
    for (int c = 0; c < 2; ++c) {
      const int32_t x = children[2 * k + c];
      if (x < 0 || x >= M) return set_err(ctx, PHY_EINVAL, "topology: child id out of range");
      if (x == N + k) return set_err(ctx, PHY_EINVAL, "topology: node is its own child");
      if (parent[x] >= 0) return set_err(ctx, PHY_EINVAL, "topology: node with two parents");
      parent[x] = N + k;
    }
  if (parent[root] >= 0) return set_err(ctx, PHY_EINVAL, "topology: root has a parent");
  for (int32_t v = 0; v < M; ++v)
    if (v != root && parent[v] < 0) return set_err(ctx, PHY_EINVAL, "topology: node without parent");
  // every node must reach the root without a cycle
  for (int32_t v = 0; v < M; ++v) {
    int32_t x = v, steps = 0;
    while (x != root) {
      x = parent[x];
      if (x < 0 || ++steps > M) return set_err(ctx, PHY_EINVAL, "topology: cycle or disconnected node");
    }
  }
  phy_topo* t = new phy_topo();
  t->ctx = ctx;
  t->n_leaves = N;
  t->root = root;
  t->children.assign(children, children + 2 * (N - 1));
  t->parent = parent;
  if (phy_compile_topo(t) != 0) {
    delete t;
    return set_err(ctx, PHY_EINVAL, "topology: program compilation failed");
  }
  if (!ctx) {
    *out = t;
    return PHY_OK;
  }
  hipError_t e = hipSuccess;
  std::vector<int32_t> dfifo(phy_fifo_padded(t->down_fifo.size()), t->down_fifo[0]);
  std::vector<int32_t> ufifo(phy_fifo_padded(t->up_fifo.size()), t->up_fifo[0]);
  std::copy(t->down_fifo.begin(), t->down_fifo.end(), dfifo.begin());
  std::copy(t->up_fifo.begin(), t->up_fifo.end(), ufifo.begin());
  const size_t pd = t->down_prog.size() * sizeof(int32_t), fd = dfifo.size() * sizeof(int32_t);
  const size_t pu = t->up_prog.size() * sizeof(int32_t), fu = ufifo.size() * sizeof(int32_t);
  if (e == hipSuccess) e = hipMalloc((void**)&t->d_down_prog, pd);
  if (e == hipSuccess) e = hipMalloc((void**)&t->d_down_fifo, fd);
  if (e == hipSuccess) e = hipMalloc((void**)&t->d_up_prog, pu);
  if (e == hipSuccess) e = hipMalloc((void**)&t->d_up_fifo, fu);
  if (e == hipSuccess) e = hipMemcpy(t->d_down_prog, t->down_prog.data(), pd, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->d_down_fifo, dfifo.data(), fd, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->d_up_prog, t->up_prog.data(), pu, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->d_up_fifo, ufifo.data(), fu, hipMemcpyHostToDevice);
  if (e == hipSuccess && !t->parts.empty()) {
    std::vector<int32_t> pp(t->parts.size());
    for (size_t k = 0; k < t->parts.size(); ++k) pp[k] = t->parent[t->parts[k]];
    for (SplitProg* sp : {&t->split_down, &t->split_up})
      if (e == hipSuccess) e = upload_split(sp);
    if (e == hipSuccess) e = hipMalloc((void**)&t->d_part_parent, pp.size() * 4);
    if (e == hipSuccess) e = hipMemcpy(t->d_part_parent, pp.data(), pp.size() * 4, hipMemcpyHostToDevice);
  }
  if (e == hipSuccess && !t->pparts.empty()) {
    if (e == hipSuccess) e = upload_split(&t->psplit);
    if (e == hipSuccess) e = hipMalloc((void**)&t->d_ptop_prog, t->ptop_prog.size() * 4);
    if (e == hipSuccess) e = hipMalloc((void**)&t->d_ptop_fifo, t->ptop_fifo.size() * 4);
    if (e == hipSuccess)
      e = hipMemcpy(t->d_ptop_prog, t->ptop_prog.data(), t->ptop_prog.size() * 4, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = hipMemcpy(t->d_ptop_fifo, t->ptop_fifo.data(), t->ptop_fifo.size() * 4, hipMemcpyHostToDevice);
  }
  if (e != hipSuccess) {
    phy_topo_destroy(t);
    return set_err(ctx, PHY_EDEVICE, std::string("topology upload: ") + hipGetErrorString(e));
  }
  *out = t;
  return PHY_OK;
}

void phy_topo_destroy(phy_topo* t) {
  if (!t) return;
  if (t->d_down_prog) (void)hipFree(t->d_down_prog);
  if (t->d_down_fifo) (void)hipFree(t->d_down_fifo);
  if (t->d_up_prog) (void)hipFree(t->d_up_prog);
  if (t->d_up_fifo) (void)hipFree(t->d_up_fifo);
  for (SplitProg* sp : {&t->split_down, &t->split_up, &t->psplit}) {
    if (sp->d_prog) (void)hipFree(sp->d_prog);
    if (sp->d_fifo) (void)hipFree(sp->d_fifo);
    if (sp->d_meta) (void)hipFree(sp->d_meta);
  }
  if (t->d_part_parent) (void)hipFree(t->d_part_parent);
  if (t->d_ptop_prog) (void)hipFree(t->d_ptop_prog);
  if (t->d_ptop_fifo) (void)hipFree(t->d_ptop_fifo);
  if (t->d_scratch) (void)hipFree(t->d_scratch);
  delete t;
}

int phy_topo_programs(const phy_topo* t, int32_t* down_prog, int32_t* down_fifo,
                      int32_t* up_prog, int32_t* up_fifo) {
  if (!t) return PHY_EINVAL;
  if (down_prog) std::copy(t->down_prog.begin(), t->down_prog.end(), down_prog);
  if (down_fifo) std::copy(t->down_fifo.begin(), t->down_fifo.end(), down_fifo);
  if (up_prog) std::copy(t->up_prog.begin(), t->up_prog.end(), up_prog);
  if (up_fifo) std::copy(t->up_fifo.begin(), t->up_fifo.end(), up_fifo);
  return PHY_OK;
}

int phy_topo_split(const phy_topo* t, int32_t* n_parts, int32_t* cut_nodes) {
  if (!t) return PHY_EINVAL;
  if (n_parts) *n_parts = (int32_t)t->parts.size();
  if (cut_nodes) std::copy(t->parts.begin(), t->parts.end(), cut_nodes);
  return PHY_OK;
}

int phy_topo_split_program(const phy_topo* t, int32_t up, int32_t k, int32_t* prog, int32_t* n_ops,
                           int32_t* fifo, int32_t* n_fifo) {
  if (!t || t->parts.empty()) return PHY_EINVAL;
  const SplitProg& sp = up ? t->split_up : t->split_down;
  if (k < 0 || k > (int32_t)t->parts.size()) return PHY_EINVAL;
  const int32_t n = sp.n_ops[k];
  // the fifo of program k ends where the next one's padded fifo begins; its
  // unpadded length is the number of operands the program reads
  int32_t nf = 0;
  for (int32_t e = 0; e < n; ++e) {
    const int fl = sp.prog[4 * (sp.prog_off[k] + e) + 3];
    if (up) nf += 2 + (((fl & 3) == SRC_LEAF || (fl & 3) == SRC_EXT) ? 1 : 0);
    else nf += (((fl & 3) == SRC_LEAF) ? 1 : 0) + ((((fl >> 2) & 3) == SRC_LEAF) ? 1 : 0);
  }
  if (n_ops) *n_ops = n;
  if (n_fifo) *n_fifo = nf;
  if (prog) std::copy(sp.prog.begin() + 4 * (size_t)sp.prog_off[k], sp.prog.begin() + 4 * ((size_t)sp.prog_off[k] + n), prog);
  if (fifo) std::copy(sp.fifo.begin() + sp.fifo_off[k], sp.fifo.begin() + sp.fifo_off[k] + nf, fifo);
  return PHY_OK;
}

int phy_topo_info(const phy_topo* t, int32_t* dd, int32_t* ud, int32_t* parent_out) {
  if (!t) return PHY_EINVAL;
  if (dd) *dd = t->down_depth;
  if (ud) *ud = t->up_depth;
  if (parent_out) std::copy(t->parent.begin(), t->parent.end(), parent_out);
  return PHY_OK;
}

}  // extern "C"
