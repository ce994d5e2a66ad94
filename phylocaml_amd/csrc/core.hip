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
  *out = c;
  return PHY_OK;
}

void phy_ctx_destroy(phy_ctx* ctx) { delete ctx; }

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

struct Compiler {
  int32_t N;
  const std::vector<int32_t>& ch;
  const std::vector<int32_t>& size;
  std::vector<int32_t> need;  // Sethi-Ullman register need
  std::vector<int32_t>* prog;
  std::vector<int32_t>* fifo;
  bool acc_live = false;
  int32_t depth = 0, max_depth = 0;

  Compiler(int32_t n, const std::vector<int32_t>& c, const std::vector<int32_t>& s)
      : N(n), ch(c), size(s), need(2 * n - 1, 0) {}

  void compute_need(const std::vector<int32_t>& post) {
    for (int32_t v : post) {
      if (v < N) { need[v] = 0; continue; }
      const int32_t a = ch[2 * (v - N)], b = ch[2 * (v - N) + 1];
      const int32_t na = a < N ? 0 : need[a], nb = b < N ? 0 : need[b];
      need[v] = (na == nb) ? na + 1 : std::max(na, nb);
    }
  }
  // first-evaluated child A, second B
  void order(int32_t v, int32_t* A, int32_t* B) const {
    int32_t a = ch[2 * (v - N)], b = ch[2 * (v - N) + 1];
    const bool la = a < N, lb = b < N;
    if (la && !lb) std::swap(a, b);                // internal child first
    else if (!la && !lb && need[b] > need[a]) std::swap(a, b);  // heavier first
    *A = a;
    *B = b;
  }
};

bool build_post_order(int32_t N, int32_t root, const std::vector<int32_t>& ch,
                      std::vector<int32_t>* post) {
  // iterative post-order
  post->clear();
  std::vector<std::pair<int32_t, int>> st = {{root, 0}};
  while (!st.empty()) {
    auto& top = st.back();
    const int32_t v = top.first;
    if (v < N) { post->push_back(v); st.pop_back(); continue; }
    if (top.second < 2) {
      const int32_t c = ch[2 * (v - N) + top.second];
      top.second++;
      st.push_back({c, 0});
    } else {
      post->push_back(v);
      st.pop_back();
    }
  }
  return (int32_t)post->size() == 2 * N - 1;
}

}  // namespace

int phy_compile_topo(phy_topo* t) {
  const int32_t N = t->n_leaves, M = 2 * N - 1;
  const std::vector<int32_t>& ch = t->children;
  std::vector<int32_t> post;
  if (!build_post_order(N, t->root, ch, &post)) return 1;
  t->subtree.assign(M, 0);
  for (int32_t v : post)
    t->subtree[v] = v < N ? 1 : t->subtree[ch[2 * (v - N)]] + t->subtree[ch[2 * (v - N) + 1]];

  // ---- post-order program with accumulator + stack ----
  Compiler c(N, ch, t->subtree);
  c.compute_need(post);
  t->down_prog.clear();
  t->down_fifo.clear();
  // recursive evaluation, made iterative with an explicit frame stack
  struct Frame { int32_t v; int stage; int32_t A, B; };
  std::vector<Frame> fs = {{t->root, 0, -1, -1}};
  while (!fs.empty()) {
    Frame& f = fs.back();
    if (f.stage == 0) {
      c.order(f.v, &f.A, &f.B);
      f.stage = 1;
      if (f.A >= N) { fs.push_back({f.A, 0, -1, -1}); continue; }
    }
    if (f.stage == 1) {
      f.stage = 2;
      if (f.B >= N) { fs.push_back({f.B, 0, -1, -1}); continue; }
    }
    // emit v
    const bool la = f.A < N, lb = f.B < N;
    int flags = 0;
    int srcA, srcB;
    if (la && lb) {
      if (c.acc_live) {
        flags |= DOWN_SPILL;
        c.depth++;
        c.max_depth = std::max(c.max_depth, c.depth);
      }
      srcA = SRC_LEAF;
      srcB = SRC_LEAF;
    } else if (!la && lb) {
      srcA = SRC_ACC;
      srcB = SRC_LEAF;
    } else {  // both internal (A leaf & B internal never happens)
      srcA = SRC_STACK;
      srcB = SRC_ACC;
      c.depth--;
    }
    if (srcA == SRC_LEAF) t->down_fifo.push_back(f.A);
    if (srcB == SRC_LEAF) t->down_fifo.push_back(f.B);
    flags |= srcA | (srcB << 2);
    t->down_prog.insert(t->down_prog.end(), {f.v, f.A, f.B, flags});
    c.acc_live = true;
    fs.pop_back();
  }
  t->down_depth = c.max_depth;
  if ((int32_t)t->down_prog.size() != 4 * (N - 1) || (int32_t)t->down_fifo.size() != N || c.depth != 0)
    return 2;

  // ---- pre-order program ----
  t->up_prog.clear();
  t->up_fifo.clear();
  t->up_fifo.push_back(t->root);
  // pending internal nodes: (node) on the stack; the accumulator holds at most one
  std::vector<int32_t> stack;
  int32_t cur = t->root;
  int src = SRC_LEAF;  // root
  int32_t depth = 0, max_depth = 0;
  for (;;) {
    const int32_t a = ch[2 * (cur - N)], b = ch[2 * (cur - N) + 1];
    t->up_fifo.push_back(a);
    t->up_fifo.push_back(b);
    int flags = src;
    if (a < N) flags |= UP_A_LEAF;
    if (b < N) flags |= UP_B_LEAF;
    int32_t next = -1;
    if (a >= N && b >= N) {
      // continue with the smaller subtree, push the larger
      const bool contb = t->subtree[b] < t->subtree[a];
      if (contb) flags |= UP_CONT_B;
      flags |= UP_PUSH;
      stack.push_back(contb ? a : b);
      depth++;
      max_depth = std::max(max_depth, depth);
      next = contb ? b : a;
    } else if (a >= N) {
      next = a;
    } else if (b >= N) {
      flags |= UP_CONT_B;
      next = b;
    }
    t->up_prog.insert(t->up_prog.end(), {cur, a, b, flags});
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
  t->up_depth = max_depth;
  if ((int32_t)t->up_prog.size() != 4 * (N - 1) || (int32_t)t->up_fifo.size() != M) return 3;
  return 0;
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

int phy_topo_info(const phy_topo* t, int32_t* dd, int32_t* ud, int32_t* parent_out) {
  if (!t) return PHY_EINVAL;
  if (dd) *dd = t->down_depth;
  if (ud) *ud = t->up_depth;
  if (parent_out) std::copy(t->parent.begin(), t->parent.end(), parent_out);
  return PHY_OK;
}

}  // extern "C"
