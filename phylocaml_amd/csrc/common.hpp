// Shared definitions of the gfx950 library: context, topology programs,
// error reporting.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/phylo_hip.h"

struct phy_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  // per-kernel timing (phy_ctx_kernel_timer): HIP event pairs recorded on the
  // stream around the dominant launches of each entry point, resolved into
  // per-label totals by phy_ctx_kernel_time
  bool timing = false;
  struct Span {
    const char* label;
    hipEvent_t start, stop;
  };
  std::vector<Span> spans;
  std::vector<std::pair<std::string, std::pair<double, int64_t>>> totals;
  // this device's copy of the fixed scratch array of phy_lk_tree_noalloc
  // (resolved at creation; null if that lookup failed)
  void* noalloc_scratch = nullptr;
};

// Records a HIP event pair around the launches in its scope when the context's
// timing is on (one span per scope, labelled by the kernel it covers).
struct KernelSpan {
  phy_ctx* c;
  const char* label;
  hipEvent_t a = nullptr, b = nullptr;
  KernelSpan(phy_ctx* ctx, const char* lbl) : c(ctx), label(lbl) {
    if (!c || !c->timing) return;
    if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess ||
        hipEventRecord(a, c->stream) != hipSuccess) {
      if (a) (void)hipEventDestroy(a);
      if (b) (void)hipEventDestroy(b);
      a = b = nullptr;
      (void)hipGetLastError();
    }
  }
  ~KernelSpan() {
    if (!a) return;
    if (hipEventRecord(b, c->stream) == hipSuccess) {
      c->spans.push_back({label, a, b});
    } else {
      (void)hipEventDestroy(a);
      (void)hipEventDestroy(b);
      (void)hipGetLastError();
    }
  }
};

// Resolve (once per device, with that device current) the device address of
// the fixed scratch array behind phy_lk_tree_noalloc; *out receives it.
int phy_noalloc_scratch_init(int device, void** out);

// Operand sources of a program entry (SRC_EXT: pre-order entry of a part,
// whose parent's final set is read from memory).
enum : int { SRC_LEAF = 0, SRC_ACC = 1, SRC_STACK = 2, SRC_EXT = 3 };
// Flag bits of a down entry (int4.w): A source in bits 0-1, B source in bits
// 2-3, bit 4 = push the accumulator to the LDS stack before the entry.
enum : int { DOWN_SPILL = 16 };
// Flag bits of an up entry (int4.w): input source in bits 0-1 (SRC_LEAF means
// "root: read P from the fifo"), bit 2 = child a is a leaf, bit 3 = child b
// is a leaf, bit 4 = child b (not a) continues in the accumulator, bit 5 =
// push the other internal child, bits 6/7 = child a/b is a cut node (its
// final set is computed by the program of its own part).
enum : int { UP_A_LEAF = 4, UP_B_LEAF = 8, UP_CONT_B = 16, UP_PUSH = 32, UP_A_CUT = 64, UP_B_CUT = 128 };

// Programs of a tree split at cut nodes into parts that run side by side
// (one launch, grid.y = part) plus a top program with the cut nodes as
// leaves.  Down: parts 0..P-1, then the top at index P.  Up: the top at
// index 0, then the parts.  Programs and padded fifos are concatenated.
constexpr int kMaxParts = 32;  // the default is 16 (PHY_FITCH_PARTS at topology creation)
struct SplitProg {
  std::vector<int32_t> prog, fifo, prog_off, n_ops, fifo_off;
  int32_t depth = 0;
  int4* d_prog = nullptr;
  int32_t* d_fifo = nullptr;
  int32_t* d_meta = nullptr;  // prog_off | n_ops | fifo_off, parts + 1 each
};

struct phy_topo {
  phy_ctx* ctx = nullptr;
  int32_t n_leaves = 0;
  int32_t root = -1;
  std::vector<int32_t> children;   // 2(N-1)
  std::vector<int32_t> parent;     // 2N-1
  std::vector<int32_t> subtree;    // leaves below each node
  // post-order program: {v, A, B, flags}; A/B are child node ids
  std::vector<int32_t> down_prog;  // 4 per entry, N-1 entries
  std::vector<int32_t> down_fifo;  // leaf ids in consumption order (N)
  int32_t down_depth = 0;          // LDS stack entries needed
  // pre-order program: {v, a, b, flags}
  std::vector<int32_t> up_prog;
  std::vector<int32_t> up_fifo;    // node ids in load order (2N-1)
  int32_t up_depth = 0;
  // device copies
  int4* d_down_prog = nullptr;
  int32_t* d_down_fifo = nullptr;
  int4* d_up_prog = nullptr;
  int32_t* d_up_fifo = nullptr;
  // split programs (empty parts: the tree is run whole)
  std::vector<int32_t> parts;      // cut node of each part
  SplitProg split_down, split_up;
  int32_t* d_part_parent = nullptr;
  // split form of the fused pruning pass (its own cuts, fewer than the Fitch
  // passes take: every part adds a stored root partial per pattern):
  // psplit holds the part programs; the top program reads the cut nodes as
  // stored operands (SRC_EXT; the part of A in bits 8-15 and that of B in
  // bits 16-23 of the entry's flags), its fifo the real leaves only (padded)
  std::vector<int32_t> pparts;
  SplitProg psplit;
  std::vector<int32_t> ptop_prog, ptop_fifo;
  int32_t ptop_depth = 0;
  int4* d_ptop_prog = nullptr;
  int32_t* d_ptop_fifo = nullptr;
  // per-part partial site costs of the split down-pass, grown on demand
  mutable uint32_t* d_scratch = nullptr;
  mutable size_t scratch_bytes = 0;
};

inline int set_err(phy_ctx* ctx, int code, const std::string& msg) {
  if (ctx) ctx->err = msg;
  return code;
}

#define PHY_HIP(ctx, call)                                                           \
  do {                                                                               \
    hipError_t e_ = (call);                                                          \
    if (e_ != hipSuccess)                                                            \
      return set_err((ctx), PHY_EDEVICE, std::string(#call ": ") + hipGetErrorString(e_)); \
  } while (0)

#define PHY_LAUNCHED(ctx)                                                            \
  do {                                                                               \
    hipError_t e_ = hipGetLastError();                                               \
    if (e_ != hipSuccess)                                                            \
      return set_err((ctx), PHY_EDEVICE, std::string("kernel launch: ") + hipGetErrorString(e_)); \
  } while (0)

// Compile the post-order / pre-order programs of t (children, root, n_leaves
// set); returns 0 on success.  Host only.
int phy_compile_topo(phy_topo* t);

// Pre-order program ({v, a, b, flags} per internal node, root first with
// source SRC_LEAF) of the tree under `top` in the children layout ch of a
// tree with N leaves; nodes not below `top` are ignored, so a tree with
// nodes removed (a pruned subtree) compiles on the original node ids.  size
// orders the walk (the smaller subtree continues in the accumulator); depth
// receives the stack depth.  With `enter`, the walk enters an internal child
// only where enter[child] is set; the others are flagged UP_A_CUT /
// UP_B_CUT (their partials are still read, their subtrees not walked).
// Host only.
void phy_compile_up_program(int32_t N, const std::vector<int32_t>& ch, const std::vector<int32_t>& size,
                            int32_t top, std::vector<int32_t>* prog, int32_t* depth,
                            const std::vector<char>* enter = nullptr);

// Device copies of leaf / node id streams are padded past their end with a
// valid id (the stream's first), so kernels prefetch whole batches of ids and
// rows with no bounds check: two batches of at most 64 ids may be read past
// the last one consumed.  The padded length is a multiple of 64, which keeps
// every batch start 16-byte aligned for vector id loads.
constexpr size_t kFifoSlack = 128;
__host__ __device__ constexpr size_t phy_fifo_padded(size_t n) { return ((n + kFifoSlack + 63) / 64) * 64; }

// Zero n uint64 values on the stream with a kernel (graph-capture safe).
int phy_zero_u64(phy_ctx* ctx, uint64_t* d, int64_t n);
