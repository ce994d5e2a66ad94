/* A plain-C consumer of the whole-tree Sankoff pass of the C ABI
 * (phy_sankoff_downpass in include/phylo_hip.h): context, topology, an
 * integer cost matrix and byte state sets in device memory, one launch, the
 * per-site costs and the weighted total back.
 *
 * Input, little-endian:
 *   int64  N, root, n_sites, K
 *   int32  children[2(N-1)];  int32 tcm[K][K]
 *   uint8  leaves[N][n_sites]  (bit k = state k; an empty set means every state)
 *   uint32 weights[n_sites]
 * Output:
 *   uint64 total;  uint32 site_cost[n_sites]
 * Exit status 0 on success; on a library error the message of
 * phy_last_error goes to stderr and the status is 1. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "phylo_hip.h"

static phy_ctx* g_ctx = NULL;

static void check(int rc, const char* what) {
  if (rc != PHY_OK) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, g_ctx ? phy_last_error(g_ctx) : "");
    exit(1);
  }
}

static void* xread(FILE* f, size_t n) {
  void* p = calloc(n ? n : 1, 1);
  if (!p || (n && fread(p, 1, n, f) != n)) {
    fprintf(stderr, "short input\n");
    exit(1);
  }
  return p;
}

static void* dev(size_t n) {
  void* p = NULL;
  check(phy_dev_alloc(g_ctx, n ? n : 8, &p), "phy_dev_alloc");
  return p;
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s input.bin output.bin\n", argv[0]);
    return 2;
  }
  FILE* in = fopen(argv[1], "rb");
  if (!in) {
    perror(argv[1]);
    return 2;
  }
  int64_t* hdr = (int64_t*)xread(in, 4 * sizeof(int64_t));
  const int64_t N = hdr[0], root = hdr[1], S = hdr[2], K = hdr[3];
  if (N < 2 || S < 1 || K < 2 || K > 8) {
    fprintf(stderr, "bad header\n");
    return 2;
  }
  int32_t* children = (int32_t*)xread(in, (size_t)(2 * (N - 1)) * sizeof(int32_t));
  int32_t* tcm = (int32_t*)xread(in, (size_t)(K * K) * sizeof(int32_t));
  uint8_t* leaves = (uint8_t*)xread(in, (size_t)(N * S));
  uint32_t* w = (uint32_t*)xread(in, (size_t)S * sizeof(uint32_t));
  fclose(in);

  check(phy_ctx_create(0, NULL, &g_ctx), "phy_ctx_create");
  phy_topo* topo = NULL;
  check(phy_topo_create(g_ctx, (int32_t)N, children, (int32_t)root, &topo), "phy_topo_create");
  int32_t* d_tcm = (int32_t*)dev((size_t)(K * K) * sizeof(int32_t));
  uint8_t* d_leaves = (uint8_t*)dev((size_t)(N * S));
  uint32_t* d_w = (uint32_t*)dev((size_t)S * sizeof(uint32_t));
  uint32_t* d_sc = (uint32_t*)dev((size_t)S * sizeof(uint32_t));
  uint64_t* d_total = (uint64_t*)dev(sizeof(uint64_t));
  check(phy_copy_h2d(g_ctx, d_tcm, tcm, (size_t)(K * K) * sizeof(int32_t)), "phy_copy_h2d");
  check(phy_copy_h2d(g_ctx, d_leaves, leaves, (size_t)(N * S)), "phy_copy_h2d");
  check(phy_copy_h2d(g_ctx, d_w, w, (size_t)S * sizeof(uint32_t)), "phy_copy_h2d");
  check(phy_sankoff_downpass(g_ctx, topo, (int32_t)K, d_tcm, d_leaves, S, S, d_w, d_sc, d_total),
        "phy_sankoff_downpass");
  uint64_t total = 0;
  uint32_t* sc = (uint32_t*)calloc((size_t)S, sizeof(uint32_t));
  if (!sc) return 1;
  check(phy_copy_d2h(g_ctx, &total, d_total, sizeof total), "phy_copy_d2h");
  check(phy_copy_d2h(g_ctx, sc, d_sc, (size_t)S * sizeof(uint32_t)), "phy_copy_d2h");

  FILE* out = fopen(argv[2], "wb");
  if (!out) {
    perror(argv[2]);
    return 2;
  }
  if (fwrite(&total, sizeof total, 1, out) != 1 || fwrite(sc, sizeof(uint32_t), (size_t)S, out) != (size_t)S) {
    fprintf(stderr, "short write\n");
    return 1;
  }
  fclose(out);
  void* bufs[] = {d_tcm, d_leaves, d_w, d_sc, d_total};
  for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; ++i) check(phy_dev_free(g_ctx, bufs[i]), "phy_dev_free");
  phy_topo_destroy(topo);
  phy_ctx_destroy(g_ctx);
  free(hdr); free(children); free(tcm); free(leaves); free(w); free(sc);
  printf("ok N=%lld sites=%lld states=%lld total=%llu\n", (long long)N, (long long)S, (long long)K,
         (unsigned long long)total);
  return 0;
}
