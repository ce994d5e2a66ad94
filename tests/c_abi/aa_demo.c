/* A plain-C consumer of the batched 20-state pruning of the C ABI
 * (phy_aa_batch_create / _info / _run / _destroy in include/phylo_hip.h):
 * many trees, each with its own GTR-20 model and rate classes, in one
 * persistent launch.  It reads the concatenated inputs from a binary file,
 * evaluates the batch twice (the work queue must reset between runs) and
 * writes the per-tree lnL.
 *
 * Input, little-endian:
 *   int64  T, R, waves
 *   int32  n_leaves[T];  int64 n_patterns[T];  int32 roots[T]
 *   int32  children[sum 2(N_t - 1)]
 *   double blen[sum (2N_t - 1)], pi[T][20], Q[T][400], rates[T][R], probs[T][R]
 *   uint8  tips[sum N_t P_t]       state index 0..19, >= 20 missing
 *   uint32 weights[sum P_t]
 * Output:
 *   int64  pattern_nodes
 *   double lnl_first[T], lnl_second[T]
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
  int64_t* hdr = (int64_t*)xread(in, 3 * sizeof(int64_t));
  const int64_t T = hdr[0], R = hdr[1], waves = hdr[2];
  if (T < 1 || R < 1 || R > 8) {
    fprintf(stderr, "bad header\n");
    return 2;
  }
  int32_t* nl = (int32_t*)xread(in, (size_t)T * sizeof(int32_t));
  int64_t* np = (int64_t*)xread(in, (size_t)T * sizeof(int64_t));
  int32_t* roots = (int32_t*)xread(in, (size_t)T * sizeof(int32_t));
  int64_t n_ch = 0, n_bl = 0, n_tip = 0, n_w = 0;
  for (int64_t t = 0; t < T; ++t) {
    if (nl[t] < 2 || np[t] < 1) {
      fprintf(stderr, "bad tree %lld\n", (long long)t);
      return 2;
    }
    n_ch += 2 * (int64_t)(nl[t] - 1);
    n_bl += 2 * (int64_t)nl[t] - 1;
    n_tip += (int64_t)nl[t] * np[t];
    n_w += np[t];
  }
  int32_t* children = (int32_t*)xread(in, (size_t)n_ch * sizeof(int32_t));
  double* blen = (double*)xread(in, (size_t)n_bl * sizeof(double));
  double* pi = (double*)xread(in, (size_t)T * 20 * sizeof(double));
  double* Q = (double*)xread(in, (size_t)T * 400 * sizeof(double));
  double* rates = (double*)xread(in, (size_t)(T * R) * sizeof(double));
  double* probs = (double*)xread(in, (size_t)(T * R) * sizeof(double));
  uint8_t* tips = (uint8_t*)xread(in, (size_t)n_tip);
  uint32_t* w = (uint32_t*)xread(in, (size_t)n_w * sizeof(uint32_t));
  fclose(in);

  check(phy_ctx_create(0, NULL, &g_ctx), "phy_ctx_create");
  phy_aa_batch* b = NULL;
  check(phy_aa_batch_create(g_ctx, (int32_t)T, nl, np, children, roots, blen, pi, Q, (int32_t)R,
                            rates, probs, tips, w, (int32_t)waves, &b),
        "phy_aa_batch_create");
  int64_t pattern_nodes = 0, n_items = 0;
  int32_t depth = 0;
  check(phy_aa_batch_info(b, &pattern_nodes, &n_items, &depth), "phy_aa_batch_info");
  double* d_lnl = NULL;
  check(phy_dev_alloc(g_ctx, (size_t)T * sizeof(double), (void**)&d_lnl), "phy_dev_alloc");
  double* lnl = (double*)calloc((size_t)(2 * T), sizeof(double));
  if (!lnl) return 1;
  for (int k = 0; k < 2; ++k) {
    check(phy_aa_batch_run(g_ctx, b, d_lnl), "phy_aa_batch_run");
    check(phy_copy_d2h(g_ctx, lnl + k * T, d_lnl, (size_t)T * sizeof(double)), "phy_copy_d2h");
  }
  FILE* out = fopen(argv[2], "wb");
  if (!out) {
    perror(argv[2]);
    return 2;
  }
  if (fwrite(&pattern_nodes, sizeof pattern_nodes, 1, out) != 1 ||
      fwrite(lnl, sizeof(double), (size_t)(2 * T), out) != (size_t)(2 * T)) {
    fprintf(stderr, "short write\n");
    return 1;
  }
  fclose(out);
  check(phy_dev_free(g_ctx, d_lnl), "phy_dev_free");
  phy_aa_batch_destroy(b);
  phy_ctx_destroy(g_ctx);
  free(hdr); free(nl); free(np); free(roots); free(children); free(blen); free(pi); free(Q);
  free(rates); free(probs); free(tips); free(w); free(lnl);
  printf("ok trees=%lld rates=%lld waves=%lld pattern_nodes=%lld depth=%d\n", (long long)T,
         (long long)R, (long long)waves, (long long)pattern_nodes, (int)depth);
  return 0;
}
