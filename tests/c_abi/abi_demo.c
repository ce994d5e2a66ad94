/* A plain-C consumer of the drop-in C ABI (include/phylo_hip.h): the call
 * sequence of the OCaml stubs (context, topology, Fitch down- and up-pass,
 * likelihood model, whole-tree likelihood with branch derivatives), with no
 * Python and no torch in the process.  It reads one case from a binary file
 * and writes the results to another; the GPU test compares them with the
 * oracle.
 *
 * Input, little-endian:
 *   int64  N, root, n_sites, n_patterns, n_rates
 *   int32  children[2(N-1)]            (phy_topo_create layout)
 *   uint8  fitch_leaves[N][n_sites]    4-bit state sets
 *   uint32 fitch_weights[n_sites]
 *   double pi[4], Q[16], rates[R], probs[R]
 *   uint8  tips[N][n_patterns]         4-bit state sets
 *   uint32 weights[n_patterns]
 *   double blen[2N-1]
 * Output:
 *   uint64 fitch_total
 *   uint32 site_cost[n_sites]
 *   uint8  prelim[N-1][n_sites]        internal nodes N .. 2N-2
 *   uint8  final[2N-1][n_sites]
 *   double lnl_f64, lnl_f32, lnl_f32_fused
 *   double d1[2N-1], d2[2N-1]          FP64 form
 * Exit status 0 on success; on a library error the message of
 * phy_last_error goes to stderr and the status is 1. */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "phylo_hip.h"

static phy_ctx* g_ctx = NULL;

static void check(int rc, const char* what) {
  if (rc != PHY_OK) {
    fprintf(stderr, "%s failed (%d): %s\n", what, rc, g_ctx ? phy_last_error(g_ctx) : "");
    exit(1);
  }
}

static void* xmalloc(size_t n) {
  void* p = calloc(n ? n : 1, 1);
  if (!p) {
    fprintf(stderr, "out of host memory\n");
    exit(1);
  }
  return p;
}

static void get(FILE* f, void* dst, size_t n) {
  if (n && fread(dst, 1, n, f) != n) {
    fprintf(stderr, "short input\n");
    exit(1);
  }
}

static void put(FILE* f, const void* src, size_t n) {
  if (n && fwrite(src, 1, n, f) != n) {
    fprintf(stderr, "short write\n");
    exit(1);
  }
}

/* rows x n 4-bit sets -> rows x words(n) nibble blocks */
static uint64_t* pack(const uint8_t* s, int64_t rows, int64_t n, int64_t words) {
  uint64_t* w = (uint64_t*)xmalloc((size_t)(rows * words) * sizeof(uint64_t));
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t i = 0; i < n; ++i)
      w[r * words + i / 16] |= (uint64_t)(s[r * n + i] & 15u) << (4 * (i % 16));
  return w;
}

static void unpack(const uint64_t* w, int64_t rows, int64_t n, int64_t words, uint8_t* s) {
  for (int64_t r = 0; r < rows; ++r)
    for (int64_t i = 0; i < n; ++i)
      s[r * n + i] = (uint8_t)((w[r * words + i / 16] >> (4 * (i % 16))) & 15u);
}

static void* dev(size_t bytes) {
  void* p = NULL;
  check(phy_dev_alloc(g_ctx, bytes ? bytes : 8, &p), "phy_dev_alloc");
  return p;
}

static double lnl_form(const phy_topo* topo, const phy_model* model, int32_t form,
                       int64_t P, const uint64_t* d_tips, int64_t stride, const uint32_t* d_w,
                       const double* d_blen, double* d_d1, double* d_d2) {
  size_t wb = 0;
  const int derivs = d_d1 != NULL;
  check(phy_lk_tree_work_size(topo, model, P, form, derivs, &wb), "phy_lk_tree_work_size");
  void* d_work = dev(wb);
  double* d_lnl = (double*)dev(sizeof(double));
  check(phy_lk_tree(g_ctx, topo, model, form, P, d_tips, stride, d_w, d_blen, d_lnl, NULL, d_d1,
                    d_d2, d_work, wb),
        "phy_lk_tree");
  double v = 0.0;
  check(phy_copy_d2h(g_ctx, &v, d_lnl, sizeof(double)), "phy_copy_d2h");
  check(phy_dev_free(g_ctx, d_work), "phy_dev_free");
  check(phy_dev_free(g_ctx, d_lnl), "phy_dev_free");
  return v;
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
  int64_t hdr[5];
  get(in, hdr, sizeof hdr);
  const int64_t N = hdr[0], root = hdr[1], S = hdr[2], P = hdr[3], R = hdr[4];
  if (N < 2 || S < 1 || P < 1 || R < 1 || R > 8) {
    fprintf(stderr, "bad header\n");
    return 2;
  }
  const int64_t M = 2 * N - 1;
  int32_t* children = (int32_t*)xmalloc((size_t)(2 * (N - 1)) * sizeof(int32_t));
  uint8_t* fl = (uint8_t*)xmalloc((size_t)(N * S));
  uint32_t* fw = (uint32_t*)xmalloc((size_t)S * sizeof(uint32_t));
  double pi[4], Q[16], rates[8], probs[8];
  uint8_t* tips = (uint8_t*)xmalloc((size_t)(N * P));
  uint32_t* w = (uint32_t*)xmalloc((size_t)P * sizeof(uint32_t));
  double* blen = (double*)xmalloc((size_t)M * sizeof(double));
  get(in, children, (size_t)(2 * (N - 1)) * sizeof(int32_t));
  get(in, fl, (size_t)(N * S));
  get(in, fw, (size_t)S * sizeof(uint32_t));
  get(in, pi, sizeof pi);
  get(in, Q, sizeof Q);
  get(in, rates, (size_t)R * sizeof(double));
  get(in, probs, (size_t)R * sizeof(double));
  get(in, tips, (size_t)(N * P));
  get(in, w, (size_t)P * sizeof(uint32_t));
  get(in, blen, (size_t)M * sizeof(double));
  fclose(in);

  check(phy_ctx_create(0, NULL, &g_ctx), "phy_ctx_create");
  phy_topo* topo = NULL;
  check(phy_topo_create(g_ctx, (int32_t)N, children, (int32_t)root, &topo), "phy_topo_create");

  /* Fitch: down-pass (prelims, per-site and total cost), then up-pass */
  const int64_t fs = phy_words(S);
  uint64_t* h_leaves = pack(fl, N, S, fs);
  uint64_t* d_sets = (uint64_t*)dev((size_t)(M * fs) * sizeof(uint64_t));
  uint64_t* d_final = (uint64_t*)dev((size_t)(M * fs) * sizeof(uint64_t));
  uint32_t* d_fw = (uint32_t*)dev((size_t)S * sizeof(uint32_t));
  uint32_t* d_sc = (uint32_t*)dev((size_t)S * sizeof(uint32_t));
  uint64_t* d_total = (uint64_t*)dev(sizeof(uint64_t));
  check(phy_copy_h2d(g_ctx, d_sets, h_leaves, (size_t)(N * fs) * sizeof(uint64_t)), "phy_copy_h2d");
  check(phy_copy_h2d(g_ctx, d_fw, fw, (size_t)S * sizeof(uint32_t)), "phy_copy_h2d");
  check(phy_fitch_downpass(g_ctx, topo, d_sets, fs, S, d_fw, d_sc, NULL, d_total),
        "phy_fitch_downpass");
  check(phy_fitch_uppass(g_ctx, topo, d_sets, d_final, fs, S), "phy_fitch_uppass");
  uint64_t total = 0;
  uint32_t* sc = (uint32_t*)xmalloc((size_t)S * sizeof(uint32_t));
  uint64_t* h_sets = (uint64_t*)xmalloc((size_t)(M * fs) * sizeof(uint64_t));
  uint64_t* h_final = (uint64_t*)xmalloc((size_t)(M * fs) * sizeof(uint64_t));
  check(phy_copy_d2h(g_ctx, &total, d_total, sizeof total), "phy_copy_d2h");
  check(phy_copy_d2h(g_ctx, sc, d_sc, (size_t)S * sizeof(uint32_t)), "phy_copy_d2h");
  check(phy_copy_d2h(g_ctx, h_sets, d_sets, (size_t)(M * fs) * sizeof(uint64_t)), "phy_copy_d2h");
  check(phy_copy_d2h(g_ctx, h_final, d_final, (size_t)(M * fs) * sizeof(uint64_t)), "phy_copy_d2h");

  /* likelihood: model, tips, weights, branch lengths; three forms */
  phy_model* model = NULL;
  check(phy_model_create(g_ctx, 4, (int32_t)R, pi, Q, rates, probs, &model), "phy_model_create");
  const int64_t ls = phy_words(P);
  uint64_t* h_tips = pack(tips, N, P, ls);
  uint64_t* d_tips = (uint64_t*)dev((size_t)(N * ls) * sizeof(uint64_t));
  uint32_t* d_w = (uint32_t*)dev((size_t)P * sizeof(uint32_t));
  double* d_blen = (double*)dev((size_t)M * sizeof(double));
  double* d_d1 = (double*)dev((size_t)M * sizeof(double));
  double* d_d2 = (double*)dev((size_t)M * sizeof(double));
  check(phy_copy_h2d(g_ctx, d_tips, h_tips, (size_t)(N * ls) * sizeof(uint64_t)), "phy_copy_h2d");
  check(phy_copy_h2d(g_ctx, d_w, w, (size_t)P * sizeof(uint32_t)), "phy_copy_h2d");
  check(phy_copy_h2d(g_ctx, d_blen, blen, (size_t)M * sizeof(double)), "phy_copy_h2d");
  const double l64 = lnl_form(topo, model, PHY_F64, P, d_tips, ls, d_w, d_blen, d_d1, d_d2);
  double* d1 = (double*)xmalloc((size_t)M * sizeof(double));
  double* d2 = (double*)xmalloc((size_t)M * sizeof(double));
  check(phy_copy_d2h(g_ctx, d1, d_d1, (size_t)M * sizeof(double)), "phy_copy_d2h");
  check(phy_copy_d2h(g_ctx, d2, d_d2, (size_t)M * sizeof(double)), "phy_copy_d2h");
  const double l32 = lnl_form(topo, model, PHY_F32, P, d_tips, ls, d_w, d_blen, NULL, NULL);
  const double l32f =
      lnl_form(topo, model, PHY_F32 | PHY_FUSED, P, d_tips, ls, d_w, d_blen, NULL, NULL);

  FILE* out = fopen(argv[2], "wb");
  if (!out) {
    perror(argv[2]);
    return 2;
  }
  uint8_t* rows = (uint8_t*)xmalloc((size_t)(M * S));
  put(out, &total, sizeof total);
  put(out, sc, (size_t)S * sizeof(uint32_t));
  unpack(h_sets + N * fs, N - 1, S, fs, rows);
  put(out, rows, (size_t)((N - 1) * S));
  unpack(h_final, M, S, fs, rows);
  put(out, rows, (size_t)(M * S));
  const double l3[3] = {l64, l32, l32f};
  put(out, l3, sizeof l3);
  put(out, d1, (size_t)M * sizeof(double));
  put(out, d2, (size_t)M * sizeof(double));
  fclose(out);

  void* bufs[] = {d_sets, d_final, d_fw, d_sc, d_total, d_tips, d_w, d_blen, d_d1, d_d2};
  for (size_t i = 0; i < sizeof bufs / sizeof bufs[0]; ++i) check(phy_dev_free(g_ctx, bufs[i]), "phy_dev_free");
  phy_model_destroy(model);
  phy_topo_destroy(topo);
  phy_ctx_destroy(g_ctx);
  free(children); free(fl); free(fw); free(tips); free(w); free(blen); free(h_leaves);
  free(sc); free(h_sets); free(h_final); free(h_tips); free(d1); free(d2); free(rows);
  printf("ok N=%lld sites=%lld patterns=%lld rates=%lld\n", (long long)N, (long long)S,
         (long long)P, (long long)R);
  return 0;
}
