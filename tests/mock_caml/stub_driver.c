/* Drives the OCaml stubs (ocaml/phylo_hip_stubs.c) through the mock OCaml
 * runtime exactly as ocaml/nonAdditive_c.ml and ocaml/likelihood_c.ml call
 * them: values are built as those modules' records (same field order), the
 * externals are called with OCaml values, and exceptions come back through
 * the runtime's handler.  Prints "key value ..." lines that the GPU test
 * compares with the oracle.  Test infrastructure only.
 *
 * usage: stub_driver N_fitch S_fitch N_lk S_lk seed */
#include <math.h>
#include <setjmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "caml/alloc.h"
#include "caml/bigarray.h"
#include "caml/memory.h"
#include "caml/mlvalues.h"
#include "mock_caml.h"
#include "phylo_hip.h"
#include "phylo_synth.h"

/* the externals, as the .ml files declare them */
value phylo_hip_CAML_register(value);
value phylo_hip_CAML_topo_create(value, value, value);
value nonadditive_CAML_of_codes(value);
value nonadditive_CAML_weights(value);
value nonadditive_CAML_to_codes(value, value);
value nonadditive_CAML_median2(value, value, value);
value nonadditive_CAML_median3(value, value, value, value);
value nonadditive_CAML_distance1(value, value);
value nonadditive_CAML_distance2(value, value, value);
value nonadditive_CAML_adjust3(value, value, value, value);
value nonadditive_CAML_select(value, value);
value nonadditive_CAML_tree_rows(value, value);
value nonadditive_CAML_downpass(value, value, value, value, value);
value nonadditive_CAML_uppass(value, value, value, value, value);
value nonadditive_CAML_row(value, value, value, value);
value likelihood_CAML_model(value);
value likelihood_CAML_of_tips(value, value, value);
value likelihood_CAML_median2(value, value, value);
value likelihood_CAML_median3(value, value, value);
value likelihood_CAML_root(value);
value likelihood_CAML_edge(value, value, value);
value likelihood_CAML_adjust3(value, value, value);
value likelihood_CAML_derivs(value, value, value, value, value);
value likelihood_CAML_diagonalize_gtr(value, value, value);
value likelihood_CAML_compose_gtr(value, value, value, value);

static void die(const char* what) {
  fprintf(stderr, "driver: %s\n", what);
  exit(2);
}

/* NonAdditive_c.t = {states; weights; n_sites; cost; codes} */
static value na_record(value states, value weights, int64_t n, double cost, value codes) {
  value r = caml_alloc_tuple(5);
  Store_field(r, 0, states);
  Store_field(r, 1, weights);
  Store_field(r, 2, Val_long(n));
  Store_field(r, 3, caml_copy_double(cost));
  Store_field(r, 4, codes);
  return r;
}

/* Likelihood_c.t = {part; exps; blen; n_pat; pitch; weights; f32; lnl;
 * codes; model} */
static value lk_record(value part, value exps, double blen, int64_t n, int64_t pitch,
                       value weights, int f32, value codes, value model) {
  value r = caml_alloc_tuple(10);
  Store_field(r, 0, part);
  Store_field(r, 1, exps);
  Store_field(r, 2, caml_copy_double(blen));
  Store_field(r, 3, Val_long(n));
  Store_field(r, 4, Val_long(pitch));
  Store_field(r, 5, weights);
  Store_field(r, 6, Val_bool(f32));
  Store_field(r, 7, caml_copy_double(NAN));
  Store_field(r, 8, codes);
  Store_field(r, 9, model);
  return r;
}

static int32_t* g_children;
static int32_t g_N;
/* post-order of the rooted layout */
static void post(int32_t v, int32_t* out, int* k) {
  if (v >= g_N) {
    post(g_children[2 * (v - g_N)], out, k);
    post(g_children[2 * (v - g_N) + 1], out, k);
  }
  out[(*k)++] = v;
}

static uint64_t fnv(const value arr) {
  uint64_t h = 1469598103934665603ull;
  mlsize_t i;
  for (i = 0; i < Wosize_val(arr); ++i) {
    h ^= (uint64_t)Long_val(Field(arr, i));
    h *= 1099511628211ull;
  }
  return h;
}

static void fitch(int32_t N, int64_t S, uint64_t seed) {
  int32_t* ch = (int32_t*)malloc(sizeof(int32_t) * 2 * (N - 1));
  double* bl = (double*)malloc(sizeof(double) * (2 * N - 1));
  const int64_t W = phy_words(S);
  uint64_t* packed = (uint64_t*)malloc(sizeof(uint64_t) * N * W);
  int64_t* sets = (int64_t*)malloc(sizeof(int64_t) * S);
  double* w = (double*)malloc(sizeof(double) * S);
  value* leaf = (value*)malloc(sizeof(value) * N);
  value* node = (value*)calloc(2 * N - 1, sizeof(value));
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (2 * N - 1));
  int32_t root = 0;
  int k = 0, i;
  int64_t s;
  double total = 0.0;
  value codes, wv, rows_pair, leaves_arr, topo, rows, final, tot, r;
  if (phy_synth_yule(N, seed, 0.1, ch, &root, bl) || phy_synth_fitch_leaves(N, S, seed, packed, W))
    die("synth");
  for (s = 0; s < S; ++s) {
    w[s] = (double)(1 + (s % 3));
    sets[s] = s;
  }
  codes = mock_int_array(sets, (size_t)S);
  wv = nonadditive_CAML_weights(mock_float_array(w, (size_t)S));
  leaves_arr = caml_alloc(N, 0);
  for (i = 0; i < N; ++i) {
    value a;
    for (s = 0; s < S; ++s) sets[s] = (int64_t)((packed[i * W + s / 16] >> (4 * (s % 16))) & 15);
    a = mock_int_array(sets, (size_t)S);
    Store_field(leaves_arr, i, a);
    leaf[i] = na_record(nonadditive_CAML_of_codes(a), wv, S, 0.0, codes);
    node[i] = leaf[i];
  }
  /* node by node: Node.Make1D.median_2 in post-order */
  g_children = ch;
  g_N = N;
  post(root, order, &k);
  for (i = 0; i < k; ++i) {
    const int32_t v = order[i];
    if (v < N) continue;
    r = nonadditive_CAML_median2(Val_none, node[ch[2 * (v - N)]], node[ch[2 * (v - N) + 1]]);
    node[v] = na_record(Field(r, 0), wv, S, Double_val(Field(r, 1)), codes);
    total += Double_val(Field(r, 1));
  }
  printf("fitch_nodes_total %.1f\n", total);
  printf("fitch_root_hash %llu\n",
         (unsigned long long)fnv(nonadditive_CAML_to_codes(Field(node[root], 0), Val_long(S))));
  /* reuse of the previous block (t option argument) gives the same sets */
  {
    const int32_t v = order[k - 1];
    value again = nonadditive_CAML_median2(mock_some(node[v]), node[ch[2 * (v - N)]],
                                           node[ch[2 * (v - N) + 1]]);
    printf("fitch_prev_cost %.1f\n", Double_val(Field(again, 1)));
  }
  printf("fitch_distance1 %.1f\n",
         Double_val(nonadditive_CAML_distance1(node[ch[2 * (root - N)]], node[ch[2 * (root - N) + 1]])));
  /* whole tree */
  rows_pair = nonadditive_CAML_tree_rows(leaves_arr, Val_long(2 * N - 1));
  rows = Field(rows_pair, 0);
  final = Field(nonadditive_CAML_tree_rows(leaves_arr, Val_long(2 * N - 1)), 0);
  {
    value chv;
    int64_t* c64 = (int64_t*)malloc(sizeof(int64_t) * 2 * (N - 1));
    for (i = 0; i < 2 * (N - 1); ++i) c64[i] = ch[i];
    chv = mock_int_array(c64, 2 * (size_t)(N - 1));
    free(c64);
    topo = phylo_hip_CAML_topo_create(Val_long(N), chv, Val_long(root));
  }
  tot = nonadditive_CAML_downpass(topo, rows, Field(rows_pair, 1), wv, Val_long(S));
  nonadditive_CAML_uppass(topo, rows, final, Field(rows_pair, 1), Val_long(S));
  printf("fitch_tree_total %.1f\n", Double_val(Field(tot, 0)));
  printf("fitch_site_cost_hash %llu\n", (unsigned long long)fnv(Field(tot, 1)));
  {
    uint64_t hp = 0, hf = 0;
    for (i = N; i < 2 * N - 1; ++i) {
      hp = hp * 31 + fnv(nonadditive_CAML_row(rows, Field(rows_pair, 1), Val_long(i), Val_long(S)));
      hf = hf * 31 + fnv(nonadditive_CAML_row(final, Field(rows_pair, 1), Val_long(i), Val_long(S)));
    }
    printf("fitch_prelim_hash %llu\nfitch_final_hash %llu\n", (unsigned long long)hp,
           (unsigned long long)hf);
  }
  /* filter_codes keeps the even sites: distance over them */
  {
    int64_t* idx = (int64_t*)malloc(sizeof(int64_t) * S);
    int64_t n = 0;
    value sel, a, b;
    for (s = 0; s < S; s += 2) idx[n++] = s;
    sel = nonadditive_CAML_select(leaf[0], mock_int_array(idx, (size_t)n));
    a = na_record(Field(sel, 0), Field(sel, 1), n, 0.0, codes);
    sel = nonadditive_CAML_select(leaf[1], mock_int_array(idx, (size_t)n));
    b = na_record(Field(sel, 0), Field(sel, 1), n, 0.0, codes);
    printf("fitch_select_distance1 %.1f\n", Double_val(nonadditive_CAML_distance1(a, b)));
    free(idx);
  }
  /* an exception: mismatched site counts raise Invalid_argument */
  {
    jmp_buf jb;
    value small = na_record(Field(leaf[0], 0), wv, S - 1, 0.0, codes);
    mock_caml_handler = &jb;
    if (setjmp(jb) == 0) {
      nonadditive_CAML_median2(Val_none, small, leaf[1]);
      printf("fitch_exception none\n");
    } else {
      printf("fitch_exception %d\n", mock_caml_exn_kind);
    }
    mock_caml_handler = NULL;
  }
  free(ch);
  free(bl);
  free(packed);
  free(sets);
  free(w);
  free(leaf);
  free(node);
  free(order);
}

static void likelihood(int32_t N, int64_t S, uint64_t seed) {
  int32_t* ch = (int32_t*)malloc(sizeof(int32_t) * 2 * (N - 1));
  double* bl = (double*)malloc(sizeof(double) * (2 * N - 1));
  uint8_t* st = (uint8_t*)malloc((size_t)N * S);
  uint8_t* pats = (uint8_t*)malloc((size_t)N * S);
  uint32_t* wts = (uint32_t*)malloc(sizeof(uint32_t) * S);
  double pi[4], co[5], Q[16], rates[4], probs[4] = {0.25, 0.25, 0.25, 0.25};
  int64_t P = 0, p;
  int32_t root = 0;
  int i, k = 0;
  value model, mlm, dm, wv, codes, *node, U, D, Ui, P3;
  int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (2 * N - 1));
  int64_t* tmp;
  double* wd;
  if (phy_synth_yule(N, seed, 0.1, ch, &root, bl) || phy_synth_gtr_params(4, 5.0, seed, pi, co) ||
      phy_gtr_q(4, pi, co, Q) || phy_gamma_rates(PHY_GAMMA_MEAN, 0.5, 0.5, 4, rates))
    die("model");
  if (phy_synth_simulate(N, ch, root, bl, 4, pi, Q, 4, rates, S, seed, 0, st) ||
      phy_synth_compress(N, S, st, pats, wts, &P))
    die("simulate");
  printf("lk_patterns %lld\n", (long long)P);
  /* MlModel.t = {spec; priors; pinvar; rates; probs; q; u; d; ui} */
  mlm = caml_alloc_tuple(9);
  Store_field(mlm, 0, Val_unit);
  Store_field(mlm, 1, mock_vector(pi, 4));
  Store_field(mlm, 2, Val_none);
  Store_field(mlm, 3, mock_vector(rates, 4));
  Store_field(mlm, 4, mock_vector(probs, 4));
  Store_field(mlm, 5, mock_matrix(Q, 4));
  Store_field(mlm, 6, mock_matrix(Q, 4));
  Store_field(mlm, 7, mock_matrix(NULL, 4));
  Store_field(mlm, 8, Val_none);
  model = mlm;
  dm = likelihood_CAML_model(model);
  tmp = (int64_t*)malloc(sizeof(int64_t) * P);
  wd = (double*)malloc(sizeof(double) * P);
  for (p = 0; p < P; ++p) {
    tmp[p] = p;
    wd[p] = (double)wts[p];
  }
  codes = mock_int_array(tmp, (size_t)P);
  wv = nonadditive_CAML_weights(mock_float_array(wd, (size_t)P));
  node = (value*)calloc(2 * N - 1, sizeof(value));
  for (i = 0; i < N; ++i) {
    value r;
    for (p = 0; p < P; ++p) tmp[p] = 1 << pats[(size_t)i * S + p];
    r = likelihood_CAML_of_tips(dm, mock_int_array(tmp, (size_t)P), Val_false);
    node[i] = lk_record(Field(r, 0), Field(r, 1), bl[i], P, Long_val(Field(r, 2)), wv, 0, codes, dm);
  }
  g_children = ch;
  g_N = N;
  post(root, order, &k);
  for (i = 0; i < k; ++i) {
    const int32_t v = order[i];
    value r;
    if (v < N) continue;
    r = likelihood_CAML_median2(Val_none, node[ch[2 * (v - N)]], node[ch[2 * (v - N) + 1]]);
    node[v] = lk_record(Field(r, 0), Field(r, 1), bl[v], P, Long_val(Field(node[0], 4)), wv, 0,
                        codes, dm);
  }
  printf("lk_nodes_lnl %.17g\n", Double_val(likelihood_CAML_root(node[root])));
  {
    const int32_t a = ch[2 * (root - N)], b = ch[2 * (root - N) + 1];
    value e = likelihood_CAML_edge(node[a], node[b], caml_copy_double(bl[a] + bl[b]));
    value m3, adj;
    printf("lk_edge_root %.17g %.17g %.17g\n", Double_field(e, 0), Double_field(e, 1),
           Double_field(e, 2));
    if (a >= N) {
      /* the three neighbours of a: its children and b (across the root edge) */
      value bb = lk_record(Field(node[b], 0), Field(node[b], 1), bl[a] + bl[b], P,
                           Long_val(Field(node[0], 4)), wv, 0, codes, dm);
      m3 = likelihood_CAML_median3(node[ch[2 * (a - N)]], node[ch[2 * (a - N) + 1]], bb);
      printf("lk_median3 %.17g\n", Double_val(Field(m3, 2)));
    } else {
      printf("lk_median3 nan\n");
    }
    adj = likelihood_CAML_adjust3(node[a], node[b], caml_copy_double(bl[a] + bl[b]));
    printf("lk_adjust3 %.17g %.17g\n", Double_field(adj, 0), Double_field(adj, 1));
  }
  /* whole tree with derivatives */
  {
    value tips_arr = caml_alloc(N, 0), rows_pair, data, topo, res, chv, blv;
    int64_t* c64 = (int64_t*)malloc(sizeof(int64_t) * 2 * (N - 1));
    for (i = 0; i < N; ++i) {
      for (p = 0; p < P; ++p) tmp[p] = 1 << pats[(size_t)i * S + p];
      Store_field(tips_arr, i, mock_int_array(tmp, (size_t)P));
    }
    rows_pair = nonadditive_CAML_tree_rows(tips_arr, Val_long(N));
    /* Likelihood_c.tree_data = {tips; stride; n_patterns; tree_weights} */
    data = caml_alloc_tuple(4);
    Store_field(data, 0, Field(rows_pair, 0));
    Store_field(data, 1, Field(rows_pair, 1));
    Store_field(data, 2, Val_long(P));
    Store_field(data, 3, wv);
    for (i = 0; i < 2 * (N - 1); ++i) c64[i] = ch[i];
    chv = mock_int_array(c64, 2 * (size_t)(N - 1));
    topo = phylo_hip_CAML_topo_create(Val_long(N), chv, Val_long(root));
    blv = mock_float_array(bl, (size_t)(2 * N - 1));
    res = likelihood_CAML_derivs(topo, dm, data, blv, Val_false);
    printf("lk_tree_lnl %.17g\n", Double_val(Field(res, 0)));
    printf("lk_tree_d1");
    for (i = 0; i < 2 * N - 1; ++i) printf(" %.17g", Double_field(Field(res, 1), i));
    printf("\nlk_tree_d2");
    for (i = 0; i < 2 * N - 1; ++i) printf(" %.17g", Double_field(Field(res, 2), i));
    printf("\n");
    res = likelihood_CAML_derivs(topo, dm, data, blv, Val_true);
    printf("lk_tree_lnl_f32 %.17g\n", Double_val(Field(res, 0)));
    free(c64);
  }
  /* MlModel stubs: diagonalize Q, compose P(0.3) */
  U = mock_matrix(Q, 4);
  D = mock_matrix(NULL, 4);
  Ui = mock_matrix(NULL, 4);
  likelihood_CAML_diagonalize_gtr(U, D, Ui);
  P3 = likelihood_CAML_compose_gtr(U, D, Ui, caml_copy_double(0.3));
  printf("compose_0.3");
  for (i = 0; i < 16; ++i) printf(" %.17g", ((double*)Caml_ba_data_val(P3))[i]);
  printf("\n");
  printf("blocking_depth %d\n", mock_caml_blocking());
  free(ch);
  free(bl);
  free(st);
  free(pats);
  free(wts);
  free(order);
  free(tmp);
  free(wd);
  free(node);
}

int main(int argc, char** argv) {
  if (argc != 6) die("usage: stub_driver N_fitch S_fitch N_lk S_lk seed");
  phylo_hip_CAML_register(Val_unit);
  fitch(atoi(argv[1]), atoll(argv[2]), (uint64_t)atoll(argv[5]));
  likelihood(atoi(argv[3]), atoll(argv[4]), (uint64_t)atoll(argv[5]));
  mock_finalize_all();
  printf("done\n");
  fflush(stdout);
  return 0;
}
