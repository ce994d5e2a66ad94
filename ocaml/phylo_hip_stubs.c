/*
 * phylo_hip_stubs.c -- OCaml C stubs binding libphylo_hip (include/phylo_hip.h)
 * as the reference binds its own C: `external ... = "<module>_CAML_<fn>"`
 * (lib/bitvector.ml:72-100, lib/mlModel.ml:73-90), CAMLparam / CAMLlocal /
 * CAMLreturn (lib/bitvector/bv.c:463-480), device memory owned by custom
 * blocks with a finalizer and GC pressure (lib/bitvector/bv.c:22,181-256),
 * and a non-zero status turned into Failure (lib/mlmodel.c:110-120).
 *
 * Modules served (ocaml/nonAdditive_c.ml, ocaml/likelihood_c.ml):
 *   NonAdditive_c  Fitch values {states; weights; n_sites; cost; codes}
 *   Likelihood_c   pruning values {part; exps; blen; n_pat; pitch; weights;
 *                  f32; lnl; codes; model}
 *   MlModel        likelihood_CAML_diagonalize_gtr/_sym, compose_gtr/_sym:
 *                  the names MlModel's externals bind (lib/mlModel.ml:73-90),
 *                  served without LAPACK.
 * Every stub takes at most five arguments (no bytecode/native pair needed).
 * One process-wide context on device 0; the stubs never touch the OCaml heap
 * between caml_enter_blocking_section and caml_leave_blocking_section.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <caml/alloc.h>
#include <caml/bigarray.h>
#include <caml/custom.h>
#include <caml/fail.h>
#include <caml/memory.h>
#include <caml/mlvalues.h>
#include <caml/signals.h>

#include "phylo_hip.h"

/* ------------------------------------------------------------------------- */
/* Context and errors                                                        */
/* ------------------------------------------------------------------------- */
static phy_ctx* ctx = NULL;

static phy_ctx* get_ctx(void) {
  if (!ctx && phy_ctx_create(0, NULL, &ctx) != PHY_OK) {
    ctx = NULL;
    caml_failwith("phylo_hip: no usable gfx950 device");
  }
  return ctx;
}

/* a non-zero status becomes Failure with the library's message */
static void check(int rc) {
  if (rc != PHY_OK) caml_failwith(phy_last_error(ctx));
}

/* synchronous device-to-host copy with the runtime lock released (no OCaml
 * value is read or written while it is released) */
static void copy_d2h(void* h, const void* d, size_t bytes) {
  int rc;
  caml_enter_blocking_section();
  rc = phy_copy_d2h(ctx, h, d, bytes);
  caml_leave_blocking_section();
  check(rc);
}

/* ------------------------------------------------------------------------- */
/* Custom blocks: device buffers, topologies, likelihood models              */
/* ------------------------------------------------------------------------- */
typedef struct {
  void* d;
  size_t bytes;
} dbuf;
#define Dbuf_val(v) ((dbuf*)Data_custom_val(v))
#define Dptr(v) (Dbuf_val(v)->d)

static int dbuf_freq = 10000; /* device buffers allocated before a major slice */

static void dbuf_finalize(value v) {
  dbuf* b = Dbuf_val(v);
  if (b->d && ctx) (void)phy_dev_free(ctx, b->d);
  b->d = NULL;
}
static struct custom_operations dbuf_ops = {
    .identifier = "phylo_hip/dbuf/1",
    .finalize = dbuf_finalize,
    .compare = custom_compare_default,
    .hash = custom_hash_default,
    .serialize = custom_serialize_default,
    .deserialize = custom_deserialize_default,
    .compare_ext = custom_compare_ext_default};

static value dbuf_alloc(size_t bytes) {
  CAMLparam0();
  CAMLlocal1(v);
  void* d = NULL;
  phy_ctx* c = get_ctx();
  v = caml_alloc_custom(&dbuf_ops, sizeof(dbuf), 1, dbuf_freq);
  Dbuf_val(v)->d = NULL;
  Dbuf_val(v)->bytes = 0;
  check(phy_dev_alloc(c, bytes > 0 ? bytes : 8, &d));
  Dbuf_val(v)->d = d;
  Dbuf_val(v)->bytes = bytes;
  CAMLreturn(v);
}

#define Topo_val(v) (*(phy_topo**)Data_custom_val(v))
static void topo_finalize(value v) {
  if (Topo_val(v)) phy_topo_destroy(Topo_val(v));
  Topo_val(v) = NULL;
}
static struct custom_operations topo_ops = {
    .identifier = "phylo_hip/topo/1",
    .finalize = topo_finalize,
    .compare = custom_compare_default,
    .hash = custom_hash_default,
    .serialize = custom_serialize_default,
    .deserialize = custom_deserialize_default,
    .compare_ext = custom_compare_ext_default};

typedef struct {
  phy_model* h;
  int32_t K, R;
} dmodel;
#define Dmodel_val(v) ((dmodel*)Data_custom_val(v))
#define Model_val(v) (Dmodel_val(v)->h)
static void model_finalize(value v) {
  if (Model_val(v)) phy_model_destroy(Model_val(v));
  Model_val(v) = NULL;
}
static struct custom_operations model_ops = {
    .identifier = "phylo_hip/model/1",
    .finalize = model_finalize,
    .compare = custom_compare_default,
    .hash = custom_hash_default,
    .serialize = custom_serialize_default,
    .deserialize = custom_deserialize_default,
    .compare_ext = custom_compare_ext_default};

value phylo_hip_CAML_register(value unit) {
  CAMLparam1(unit);
  caml_register_custom_operations(&dbuf_ops);
  caml_register_custom_operations(&topo_ops);
  caml_register_custom_operations(&model_ops);
  CAMLreturn(Val_unit);
}

value phylo_hip_CAML_custom_max(value n) {
  CAMLparam1(n);
  dbuf_freq = Int_val(n);
  CAMLreturn(Val_unit);
}

/* Tree.t reduced to the rooted layout of phy_topo_create: children of
 * internal node N + k at 2k, 2k + 1. */
value phylo_hip_CAML_topo_create(value n_leaves, value children, value root) {
  CAMLparam3(n_leaves, children, root);
  CAMLlocal1(v);
  const int32_t N = Int_val(n_leaves);
  const mlsize_t n = Wosize_val(children);
  int32_t* ch;
  phy_topo* t = NULL;
  mlsize_t i;
  int rc;
  if (N < 2 || n != (mlsize_t)(2 * (N - 1))) caml_invalid_argument("topo_create: children size");
  ch = (int32_t*)malloc(sizeof(int32_t) * n);
  if (!ch) caml_failwith("topo_create: out of memory");
  for (i = 0; i < n; ++i) ch[i] = (int32_t)Int_val(Field(children, i));
  rc = phy_topo_create(get_ctx(), N, ch, Int_val(root), &t);
  free(ch);
  check(rc);
  v = caml_alloc_custom(&topo_ops, sizeof(phy_topo*), 0, 1);
  Topo_val(v) = t;
  CAMLreturn(v);
}

/* ------------------------------------------------------------------------- */
/* Host helpers                                                              */
/* ------------------------------------------------------------------------- */
/* pack one 4-bit set per site (an int array) into nibble blocks */
static void pack_sets(value sets, uint64_t* words, int64_t n_words) {
  const mlsize_t n = Wosize_val(sets);
  mlsize_t s;
  memset(words, 0, sizeof(uint64_t) * (size_t)n_words);
  for (s = 0; s < n; ++s) {
    const uint64_t x = (uint64_t)(Long_val(Field(sets, s)) & 15);
    words[s / 16] |= x << (4 * (s % 16));
  }
}

static value pack_upload(value sets) {
  CAMLparam1(sets);
  CAMLlocal1(res);
  const int64_t n_words = phy_words((int64_t)Wosize_val(sets));
  uint64_t* h = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(n_words > 0 ? n_words : 1));
  int rc;
  if (!h) caml_failwith("phylo_hip: out of memory");
  pack_sets(sets, h, n_words);
  res = dbuf_alloc(sizeof(uint64_t) * (size_t)n_words);
  rc = phy_copy_h2d(ctx, Dptr(res), h, sizeof(uint64_t) * (size_t)n_words);
  free(h);
  check(rc);
  CAMLreturn(res);
}

static value pair(value a, value b) {
  CAMLparam2(a, b);
  CAMLlocal1(t);
  t = caml_alloc_tuple(2);
  Store_field(t, 0, a);
  Store_field(t, 1, b);
  CAMLreturn(t);
}

/* ------------------------------------------------------------------------- */
/* NonAdditive_c (Fitch): t = {states; weights; n_sites; cost; codes}        */
/* ------------------------------------------------------------------------- */
enum { NA_STATES = 0, NA_WEIGHTS = 1, NA_NSITES = 2, NA_COST = 3, NA_CODES = 4 };
#define Na_states(v) ((uint64_t*)Dptr(Field(v, NA_STATES)))
#define Na_weights(v) ((const uint32_t*)Dptr(Field(v, NA_WEIGHTS)))
#define Na_nsites(v) ((int64_t)Long_val(Field(v, NA_NSITES)))

/* the character's state sets (4-bit, one per site) -> device nibble block */
value nonadditive_CAML_of_codes(value sets) {
  CAMLparam1(sets);
  CAMLreturn(pack_upload(sets));
}

/* float weights (non-negative integers, the weights of lib/nonAdditive_c.ml:
 * 3) -> device uint32 */
value nonadditive_CAML_weights(value w) {
  CAMLparam1(w);
  CAMLlocal1(res);
  const mlsize_t n = Wosize_val(w) / Double_wosize; /* a float array */
  uint32_t* h = (uint32_t*)malloc(sizeof(uint32_t) * (n > 0 ? n : 1));
  mlsize_t i;
  int rc;
  if (!h) caml_failwith("weights: out of memory");
  for (i = 0; i < n; ++i) {
    const double x = Double_field(w, i);
    if (!(x >= 0.0 && x <= 4294967295.0 && x == floor(x))) {
      free(h);
      caml_invalid_argument("weights: not a non-negative 32-bit integer");
    }
    h[i] = (uint32_t)x;
  }
  res = dbuf_alloc(sizeof(uint32_t) * n);
  rc = phy_copy_h2d(ctx, Dptr(res), h, sizeof(uint32_t) * n);
  free(h);
  check(rc);
  CAMLreturn(res);
}

/* the state sets of a block back on the host: int array of n_sites */
value nonadditive_CAML_to_codes(value states, value n_sites) {
  CAMLparam2(states, n_sites);
  CAMLlocal1(res);
  const int64_t n = Long_val(n_sites), nw = phy_words(n);
  uint64_t* h = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(nw > 0 ? nw : 1));
  int64_t s;
  if (!h) caml_failwith("to_codes: out of memory");
  if ((size_t)nw * sizeof(uint64_t) > Dbuf_val(states)->bytes) {
    free(h);
    caml_invalid_argument("to_codes: block smaller than n_sites");
  }
  caml_enter_blocking_section();
  s = phy_copy_d2h(ctx, h, Dptr(states), sizeof(uint64_t) * (size_t)nw);
  caml_leave_blocking_section();
  if (s != PHY_OK) {
    free(h);
    check((int)s);
  }
  res = caml_alloc((mlsize_t)n, 0);
  for (s = 0; s < n; ++s) Store_field(res, s, Val_long((h[s / 16] >> (4 * (s % 16))) & 15));
  free(h);
  CAMLreturn(res);
}

/* a device cost word read back as a float */
static double read_cost(uint64_t* d_cost) {
  uint64_t c = 0;
  copy_d2h(&c, d_cost, sizeof c);
  return (double)c;
}

/* out block: the previous value's block when given (Node.Make1D passes the
 * node's previous value, lib/node.ml:171,188), else a new one */
static value na_out(value prev, value like) {
  CAMLparam2(prev, like);
  if (Is_block(prev)) CAMLreturn(Field(Some_val(prev), NA_STATES));
  CAMLreturn(dbuf_alloc(sizeof(uint64_t) * (size_t)phy_words(Na_nsites(like))));
}

static value na_cost_buf(void) { return dbuf_alloc(sizeof(uint64_t)); }

/* median_2 (lib/nonAdditive_c.ml:19-35): (states, weighted cost) */
value nonadditive_CAML_median2(value prev, value x, value y) {
  CAMLparam3(prev, x, y);
  CAMLlocal3(out, cb, c);
  if (Na_nsites(x) != Na_nsites(y)) caml_invalid_argument("median_2: different site counts");
  out = na_out(prev, x);
  cb = na_cost_buf();
  check(phy_fitch_median2(ctx, Na_states(x), Na_states(y), (uint64_t*)Dptr(out), Na_nsites(x),
                          Na_weights(x), (uint64_t*)Dptr(cb)));
  c = caml_copy_double(read_cost((uint64_t*)Dptr(cb)));
  CAMLreturn(pair(out, c));
}

/* median_3 (the unrooted root, lib/node.ml:200-216) */
value nonadditive_CAML_median3(value prev, value a, value b, value cc) {
  CAMLparam4(prev, a, b, cc);
  CAMLlocal3(out, cb, c);
  if (Na_nsites(a) != Na_nsites(b) || Na_nsites(a) != Na_nsites(cc))
    caml_invalid_argument("median_3: different site counts");
  out = na_out(prev, a);
  cb = na_cost_buf();
  check(phy_fitch_median3(ctx, Na_states(a), Na_states(b), Na_states(cc), (uint64_t*)Dptr(out),
                          Na_nsites(a), Na_weights(a), (uint64_t*)Dptr(cb)));
  c = caml_copy_double(read_cost((uint64_t*)Dptr(cb)));
  CAMLreturn(pair(out, c));
}

/* distance_1 (bv_distance, lib/bitvector/bv.c:46-55, weighted) */
value nonadditive_CAML_distance1(value a, value b) {
  CAMLparam2(a, b);
  CAMLlocal1(cb);
  cb = na_cost_buf();
  check(phy_fitch_distance1(ctx, Na_states(a), Na_states(b), Na_nsites(a), Na_weights(a),
                            (uint64_t*)Dptr(cb)));
  CAMLreturn(caml_copy_double(read_cost((uint64_t*)Dptr(cb))));
}

/* distance_2: insertion cost of c on the edge (a, b) */
value nonadditive_CAML_distance2(value a, value b, value c) {
  CAMLparam3(a, b, c);
  CAMLlocal1(cb);
  cb = na_cost_buf();
  check(phy_fitch_distance2(ctx, Na_states(a), Na_states(b), Na_states(c), Na_nsites(a),
                            Na_weights(a), (uint64_t*)Dptr(cb)));
  CAMLreturn(caml_copy_double(read_cost((uint64_t*)Dptr(cb))));
}

/* adjust_3: Fitch (1971) final sets of n from its children's preliminary
 * sets and its parent's final sets; (states, changed) */
value nonadditive_CAML_adjust3(value n, value a, value b, value fp) {
  CAMLparam4(n, a, b, fp);
  CAMLlocal2(out, cb);
  out = dbuf_alloc(sizeof(uint64_t) * (size_t)phy_words(Na_nsites(n)));
  cb = na_cost_buf();
  check(phy_fitch_adjust3(ctx, Na_states(n), Na_states(a), Na_states(b), Na_states(fp),
                          (uint64_t*)Dptr(out), Na_nsites(n), (uint64_t*)Dptr(cb)));
  CAMLreturn(pair(out, Val_bool(read_cost((uint64_t*)Dptr(cb)) > 0.0)));
}

/* filter_codes: the sites at the given indices, (states, weights) */
value nonadditive_CAML_select(value x, value idx) {
  CAMLparam2(x, idx);
  CAMLlocal2(st, w);
  const mlsize_t n = Wosize_val(idx);
  int64_t* h = (int64_t*)malloc(sizeof(int64_t) * (n > 0 ? n : 1));
  mlsize_t i;
  int rc;
  if (!h) caml_failwith("select: out of memory");
  for (i = 0; i < n; ++i) h[i] = Long_val(Field(idx, i));
  st = dbuf_alloc(sizeof(uint64_t) * (size_t)phy_words((int64_t)n));
  w = dbuf_alloc(sizeof(uint32_t) * n);
  rc = phy_fitch_select(ctx, Na_states(x), Na_nsites(x), Na_weights(x), h, (int64_t)n,
                        (uint64_t*)Dptr(st), (uint32_t*)Dptr(w));
  free(h);
  check(rc);
  CAMLreturn(pair(st, w));
}

/* Whole-tree node arrays: (2N-1) rows of `stride` words; the leaf rows are
 * the given int arrays of sets.  Returns (rows, stride). */
value nonadditive_CAML_tree_rows(value leaves, value n_nodes) {
  CAMLparam2(leaves, n_nodes);
  CAMLlocal2(res, sv);
  const mlsize_t N = Wosize_val(leaves);
  const int64_t M = Long_val(n_nodes);
  int64_t S, stride;
  uint64_t* h;
  mlsize_t i;
  int rc;
  if (N < 1 || M < (int64_t)N) caml_invalid_argument("tree_rows: sizes");
  S = (int64_t)Wosize_val(Field(leaves, 0));
  stride = phy_words(S);
  h = (uint64_t*)calloc((size_t)(M * stride > 0 ? M * stride : 1), sizeof(uint64_t));
  if (!h) caml_failwith("tree_rows: out of memory");
  for (i = 0; i < N; ++i) {
    if ((int64_t)Wosize_val(Field(leaves, i)) != S) {
      free(h);
      caml_invalid_argument("tree_rows: ragged leaves");
    }
    pack_sets(Field(leaves, i), h + i * stride, stride);
  }
  res = dbuf_alloc(sizeof(uint64_t) * (size_t)(M * stride));
  rc = phy_copy_h2d(ctx, Dptr(res), h, sizeof(uint64_t) * (size_t)(M * stride));
  free(h);
  check(rc);
  sv = Val_long(stride);
  CAMLreturn(pair(res, sv));
}

/* Node.Make1D.median_2 over the whole post-order in one launch: fills the
 * internal rows; returns the weighted total cost (lib/node.ml:183-198,272)
 * and the number of union events at each site */
value nonadditive_CAML_downpass(value topo, value rows, value stride, value weights,
                                value n_sites) {
  CAMLparam5(topo, rows, stride, weights, n_sites);
  CAMLlocal3(tot, sc, arr);
  const int64_t S = Long_val(n_sites);
  uint32_t* h = (uint32_t*)malloc(sizeof(uint32_t) * (size_t)(S > 0 ? S : 1));
  int64_t s;
  int rc;
  double total;
  if (!h) caml_failwith("downpass: out of memory");
  tot = na_cost_buf();
  sc = dbuf_alloc(sizeof(uint32_t) * (size_t)S);
  rc = phy_fitch_downpass(ctx, Topo_val(topo), (uint64_t*)Dptr(rows), Long_val(stride), S,
                          (const uint32_t*)Dptr(weights), (uint32_t*)Dptr(sc), NULL,
                          (uint64_t*)Dptr(tot));
  if (rc == PHY_OK) {
    caml_enter_blocking_section();
    rc = phy_copy_d2h(ctx, h, Dptr(sc), sizeof(uint32_t) * (size_t)S);
    caml_leave_blocking_section();
  }
  if (rc != PHY_OK) {
    free(h);
    check(rc);
  }
  total = read_cost((uint64_t*)Dptr(tot));
  arr = caml_alloc((mlsize_t)S, 0);
  for (s = 0; s < S; ++s) Store_field(arr, s, Val_long(h[s]));
  free(h);
  CAMLreturn(pair(caml_copy_double(total), arr));
}

/* final sets of every node in one pre-order launch */
value nonadditive_CAML_uppass(value topo, value rows, value final, value stride, value n_sites) {
  CAMLparam5(topo, rows, final, stride, n_sites);
  check(phy_fitch_uppass(ctx, Topo_val(topo), (const uint64_t*)Dptr(rows),
                         (uint64_t*)Dptr(final), Long_val(stride), Long_val(n_sites)));
  CAMLreturn(Val_unit);
}

/* one row of a node array back on the host */
value nonadditive_CAML_row(value rows, value stride, value row, value n_sites) {
  CAMLparam4(rows, stride, row, n_sites);
  CAMLlocal1(res);
  const int64_t st = Long_val(stride), r = Long_val(row), n = Long_val(n_sites);
  uint64_t* h = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(st > 0 ? st : 1));
  int64_t s;
  int rc;
  if (!h) caml_failwith("row: out of memory");
  if ((size_t)((r + 1) * st) * sizeof(uint64_t) > Dbuf_val(rows)->bytes || phy_words(n) > st) {
    free(h);
    caml_invalid_argument("row: out of range");
  }
  caml_enter_blocking_section();
  rc = phy_copy_d2h(ctx, h, (const uint64_t*)Dptr(rows) + r * st, sizeof(uint64_t) * (size_t)st);
  caml_leave_blocking_section();
  if (rc != PHY_OK) {
    free(h);
    check(rc);
  }
  res = caml_alloc((mlsize_t)n, 0);
  for (s = 0; s < n; ++s) Store_field(res, s, Val_long((h[s / 16] >> (4 * (s % 16))) & 15));
  free(h);
  CAMLreturn(res);
}

/* ------------------------------------------------------------------------- */
/* Likelihood_c: t = {part; exps; blen; n_pat; pitch; weights; f32; lnl;      */
/* codes; model}                                                             */
/* ------------------------------------------------------------------------- */
enum {
  LK_PART = 0, LK_EXPS = 1, LK_BLEN = 2, LK_NPAT = 3, LK_PITCH = 4,
  LK_WEIGHTS = 5, LK_F32 = 6, LK_LNL = 7, LK_CODES = 8, LK_MODEL = 9
};
#define Lk_part(v) Dptr(Field(v, LK_PART))
#define Lk_exps(v) ((int32_t*)Dptr(Field(v, LK_EXPS)))
#define Lk_blen(v) Double_val(Field(v, LK_BLEN))
#define Lk_npat(v) ((int64_t)Long_val(Field(v, LK_NPAT)))
#define Lk_pitch(v) ((int64_t)Long_val(Field(v, LK_PITCH)))
#define Lk_weights(v) ((const uint32_t*)Dptr(Field(v, LK_WEIGHTS)))
#define Lk_form(v) (Bool_val(Field(v, LK_F32)) ? PHY_F32 : PHY_F64)
#define Lk_model(v) Model_val(Field(v, LK_MODEL))

/* MlModel.t (lib/mlModel.ml:53-63: spec, priors, pinvar, rates, probs, q,
 * ...) -> device model (pi, eigen-decomposition of q, rates, probs) */
value likelihood_CAML_model(value m) {
  CAMLparam1(m);
  CAMLlocal1(v);
  struct caml_ba_array* pri = Caml_ba_array_val(Field(m, 1));
  struct caml_ba_array* rat = Caml_ba_array_val(Field(m, 3));
  struct caml_ba_array* pro = Caml_ba_array_val(Field(m, 4));
  struct caml_ba_array* q = Caml_ba_array_val(Field(m, 5));
  const int32_t K = (int32_t)pri->dim[0], R = (int32_t)rat->dim[0];
  phy_model* h = NULL;
  if (q->num_dims != 2 || q->dim[0] != K || q->dim[1] != K || pro->dim[0] != R)
    caml_invalid_argument("model: inconsistent sizes");
  check(phy_model_create(get_ctx(), K, R, (const double*)pri->data, (const double*)q->data,
                         (const double*)rat->data, (const double*)pro->data, &h));
  v = caml_alloc_custom(&model_ops, sizeof(dmodel), 0, 1);
  Dmodel_val(v)->h = h;
  Dmodel_val(v)->K = K;
  Dmodel_val(v)->R = R;
  CAMLreturn(v);
}

/* pitch of a value: the pattern count rounded up to 64 */
static int64_t lk_pitch(int64_t n) { return ((n + 63) / 64) * 64; }

/* bytes of a value's partial (R*K rows) and exponents (R rows) */
static void lk_sizes(value model, int f32, int64_t pitch, size_t* part, size_t* exps) {
  const size_t K = (size_t)Dmodel_val(model)->K, R = (size_t)Dmodel_val(model)->R;
  *part = R * K * (size_t)pitch * (f32 ? 4 : 8);
  *exps = R * (size_t)pitch * 4;
}

/* median_1 of a leaf: (part, exps, pitch) from 4-bit tip sets */
value likelihood_CAML_of_tips(value model, value tips, value f32) {
  CAMLparam3(model, tips, f32);
  CAMLlocal4(tip, part, exps, res);
  const int64_t n = (int64_t)Wosize_val(tips), pitch = lk_pitch(n);
  size_t pb, eb;
  if (n < 1) caml_invalid_argument("of_tips: no patterns");
  lk_sizes(model, Bool_val(f32), pitch, &pb, &eb);
  tip = pack_upload(tips);
  part = dbuf_alloc(pb);
  exps = dbuf_alloc(eb);
  check(phy_lk_leaf(ctx, Model_val(model), Bool_val(f32) ? PHY_F32 : PHY_F64, n,
                    (const uint64_t*)Dptr(tip), Dptr(part), (int32_t*)Dptr(exps), pitch));
  res = caml_alloc_tuple(3);
  Store_field(res, 0, part);
  Store_field(res, 1, exps);
  Store_field(res, 2, Val_long(pitch));
  CAMLreturn(res);
}

static void lk_same(value x, value y, const char* what) {
  if (Lk_npat(x) != Lk_npat(y) || Lk_pitch(x) != Lk_pitch(y) || Lk_form(x) != Lk_form(y))
    caml_invalid_argument(what);
}

/* median_2: (part, exps) of (P(t_x) L_x) * (P(t_y) L_y) per rate class */
value likelihood_CAML_median2(value prev, value x, value y) {
  CAMLparam3(prev, x, y);
  CAMLlocal2(part, exps);
  lk_same(x, y, "median_2: different pattern sets or forms");
  if (Is_block(prev) && Lk_pitch(Some_val(prev)) == Lk_pitch(x) &&
      Lk_form(Some_val(prev)) == Lk_form(x)) {
    part = Field(Some_val(prev), LK_PART);
    exps = Field(Some_val(prev), LK_EXPS);
  } else {
    size_t pb, eb;
    lk_sizes(Field(x, LK_MODEL), Lk_form(x) == PHY_F32, Lk_pitch(x), &pb, &eb);
    part = dbuf_alloc(pb);
    exps = dbuf_alloc(eb);
  }
  check(phy_lk_median2(ctx, Lk_model(x), Lk_form(x), Lk_npat(x), Lk_part(x), Lk_exps(x),
                       Lk_blen(x), Lk_part(y), Lk_exps(y), Lk_blen(y), Dptr(part),
                       (int32_t*)Dptr(exps), Lk_pitch(x)));
  CAMLreturn(pair(part, exps));
}

static double read_double(double* d) {
  double h = 0.0;
  copy_d2h(&h, d, sizeof h);
  return h;
}

/* median_3: (part, exps, lnL) of the three-neighbour join */
value likelihood_CAML_median3(value x, value y, value z) {
  CAMLparam3(x, y, z);
  CAMLlocal4(part, exps, lb, res);
  size_t pb, eb;
  lk_same(x, y, "median_3: different pattern sets or forms");
  lk_same(x, z, "median_3: different pattern sets or forms");
  lk_sizes(Field(x, LK_MODEL), Lk_form(x) == PHY_F32, Lk_pitch(x), &pb, &eb);
  part = dbuf_alloc(pb);
  exps = dbuf_alloc(eb);
  lb = dbuf_alloc(sizeof(double));
  check(phy_lk_median3(ctx, Lk_model(x), Lk_form(x), Lk_npat(x), Lk_part(x), Lk_exps(x),
                       Lk_blen(x), Lk_part(y), Lk_exps(y), Lk_blen(y), Lk_part(z), Lk_exps(z),
                       Lk_blen(z), Dptr(part), (int32_t*)Dptr(exps), Lk_pitch(x),
                       Lk_weights(x), (double*)Dptr(lb)));
  res = caml_alloc_tuple(3);
  Store_field(res, 0, part);
  Store_field(res, 1, exps);
  Store_field(res, 2, caml_copy_double(read_double((double*)Dptr(lb))));
  CAMLreturn(res);
}

/* root_cost: lnL with x as the root */
value likelihood_CAML_root(value x) {
  CAMLparam1(x);
  CAMLlocal1(lb);
  lb = dbuf_alloc(sizeof(double));
  check(phy_lk_root(ctx, Lk_model(x), Lk_form(x), Lk_npat(x), Lk_part(x), Lk_exps(x),
                    Lk_pitch(x), Lk_weights(x), (double*)Dptr(lb)));
  CAMLreturn(caml_copy_double(read_double((double*)Dptr(lb))));
}

/* (lnL, dlnL/dt, d2lnL/dt2) of x and y joined by one edge of length t */
value likelihood_CAML_edge(value x, value y, value t) {
  CAMLparam3(x, y, t);
  CAMLlocal2(lb, res);
  double h[3];
  int i;
  lk_same(x, y, "edge: different pattern sets or forms");
  lb = dbuf_alloc(3 * sizeof(double));
  check(phy_lk_edge(ctx, Lk_model(x), Lk_form(x), Lk_npat(x), Lk_part(x), Lk_exps(x),
                    Lk_part(y), Lk_exps(y), Lk_pitch(x), Double_val(t), Lk_weights(x),
                    (double*)Dptr(lb)));
  copy_d2h(h, Dptr(lb), sizeof h);
  res = caml_alloc(3 * Double_wosize, Double_array_tag);
  for (i = 0; i < 3; ++i) Store_double_field(res, i, h[i]);
  CAMLreturn(res);
}

/* adjust_3: Newton on the edge between x (a node's own value) and y (the
 * rest of the tree at its parent), from t0; (t, lnL) */
value likelihood_CAML_adjust3(value x, value y, value t0) {
  CAMLparam3(x, y, t0);
  CAMLlocal1(res);
  double t = 0.0, lnl = 0.0;
  int32_t evals = 0;
  lk_same(x, y, "adjust_3: different pattern sets or forms");
  check(phy_lk_adjust3(ctx, Lk_model(x), Lk_form(x), Lk_npat(x), Lk_part(x), Lk_exps(x),
                       Lk_part(y), Lk_exps(y), Lk_pitch(x), Lk_weights(x), Double_val(t0),
                       1e-8, 10.0, 20, 1e-8, &t, &lnl, &evals));
  res = caml_alloc(2 * Double_wosize, Double_array_tag);
  Store_double_field(res, 0, t);
  Store_double_field(res, 1, lnl);
  CAMLreturn(res);
}

/* Whole tree: lnL and dlnL/dt, d2lnL/dt2 of every branch in one traversal
 * (the derivatives behind a Newton adjust_3 of all branches at once).
 * data = {tips (dbuf, N rows of stride words); stride; n_pat; weights}. */
value likelihood_CAML_derivs(value topo, value model, value data, value blen, value f32) {
  CAMLparam5(topo, model, data, blen, f32);
  CAMLlocal4(work, out, d1v, d2v);
  const int64_t stride = Long_val(Field(data, 1)), n_pat = Long_val(Field(data, 2));
  const mlsize_t M = Wosize_val(blen) / Double_wosize;
  const int32_t form = Bool_val(f32) ? PHY_F32 : PHY_F64;
  size_t bytes = 0;
  double* hb;
  double* h;
  mlsize_t i;
  int rc;
  check(phy_lk_tree_work_size(Topo_val(topo), Model_val(model), n_pat, form, 1, &bytes));
  work = dbuf_alloc(bytes);
  out = dbuf_alloc(sizeof(double) * (1 + 3 * M));
  hb = (double*)malloc(sizeof(double) * (M > 0 ? M : 1));
  h = (double*)malloc(sizeof(double) * (1 + 2 * M));
  if (!hb || !h) {
    free(hb);
    free(h);
    caml_failwith("derivs: out of memory");
  }
  for (i = 0; i < M; ++i) hb[i] = Double_field(blen, i);
  {
    double* d = (double*)Dptr(out);
    rc = phy_copy_h2d(ctx, d + 1 + 2 * M, hb, sizeof(double) * M);
    if (rc == PHY_OK)
      rc = phy_lk_tree(ctx, Topo_val(topo), Model_val(model), form, n_pat,
                       (const uint64_t*)Dptr(Field(data, 0)), stride,
                       (const uint32_t*)Dptr(Field(data, 3)), d + 1 + 2 * M, d, NULL, d + 1,
                       d + 1 + M, Dptr(work), bytes);
    free(hb);
    if (rc != PHY_OK) {
      free(h);
      check(rc);
    }
    caml_enter_blocking_section();
    rc = phy_copy_d2h(ctx, h, d, sizeof(double) * (1 + 2 * M));
    caml_leave_blocking_section();
  }
  if (rc != PHY_OK) {
    free(h);
    check(rc);
  }
  d1v = caml_alloc(M * Double_wosize, Double_array_tag);
  d2v = caml_alloc(M * Double_wosize, Double_array_tag);
  for (i = 0; i < M; ++i) {
    Store_double_field(d1v, i, h[1 + i]);
    Store_double_field(d2v, i, h[1 + M + i]);
  }
  out = caml_alloc_tuple(3);
  Store_field(out, 0, caml_copy_double(h[0]));
  Store_field(out, 1, d1v);
  Store_field(out, 2, d2v);
  free(h);
  CAMLreturn(out);
}

/* ------------------------------------------------------------------------- */
/* MlModel: the LAPACK stubs of lib/mlmodel.c under the same names            */
/* ------------------------------------------------------------------------- */
/* diagonalize_gtr U D Ui: U (holding Q) is overwritten by the eigenvectors,
 * D by diag(lambda), Ui by U^-1 (lib/mlmodel.c:208-271) */
value likelihood_CAML_diagonalize_gtr(value U, value D, value Ui) {
  CAMLparam3(U, D, Ui);
  const int32_t n = (int32_t)Caml_ba_array_val(U)->dim[0];
  double* q = (double*)malloc(sizeof(double) * (size_t)n * (size_t)n);
  int rc;
  if (!q) caml_failwith("diagonalize_gtr: out of memory");
  memcpy(q, Caml_ba_data_val(U), sizeof(double) * (size_t)n * (size_t)n);
  rc = phy_diagonalize_gtr(n, q, (double*)Caml_ba_data_val(U), (double*)Caml_ba_data_val(D),
                           (double*)Caml_ba_data_val(Ui));
  free(q);
  if (rc != PHY_OK) caml_failwith("Imaginary eigenvalues or invalid rate matrix");
  CAMLreturn(Val_unit);
}

/* diagonalize_sym U D (lib/mlmodel.c:163-205) */
value likelihood_CAML_diagonalize_sym(value U, value D) {
  CAMLparam2(U, D);
  const int32_t n = (int32_t)Caml_ba_array_val(U)->dim[0];
  if (phy_diagonalize_sym(n, (double*)Caml_ba_data_val(U), (double*)Caml_ba_data_val(D)) != PHY_OK)
    caml_failwith("diagonalization failed to converge");
  CAMLreturn(Val_unit);
}

static value new_matrix(intnat n) {
  intnat dims[2];
  dims[0] = n;
  dims[1] = n;
  /* managed: freed with the value (the reference's P leaks, lib/mlmodel.c:320,362) */
  return caml_ba_alloc(CAML_BA_FLOAT64 | CAML_BA_C_LAYOUT | CAML_BA_MANAGED, 2, NULL, dims);
}

/* compose_gtr U D Ui t -> P = exp(Q t) (lib/mlmodel.c:325-365) */
value likelihood_CAML_compose_gtr(value U, value D, value Ui, value t) {
  CAMLparam4(U, D, Ui, t);
  CAMLlocal1(P);
  const intnat n = Caml_ba_array_val(U)->dim[0];
  P = new_matrix(n);
  if (phy_compose_gtr((int32_t)n, (const double*)Caml_ba_data_val(U),
                      (const double*)Caml_ba_data_val(D), (const double*)Caml_ba_data_val(Ui),
                      Double_val(t), (double*)Caml_ba_data_val(P)) != PHY_OK)
    caml_failwith("compose_gtr: bad argument");
  CAMLreturn(P);
}

/* compose_sym U D t -> P, t rounded to single precision as the reference's
 * `const float t` does (lib/mlmodel.c:280) */
value likelihood_CAML_compose_sym(value U, value D, value t) {
  CAMLparam3(U, D, t);
  CAMLlocal1(P);
  const intnat n = Caml_ba_array_val(U)->dim[0];
  P = new_matrix(n);
  if (phy_compose_sym((int32_t)n, (const double*)Caml_ba_data_val(U),
                      (const double*)Caml_ba_data_val(D), (float)Double_val(t),
                      (double*)Caml_ba_data_val(P)) != PHY_OK)
    caml_failwith("compose_sym: bad argument");
  CAMLreturn(P);
}
