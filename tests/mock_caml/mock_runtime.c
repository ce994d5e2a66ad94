/* A minimal stand-in for the OCaml runtime, enough to execute C stubs
 * compiled against the test headers in this directory: blocks with OCaml's
 * header layout, boxed floats, float arrays, custom blocks, bigarrays, and
 * exceptions as a longjmp to the driver's handler.  Nothing is ever
 * collected or moved (the driver's runs are short); finalizers run only when
 * the driver calls mock_finalize_all().  Test infrastructure only. */
#include <setjmp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "caml/alloc.h"
#include "caml/bigarray.h"
#include "caml/custom.h"
#include "caml/fail.h"
#include "caml/memory.h"
#include "caml/mlvalues.h"
#include "caml/signals.h"
#include "mock_caml.h"

jmp_buf* mock_caml_handler = NULL;
char mock_caml_exn[512];
int mock_caml_exn_kind = 0;
static int blocking = 0;

/* every block allocated, so custom blocks can be finalised at the end */
static value* all_blocks = NULL;
static size_t n_blocks = 0, cap_blocks = 0;
static header_t atoms[256];

value caml_alloc(mlsize_t wosize, tag_t tag) {
  header_t* p;
  if (wosize == 0) {
    atoms[tag] = Make_header(0, tag);
    return (value)(&atoms[tag] + 1);
  }
  p = (header_t*)calloc(wosize + 1, sizeof(header_t));
  if (!p) abort();
  p[0] = Make_header(wosize, tag);
  if (tag < String_tag) {
    mlsize_t i;
    for (i = 0; i < wosize; ++i) ((value*)(p + 1))[i] = Val_unit;
  }
  if (n_blocks == cap_blocks) {
    cap_blocks = cap_blocks ? 2 * cap_blocks : 1024;
    all_blocks = (value*)realloc(all_blocks, cap_blocks * sizeof(value));
    if (!all_blocks) abort();
  }
  all_blocks[n_blocks++] = (value)(p + 1);
  return (value)(p + 1);
}

value caml_alloc_tuple(mlsize_t wosize) { return caml_alloc(wosize, 0); }

value caml_alloc_float_array(mlsize_t len) {
  return caml_alloc(len * Double_wosize, Double_array_tag);
}

value caml_copy_double(double d) {
  value v = caml_alloc(Double_wosize, Double_tag);
  Double_val(v) = d;
  return v;
}

value caml_copy_string(const char* s) {
  const size_t n = strlen(s);
  const mlsize_t w = (n + sizeof(value)) / sizeof(value);
  value v = caml_alloc(w, String_tag);
  memcpy(Bp_val(v), s, n + 1);
  return v;
}

void caml_modify(value* fp, value v) { *fp = v; }

value caml_alloc_custom(struct custom_operations* ops, uintnat size, mlsize_t mem,
                        mlsize_t max) {
  const mlsize_t w = 1 + (size + sizeof(value) - 1) / sizeof(value);
  value v = caml_alloc(w, Custom_tag);
  (void)mem;
  (void)max;
  Custom_ops_val(v) = ops;
  return v;
}

void caml_register_custom_operations(struct custom_operations* ops) { (void)ops; }

static void raise_(int kind, const char* msg) {
  if (blocking) {
    fprintf(stderr, "mock runtime: exception raised inside a blocking section\n");
    abort();
  }
  snprintf(mock_caml_exn, sizeof mock_caml_exn, "%s", msg ? msg : "");
  mock_caml_exn_kind = kind;
  if (mock_caml_handler) longjmp(*mock_caml_handler, kind);
  fprintf(stderr, "uncaught exception %s: %s\n", kind == 1 ? "Failure" : "Invalid_argument",
          mock_caml_exn);
  exit(3);
}

void caml_failwith(const char* msg) { raise_(1, msg); exit(3); }
void caml_invalid_argument(const char* msg) { raise_(2, msg); exit(3); }

void caml_enter_blocking_section(void) {
  if (blocking++) {
    fprintf(stderr, "mock runtime: nested blocking section\n");
    abort();
  }
}
void caml_leave_blocking_section(void) {
  if (--blocking) {
    fprintf(stderr, "mock runtime: unbalanced blocking section\n");
    abort();
  }
}
int mock_caml_blocking(void) { return blocking; }

value caml_ba_alloc(int flags, int num_dims, void* data, intnat* dim) {
  const uintnat size = sizeof(struct caml_ba_array) + (uintnat)num_dims * sizeof(intnat);
  static struct custom_operations ba_ops = {.identifier = "_bigarray"};
  value v = caml_alloc_custom(&ba_ops, size, 0, 1);
  struct caml_ba_array* b = Caml_ba_array_val(v);
  intnat n = 1;
  int i;
  for (i = 0; i < num_dims; ++i) {
    b->dim[i] = dim[i];
    n *= dim[i];
  }
  b->num_dims = num_dims;
  b->flags = flags;
  b->proxy = NULL;
  if (!data) {
    const int kind = flags & CAML_BA_KIND_MASK;
    const size_t esz = kind == CAML_BA_FLOAT64 || kind == CAML_BA_INT64 ? 8 : 4;
    data = calloc((size_t)(n > 0 ? n : 1), esz);
    if (!data) abort();
    b->flags |= CAML_BA_MANAGED;
  }
  b->data = data;
  return v;
}

/* helpers for the driver */
value mock_int_array(const int64_t* x, size_t n) {
  value v = caml_alloc(n, 0);
  size_t i;
  for (i = 0; i < n; ++i) Field(v, i) = Val_long(x[i]);
  return v;
}
value mock_float_array(const double* x, size_t n) {
  value v = caml_alloc_float_array(n);
  size_t i;
  for (i = 0; i < n; ++i) Store_double_field(v, i, x[i]);
  return v;
}
value mock_some(value x) {
  value v = caml_alloc(1, 0);
  Field(v, 0) = x;
  return v;
}
value mock_vector(const double* x, intnat n) {
  value v = caml_ba_alloc(CAML_BA_FLOAT64 | CAML_BA_C_LAYOUT, 1, NULL, &n);
  memcpy(Caml_ba_data_val(v), x, sizeof(double) * (size_t)n);
  return v;
}
value mock_matrix(const double* x, intnat n) {
  intnat d[2];
  value v;
  d[0] = n;
  d[1] = n;
  v = caml_ba_alloc(CAML_BA_FLOAT64 | CAML_BA_C_LAYOUT, 2, NULL, d);
  if (x) memcpy(Caml_ba_data_val(v), x, sizeof(double) * (size_t)(n * n));
  return v;
}

void mock_finalize_all(void) {
  size_t i;
  for (i = 0; i < n_blocks; ++i) {
    const value v = all_blocks[i];
    if (Tag_val(v) == Custom_tag && Custom_ops_val(v) && Custom_ops_val(v)->finalize)
      Custom_ops_val(v)->finalize(v);
  }
}
