/* Test double of OCaml's <caml/alloc.h>. */
#ifndef MOCK_CAML_ALLOC_H
#define MOCK_CAML_ALLOC_H
#include "mlvalues.h"

value caml_alloc(mlsize_t wosize, tag_t tag);
value caml_alloc_tuple(mlsize_t wosize);
value caml_alloc_float_array(mlsize_t len);
value caml_copy_double(double d);
value caml_copy_string(const char* s);

#endif
