/* Driver-side helpers of the mock OCaml runtime (test infrastructure). */
#ifndef MOCK_CAML_H
#define MOCK_CAML_H
#include <setjmp.h>
#include <stddef.h>
#include <stdint.h>

#include "caml/mlvalues.h"

extern jmp_buf* mock_caml_handler; /* longjmp target of caml_failwith et al. */
extern char mock_caml_exn[512];    /* message of the last exception */
extern int mock_caml_exn_kind;     /* 1 Failure, 2 Invalid_argument */
int mock_caml_blocking(void);      /* depth of caml_enter_blocking_section */

value mock_int_array(const int64_t* x, size_t n);
value mock_float_array(const double* x, size_t n);
value mock_some(value x);
value mock_vector(const double* x, intnat n);
value mock_matrix(const double* x, intnat n);
void mock_finalize_all(void);
#endif
