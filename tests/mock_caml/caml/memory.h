/* Test double of OCaml's <caml/memory.h>: local-root registration macros.
 * The mock runtime never moves or collects values, so registration is
 * bookkeeping only; the macros keep OCaml's usage rules (CAMLparam first,
 * CAMLreturn to leave). */
#ifndef MOCK_CAML_MEMORY_H
#define MOCK_CAML_MEMORY_H
#include "mlvalues.h"

void caml_modify(value* fp, value v);
#define Store_field(block, offset, val) caml_modify(&Field(block, offset), (val))

#define CAMLparam0() int caml__frame = 0
#define CAMLparam1(x) CAMLparam0(); (void)(x)
#define CAMLparam2(x, y) CAMLparam1(x); (void)(y)
#define CAMLparam3(x, y, z) CAMLparam2(x, y); (void)(z)
#define CAMLparam4(x, y, z, t) CAMLparam3(x, y, z); (void)(t)
#define CAMLparam5(x, y, z, t, u) CAMLparam4(x, y, z, t); (void)(u)
#define CAMLlocal1(x) value x = Val_unit
#define CAMLlocal2(x, y) CAMLlocal1(x); CAMLlocal1(y)
#define CAMLlocal3(x, y, z) CAMLlocal2(x, y); CAMLlocal1(z)
#define CAMLlocal4(x, y, z, t) CAMLlocal3(x, y, z); CAMLlocal1(t)
#define CAMLlocal5(x, y, z, t, u) CAMLlocal4(x, y, z, t); CAMLlocal1(u)
#define CAMLreturnT(type, result) do { (void)caml__frame; return (result); } while (0)
#define CAMLreturn(result) CAMLreturnT(value, result)
#define CAMLreturn0 do { (void)caml__frame; return; } while (0)

#endif
