/* Test double of OCaml's <caml/mlvalues.h>: the subset of the public C API
 * the stubs use, with OCaml's value representation (tagged integers, a
 * header word before each block: wosize << 10 | colour << 8 | tag), so a
 * stub compiled against it reads and writes values as it would in OCaml. */
#ifndef MOCK_CAML_MLVALUES_H
#define MOCK_CAML_MLVALUES_H
#include <stdint.h>

typedef intptr_t intnat;
typedef uintptr_t uintnat;
typedef intnat value;
typedef uintnat header_t;
typedef uintnat mlsize_t;
typedef unsigned int tag_t;

#define Val_long(x) ((intnat)(((uintnat)(x) << 1)) + 1)
#define Long_val(x) ((x) >> 1)
#define Val_int(x) Val_long(x)
#define Int_val(x) ((int)Long_val(x))
#define Is_long(x) (((x) & 1) != 0)
#define Is_block(x) (((x) & 1) == 0)
#define Val_false Val_int(0)
#define Val_true Val_int(1)
#define Val_bool(x) Val_int((x) != 0)
#define Bool_val(x) Int_val(x)
#define Val_unit Val_int(0)
#define Val_none Val_int(0)
#define Some_val(v) Field(v, 0)

#define Hd_val(v) (((header_t*)(v))[-1])
#define Wosize_hd(hd) ((mlsize_t)((hd) >> 10))
#define Tag_hd(hd) ((tag_t)((hd)&0xFF))
#define Wosize_val(v) Wosize_hd(Hd_val(v))
#define Tag_val(v) Tag_hd(Hd_val(v))
#define Make_header(wosize, tag) ((header_t)(((header_t)(wosize) << 10) + (tag)))

#define Field(x, i) (((value*)(x))[i])
#define Bp_val(v) ((char*)(v))
#define String_val(x) ((const char*)Bp_val(x))

#define String_tag 252
#define Double_tag 253
#define Double_array_tag 254
#define Custom_tag 255

#define Double_wosize ((sizeof(double) / sizeof(value)))
#define Double_val(v) (*(double*)(v))
#define Double_field(v, i) (((double*)(v))[i])
#define Store_double_field(v, i, d) (((double*)(v))[i] = (d))

#endif
