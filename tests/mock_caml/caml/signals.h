/* Test double of OCaml's <caml/signals.h>. */
#ifndef MOCK_CAML_SIGNALS_H
#define MOCK_CAML_SIGNALS_H

void caml_enter_blocking_section(void);
void caml_leave_blocking_section(void);

#endif
