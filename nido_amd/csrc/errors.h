// errors.h -- per-thread error message shared by the C-ABI translation units.
#pragma once
#include <stdarg.h>
#include <stdio.h>

namespace nido {
// Records a message for nido_last_error() and returns `code`.
int set_error(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
}  // namespace nido
