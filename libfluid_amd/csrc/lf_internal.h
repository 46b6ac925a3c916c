// lf_internal.h -- shared host-side internals of libfluid_hip.so.
#pragma once

#include "lf_abi.h"

namespace lf {

// Records a printf-style message as the calling thread's last error and
// returns `code`.
int fail(int code, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();

}  // namespace lf
