// Internal declarations shared by the host codec and the HIP launch code.
#pragma once
#include <cstdint>

#include "../../include/matvans.h"

namespace mv {
// Record `fmt` as this thread's last error and return `status`.
int fail(int status, const char *fmt, ...) __attribute__((format(printf, 2, 3)));
void clear_error();
}  // namespace mv
