// status.h -- error reporting shared by the engine's C entry points: the
// thread-local message read back through fwi_last_error().
#pragma once

namespace fwi {

// Records the formatted message for this thread and returns `code`.
int fail(int code, const char *fmt, ...);

}  // namespace fwi
