// hmvm_status.hpp -- the library's per-thread error state, for the
// translation units other than the runtime that owns it
// (hmvm_hip_last_error / hmvm_hip_last_error_message report it).
#pragma once

namespace hmvm {

// Sets the state (code and message, printed once to stderr when code is not
// HMVM_OK) and returns code.
int report_error(int code, const char *msg);
// Back to HMVM_OK.
void report_clear();

}  // namespace hmvm
