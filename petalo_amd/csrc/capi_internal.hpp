// capi_internal.hpp -- error reporting shared by the C ABI translation units.
#pragma once

#include <string>

namespace petalo {

// Records the message returned by petalo_last_error() on this thread.
void set_error(const std::string& msg);
// set_error(msg) and return code.
int fail(int code, const std::string& msg);

}  // namespace petalo
