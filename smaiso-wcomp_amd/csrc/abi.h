// Internal helpers shared by the translation units that implement the C ABI.
#pragma once

#include <string>

namespace smaiso {

// Record msg as the text smaiso_last_error(NULL) returns (the calling
// thread's handle-less error) and return code.
int abi_fail(int code, const std::string& msg);

}  // namespace smaiso
