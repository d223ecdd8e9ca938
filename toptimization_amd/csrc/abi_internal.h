// Internal glue between the translation units of libtopt (not part of the
// C ABI): the thread-local error message behind topt_last_error().
#pragma once
#include <string>

namespace topt_abi {
int set_error(int code, const std::string& msg);
}
