// abi_internal.h -- shared internals of the C ABI implementation.
#pragma once

#include <string>

namespace amdat {

// Record a message for amdat_last_error() on this thread and return code.
int set_error(int code, const std::string &msg);
void clear_error();

}  // namespace amdat
