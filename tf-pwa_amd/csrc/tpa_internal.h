// Internal glue between the library's translation units (not part of the C ABI).
#pragma once
#include <string>

// Sets the calling thread's tpa_last_error() message and returns code.
int tpa_set_error(int code, const std::string& msg);
