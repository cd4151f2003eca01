// Thread-local error state of the C ABI.
#include "sie_internal.hpp"

namespace sie {
static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }
}  // namespace sie

extern "C" const char* sie_last_error(void) { return sie::g_last_error.c_str(); }
