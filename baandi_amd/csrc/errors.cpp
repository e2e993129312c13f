// Thread-local last-error message of the C ABI.
#include <string>

#include "common.hpp"

namespace bd {
namespace {
thread_local std::string g_last_error;
}
void set_last_error(const std::string& msg) { g_last_error = msg; }
void clear_last_error() { g_last_error.clear(); }
}  // namespace bd

extern "C" const char* bd_last_error(void) { return bd::g_last_error.c_str(); }
