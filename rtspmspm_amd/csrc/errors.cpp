// Error state of the library: exceptions inside, the message of the last
// failure per host thread for rt_last_error() at the C ABI.  Host code only.
#include <string>

#include "common.hpp"

namespace rt {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error_cstr() { return g_last_error.c_str(); }

[[noreturn]] void fail(int code, const std::string& msg) { throw Error(code, msg); }

}  // namespace rt
