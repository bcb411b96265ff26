#include "rn_common.h"

#include <cstdlib>
#include <cstring>

namespace rn {

namespace {
thread_local std::string g_last_error;
}

void set_last_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }
void clear_last_error() { g_last_error.clear(); }

bool fault_injected(const char* what)
{
    const char* v = std::getenv("RECONNET_FAULT");
    return v != nullptr && std::strcmp(v, what) == 0;
}

} // namespace rn
