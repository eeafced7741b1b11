// Thread-local error message behind ctlproj_last_error.
#include "errors.h"

namespace ctlrt {

namespace {
thread_local std::string g_error;
} // namespace

ctlproj_status fail(ctlproj_status st, const std::string& msg)
{
    g_error = msg;
    return st;
}

void clearError() { g_error.clear(); }

const char* lastError() { return g_error.c_str(); }

} // namespace ctlrt
