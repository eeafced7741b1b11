// Thread-local error message behind ctlproj_last_error (no device dependencies).
#pragma once

#include "ctlproj.h"

#include <string>

namespace ctlrt {

// Records `msg` as the calling thread's last error and returns `st`.
ctlproj_status fail(ctlproj_status st, const std::string& msg);
void clearError();
const char* lastError();

} // namespace ctlrt
