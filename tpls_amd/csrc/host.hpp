// Host-side helpers shared by the C ABI sources (case.cpp, io.cpp).
#pragma once

#include <string>

#include "../../include/tpls_core.h"

namespace tpls {

// Records msg as the calling thread's tpls_last_error() and returns code.
int set_error(int code, const std::string& msg);
// maxl, maxm, maxn of a case.
int case_dims(const tpls_case* c, int* maxl, int* maxm, int* maxn);

}  // namespace tpls
