// status.cpp -- the per-thread message behind gz_last_error().
#include <string>

#include "gz_common.h"

namespace gz {

namespace {
thread_local std::string g_last_error;
}

void SetError(const std::string &msg) { g_last_error = msg; }
const char *LastError() { return g_last_error.c_str(); }

}  // namespace gz
