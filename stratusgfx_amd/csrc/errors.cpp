// Thread-local error message behind sgi_last_error().
#include <string>

#include "../../include/sgi.h"

static thread_local std::string g_lastError;

void sgi_set_error(const std::string& msg) { g_lastError = msg; }

extern "C" const char* sgi_last_error(void) { return g_lastError.c_str(); }
