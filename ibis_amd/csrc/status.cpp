// status.cpp -- the library's error reporting (host only).
//
// Every entry point that fails records a message here and returns the
// status; ibis_last_error() returns the calling thread's last message.
// Kept apart from the HIP sources so host-only parts of the library (the
// grid builder) also build with a plain C++ compiler, e.g. under the
// address and undefined-behaviour sanitizers.

#include <string>

#include "ibis_internal.h"

namespace {
thread_local std::string g_last_error;
}  // namespace

namespace ibis {
ibis_status fail(ibis_status s, const char* msg) {
    g_last_error = msg ? msg : "";
    return s;
}
}  // namespace ibis

extern "C" const char* ibis_last_error(void) { return g_last_error.c_str(); }
