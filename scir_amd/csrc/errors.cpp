// Error reporting and allocation-failure injection of libscir_hip.so (host
// code only).  The per-thread message is a fixed buffer, so recording an
// error never allocates and cannot itself fail.
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <new>

#include "host_common.hpp"

namespace scir {

namespace {
thread_local char t_last_error[1024];
}  // namespace

int fail(int code, const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_last_error, sizeof t_last_error, fmt, ap);
    va_end(ap);
    return code;
}

void clear_error() { t_last_error[0] = '\0'; }

bool alloc_failure_injected()
{
    const char *v = std::getenv("SCIR_HIP_FAIL_ALLOC");
    return v && std::atoi(v) != 0;
}

void host_alloc_point()
{
    if (alloc_failure_injected()) throw std::bad_alloc();
}

}  // namespace scir

extern "C" {

int scir_hip_version(void) { return 10100; }

const char *scir_hip_last_error(void) { return scir::t_last_error; }

}  // extern "C"
