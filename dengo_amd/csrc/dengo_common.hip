// Network-independent parts of the C ABI: options and error reporting.
#include "dengo_solver.h"

namespace dengo {
namespace {
thread_local char g_err[512] = "";
thread_local double g_last_ms = 0.0;
}  // namespace

int set_error(const char *fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return -1;
}

void set_last_ms(double ms) { g_last_ms = ms; }

const char *last_error_text() { return g_err; }
}  // namespace dengo

extern "C" {

void dengo_default_options(dengo_options *opt)
{
    memset(opt, 0, sizeof(*opt));
    opt->precision = DENGO_FP64;
    opt->schedule = DENGO_PERSISTENT;
    opt->block = dengo::kBlock;
    opt->max_substeps = 50000;
}

const char *dengo_last_error(void) { return dengo::g_err; }

double dengo_last_kernel_ms(void) { return dengo::g_last_ms; }

}  // extern "C"
