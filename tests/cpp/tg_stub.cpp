// tg_stub.cpp - a device-less stand-in for the C ABI of include/teufel_gpu.h,
// linked instead of the HIP library into the sanitizer build of the host code
// (tests/test_host_sanitizers.py).  Address and undefined-behaviour
// sanitizers instrument host code only; with this stand-in the host mirror's
// file readers and writers, observers and buffers run under them without a
// GPU or the HIP runtime in the process.  Every entry point the host code
// calls reports TG_ERR_HIP ("no device"), which the adapter turns into an
// IOexception, as the real library does on a machine without a GPU.
#include "../../include/teufel_gpu.h"

extern "C" {

int tg_create(int, tg_context** ctx)
{
    if (ctx) *ctx = nullptr;
    return TG_ERR_HIP;
}
int tg_destroy(tg_context*) { return TG_ERR_ARG; }
const char* tg_last_error(const tg_context*) { return "no device (host-only stand-in of the C ABI)"; }
int tg_set_trajectories(tg_context*, int, int, const double*, const double*, const double*, const double*,
                        const double*, const double*, const double*, const double*, const double*, const double*,
                        const double*)
{
    return TG_ERR_HIP;
}
int tg_add_screen(tg_context*, int, const double*, const double*, double, int, int*) { return TG_ERR_HIP; }
int tg_remove_screens(tg_context*) { return TG_ERR_HIP; }
int tg_integrate(tg_context*, int, const tg_tuning*) { return TG_ERR_HIP; }
int tg_read_traces(tg_context*, int, double*, int) { return TG_ERR_HIP; }
int tg_fields_at(tg_context*, int, const double*, const double*, int, int, double*, int) { return TG_ERR_HIP; }
int tg_deposit_only(tg_context*, int, int, const double*, const double*, double, double, int, double*)
{
    return TG_ERR_HIP;
}
}
