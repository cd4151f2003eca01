// Thread-local error state of the C ABI.
#include "sie_internal.hpp"

namespace sie {
static thread_local std::string g_last_error;
void set_last_error(const std::string& m) { g_last_error = m; }
}  // namespace sie

extern "C" const char* sie_last_error(void) { return sie::g_last_error.c_str(); }

#include "prof.hpp"

namespace sie {
Prof& prof() {
    static Prof p;
    return p;
}
}  // namespace sie

extern "C" {

int sie_profile_enable(int kernel, int on) {
    if (kernel < 0 || kernel >= sie::PROF_COUNT) {
        sie::set_last_error("unknown kernel id");
        return SIE_ERR_INVALID;
    }
    auto& p = sie::prof();
    p.on[kernel] = on != 0;
    for (auto& r : p.recs[kernel]) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    p.recs[kernel].clear();
    return SIE_OK;
}

int sie_profile_read(int kernel, int* launches, double* total_ms, double* flops, double* bytes) {
    if (kernel < 0 || kernel >= sie::PROF_COUNT) {
        sie::set_last_error("unknown kernel id");
        return SIE_ERR_INVALID;
    }
    auto& p = sie::prof();
    double ms = 0, fl = 0, by = 0;
    for (auto& r : p.recs[kernel]) {
        (void)hipEventSynchronize(r.b);
        float t = 0;
        (void)hipEventElapsedTime(&t, r.a, r.b);
        ms += t;
        fl += r.flops;
        by += r.bytes;
    }
    *launches = (int)p.recs[kernel].size();
    *total_ms = ms;
    *flops = fl;
    *bytes = by;
    return SIE_OK;
}

}  // extern "C"
