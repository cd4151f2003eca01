// Optional per-kernel timing with HIP events on the launch stream (used by the
// benchmark to measure one kernel's average duration live). Off by default;
// when off, begin/end cost one branch.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

namespace sie {

enum ProfKernel { PROF_GEMM_F32 = 0, PROF_GEMM_F64 = 1, PROF_FAR_PAIRS = 2, PROF_PANEL = 3, PROF_NEAR_PAIRS = 4,
                  PROF_GATHER = 5, PROF_GEMV = 6, PROF_GEMM_F16X3 = 7,
                  PROF_GEMM_F16X3_WIDE = 8, PROF_SWEEPS = 9, PROF_CONTRACT = 10, PROF_MATVEC = 11,
                  PROF_COUNT = 12 };

struct ProfRec {
    hipEvent_t a, b;
    double flops, bytes;
};

struct Prof {
    bool on[PROF_COUNT] = {};
    std::vector<ProfRec> recs[PROF_COUNT];
    ProfRec* begin(int k, hipStream_t s) {
        if (!on[k]) return nullptr;
        ProfRec r{};
        (void)hipEventCreate(&r.a);
        (void)hipEventCreate(&r.b);
        (void)hipEventRecord(r.a, s);
        recs[k].push_back(r);
        return &recs[k].back();
    }
    void end(int k, hipStream_t s, double flops, double bytes) {
        if (!on[k]) return;
        ProfRec& r = recs[k].back();
        r.flops = flops;
        r.bytes = bytes;
        (void)hipEventRecord(r.b, s);
    }
};

Prof& prof();

}  // namespace sie
