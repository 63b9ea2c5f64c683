// range_plan.h -- host planner of (target range, source range) pair lists,
// the batched form of FastMultipole's direct list (fmm.direct! per pair,
// src/FLOWVPM_fmm.jl:102-168).  Host code only (no HIP), so that it builds
// with g++ and runs under the host sanitizers; the device side sees the same
// records through layout-compatible int2 / struct views.
#pragma once
#include <cstddef>
#include <cstdint>
#include <string>

namespace vpm {

// thread-local error text of the library (set on the caller's thread only)
int fail(int status, const std::string& msg);

// One unit of device work: a block (or wave) owns targets [t_begin, t_end)
// of one elementary interval and streams the source ranges
// src[l_begin .. l_end) of the plan.
struct RangeItem {
    int64_t t_begin, t_end;
    int64_t l_begin, l_end;
};

// A source range [b, e); positions fit 32 bits (fields hold < 2^31
// particles).  Same layout as the device's int2.
struct SrcRange {
    int32_t b, e;
};
static_assert(sizeof(SrcRange) == 8, "SrcRange is an int2");

struct PlanShape {
    int64_t nitems = 0, nsrc = 0;
    bool src_copied = false;   // the sink already holds the source ranges on the device
    bool wave = false;         // items of <= 64 / R targets (FMM leaves), one wave each
    int R = 1;                 // lane groups per target of the wave kernels
};

// Where a plan goes.  stage() hands out host memory for nitems items
// followed by nsrc source ranges (one block, items first).  For leaf lists
// the packed source ranges are also copied piece by piece as they are
// written: prepare_copies() is called once before the workers start, and
// copy_piece(a, b) from the worker threads for source ranges [a, b); it
// returns 0 or a nonzero code that report() turns into a status on the
// caller's thread.
struct PlanSink {
    virtual ~PlanSink() = default;
    virtual int stage(int64_t nitems, int64_t nsrc, RangeItem*& items, SrcRange*& src) = 0;
    virtual int prepare_copies(int64_t nitems, int64_t nsrc) = 0;
    virtual int copy_piece(int64_t a, int64_t b) = 0;
    virtual int report(int code) = 0;
};

// Plans npairs (target, source) range pairs over a field of np particles.
// Every pair is checked (0 <= begin <= end <= np), target ranges before
// source ranges, the lowest bad index reported.  allow_wave: the caller has
// wave kernels, taken when the mean non-empty target range is at most 128
// targets; rmax caps the lane sharing R (1, 2, 4 or 8).  Host allocation
// failures return VPM_ENOMEM; no exception leaves this function.
// threads: at most this many host threads (0: the library's default).
int plan_ranges(int64_t np, const int64_t* tb, const int64_t* te, const int64_t* sb, const int64_t* se,
                int64_t npairs, bool allow_wave, int rmax, PlanSink& sink, PlanShape& shape, int threads = 0);

constexpr int kPlanThreads = 8;                      // host threads of a plan, at most
constexpr int64_t kPlanPiece = int64_t(1) << 18;     // pairs per thread at least, and per early copy

}  // namespace vpm
