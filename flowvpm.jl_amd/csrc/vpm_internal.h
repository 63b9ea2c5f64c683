// vpm_internal.h -- shared declarations of the gfx950 P2P library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/flowvpm_hip.h"
#include "range_plan.h"

namespace vpm {

// Rows of the 46-double particle record (src/FLOWVPM_particlefield.jl:239-252)
enum : int {
    RX = 0, RG = 3, RSIG = 6, RVOL = 7, RCIRC = 8, RU = 9, RW = 12, RJ = 15,
    RPSE = 24, RM = 27, RC = 36, RSFS = 39, RSTATIC = 42, RUPREV = 43, NF = 46
};

constexpr double kPi = 3.141592653589793;
// Constants of src/FLOWVPM.jl:62-66, as the exact doubles Julia computes.
constexpr double kConst1 = 0x1.0411e71d77959p-4;    // 1/(2*pi)^1.5
constexpr double kConst2 = 0x1.9884533d43651p-1;    // sqrt(2/pi)
constexpr double kConst3 = 0x1.e8ec8a4aeacc4p-3;    // 3/(4*pi)
constexpr double kConst4 = 0x1.45f306dc9c883p-4;    // 1/(4*pi)
constexpr double kSqr2 = 0x1.6a09e667f3bcdp+0;      // sqrt(2)

constexpr int kBlock = 256;      // threads per block (4 waves of 64)
constexpr int kTile = 256;       // sources per LDS tile
constexpr int kAccF32 = 14;      // U'(3) W'(3) Jhv(8) per target, F32C path
constexpr int kAccF64 = 12;      // U(3) J(9) per target, F64 path
// internal precision of the FP64 far-form U/J kernel (vpm_uj_direct with
// VPM_PREC_F64 and VPM_UJ_F64_FAR)
constexpr int kPrecF64Far = 2;

// Grow-only device buffer owned by a field.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
};

struct Timing {
    double ms = 0.0;
    int64_t count = 0;
};

// Device memory handed out in order from one caller-provided block (the
// no-alloc entry point's static array), 256-byte aligned.  take() returns
// nullptr once the block is exhausted.
struct Arena {
    char* base = nullptr;
    size_t size = 0;
    size_t used = 0;
    void* take(size_t bytes)
    {
        const size_t at = (used + 255) & ~(size_t)255;
        if (at > size || bytes > size - at) return nullptr;
        used = at + bytes;
        return base + at;
    }
};

}  // namespace vpm

struct vpm_field_s {
    int device = 0;
    int64_t cap = 0;
    int64_t np = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    double* P = nullptr;             // cap x 46 particle records (FP64 master state)
    // staging buffers (sorted or storage order), grown on demand
    vpm::DevBuf pos;                 // double [3][cap]  positions, SoA
    vpm::DevBuf gs;                  // float4 [cap]     (-Gamma/(4 pi), 1/sigma)
    vpm::DevBuf cut;                 // float  [cap]     far-field r^2 threshold
    vpm::DevBuf g64;                 // double4 [cap]    (Gamma, sigma)
    vpm::DevBuf part;                // double partial sums [K][acc][ld]
    vpm::DevBuf keys, keys2, perm, iota, sorttmp;
    bool perm_valid = false;         // perm matches the current positions
    vpm::DevBuf bbox;                // double [6 + partials]
    vpm::DevBuf flag;                // int device error flag
    vpm::DevBuf adj;                 // adjoint / E_str staging scratch
    vpm::DevBuf rng;                 // leaf-pair range lists
    vpm::DevBuf bar;                 // cotangent records of the step adjoint
    vpm::DevBuf xfer;                // gathered rows for partial downloads
    vpm::DevBuf lrec;                // E_str / zeta list-kernel source records
    vpm::DevBuf hstage;              // pinned host staging of range plans (hipHostMalloc)
    bool profiling = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    vpm::Timing timing[5];      // U/J F32C, U/J F64, E_str, adjoint, zeta
    std::vector<vpm_tape> tapes;     // live tapes of this field (released with it)
    std::vector<vpm_plan> plans;     // live range plans of this field (released with it)
};

// Stage records of one rk3 step (state after each stage's UJ evaluation),
// the input of vpm_rk3_step_adjoint.
struct vpm_tape_s {
    vpm_field f = nullptr;
    int64_t np = 0;
    vpm::DevBuf rec;                 // double [3][np][46]
    vpm_step_config s{};
    double dt = 0.0;
    bool recorded = false;
};

namespace vpm {
// A range plan on the device: items, then source ranges (int2 = SrcRange).
struct DevPlan {
    const RangeItem* items = nullptr;
    const int2* src = nullptr;
    int64_t nitems = 0;
    bool wave = false;   // F32C wave kernels (FMM leaves)
    int R = 1;           // lane groups per target of the wave kernels
};
}  // namespace vpm

// A direct list planned once for a field (vpm_plan_ranges), served to the
// U/J, E_str and zeta list kernels.  Valid while the field holds the np it
// was made for.
struct vpm_plan_s {
    vpm_field f = nullptr;
    int64_t np = 0;
    vpm::DevBuf mem;                 // device items + source ranges
    vpm::DevPlan dp;
};

namespace vpm {

// thread-local error text, set by fail()
int fail(int status, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
// ensure a grow-only buffer holds at least `bytes`
int ensure(DevBuf& b, size_t bytes);
// the same for pinned host memory; the caller has synchronized the field's
// stream since the buffer's last use as a copy source
int ensure_pinned(DevBuf& b, size_t bytes);

// Launchers (stream-ordered on f->stream).  Return a vpm_status.
struct Staged {
    const double* pos;   // SoA positions, stride ld
    int64_t ld;
    const float4* gs;
    const float* cut;
    const double4* g64;
    const uint32_t* perm;  // sorted index -> particle index, or nullptr
};

// With an Arena, the scratch of these calls (sort keys and permutation,
// staged sources, partial sums) is taken from it instead of the field's
// grow-only buffers, nothing is allocated, and the permutation is not kept
// for later calls.  launch_uj's kmax > 0 caps the number of source splits.
int sort_field(vpm_field f, const uint32_t*& perm, Arena* ar = nullptr);
int stage_field(vpm_field f, int kernel, int precision, bool sort, Staged& out, Arena* ar = nullptr);
int launch_uj(vpm_field ft, const Staged& T, int64_t t0, int64_t nt,
              const Staged& S, int64_t s0, int64_t s1, int kernel, int precision,
              Arena* ar = nullptr, int64_t kmax = 0);
// Source splits of an all-pairs launch over nt targets and ns sources (at
// most kmax when kmax > 0): K chunks of `chunk` sources, a multiple of the tile.
void choose_split(int64_t nt, int64_t ns, int64_t kmax, int64_t& K, int64_t& chunk);
// Scratch bytes of sort_field (temp storage of the radix sort included).
int sort_scratch_bytes(vpm_field f, size_t& bytes);
int launch_estr_impl(vpm_field f, int kernel, int transposed, int precision, const uint32_t* perm);
// E_str in FP64 with a far cut, sources in spatial order (VPM_UJ_F64_FAR)
int launch_estr_f64far(vpm_field f, int kernel, int transposed);
int launch_reset(vpm_field f, int which);
int launch_update_states(vpm_field f, double a, double b, double dt, const double Uinf[3],
                         double fr, double gr, double zeta0, int transposed);
int launch_relax(vpm_field f, double rlxf, int corrected);
int launch_core_spreading(vpm_field f, double nu, double a, double b, double dt);
int launch_scale_sigma(vpm_field f, double alpha, int divide);
int launch_sfs_store_test(vpm_field f, int transposed);
int launch_sfs_coefficient(vpm_field f, int transposed, double alpha, double rlxf, double minC,
                           double maxC, double zeta0, int force_positive, int clip, int* nan_flag);
int launch_update_uprev(vpm_field f);
int launch_p2p_ranges(vpm_field f, int kernel, int precision, const int64_t* tb, const int64_t* te,
                      const int64_t* sb, const int64_t* se, int64_t npairs);
// Plans a pair list (range_plan.h) through the field's pinned staging buffer
// into the device buffer dst; the field's stream must be synchronized
// before the staging buffer is used again.
int make_dev_plan(vpm_field f, const int64_t* tb, const int64_t* te, const int64_t* sb, const int64_t* se,
                  int64_t npairs, bool allow_wave, int rmax, DevBuf& dst, DevPlan& out);
int launch_p2p_plan(vpm_field f, int kernel, int precision, const DevPlan& dp);
int launch_estr_plan(vpm_field f, int kernel, int precision, int transposed, const DevPlan& dp);
int launch_zeta_plan(vpm_field f, int kernel, int precision, const DevPlan& dp);
int launch_uj_adjoint(vpm_field f, int kernel, int precision, const double* Ubar, const double* Jbar,
                      double* Xbar, double* Gbar, double* Sbar);
// device-side U/J adjoint staging (in f->adj) and launch
struct AdjBufs {
    double *U, *J, *X, *G, *S;
    float4* cot;
    bool jbar = true;   // some Jbar may be nonzero (false: the loss is on U alone)
};
int adj_buffers(vpm_field f, AdjBufs& b);
int launch_uj_adjoint_dev(vpm_field f, int kernel, int precision, const AdjBufs& b);
int launch_update_states_adjoint(vpm_field f, double* B, double a, double b, double dt, double fr, double gr,
                                 int transposed, double* Ub, double* Jb);
int launch_adj_add_bar(vpm_field f, double* B, const double* dX, const double* dG, const double* dS);
int launch_adj_step_start(vpm_field f, double* B);
int launch_zeta(vpm_field f, int kernel, int precision);
int launch_core_spreading_reset(vpm_field f, int kernel, int precision, double sgm0, int itmax, double tol,
                                int iterror, int* iters, double res[3]);
int launch_rbf_cg(vpm_field f, int kernel, int precision, int itmax, double tol, int iterror, int* iters,
                  double res[3]);
int launch_sfs_delta(vpm_field f, int kernel, int transposed, double alpha);
int launch_gather_rows(vpm_field f, int row0, int nrows, double* out);
int launch_zeta_ranges(vpm_field f, int kernel, int precision, const int64_t* tb, const int64_t* te,
                       const int64_t* sb, const int64_t* se, int64_t npairs);
int launch_estr_ranges(vpm_field f, int kernel, int precision, int transposed, const int64_t* tb,
                       const int64_t* te, const int64_t* sb, const int64_t* se, int64_t npairs);

double zeta0_of(int kernel);

// kernel timing helpers
void timing_begin(vpm_field f);
void timing_end(vpm_field f, int which);

}  // namespace vpm
