// vpm_abi.hip -- the extern "C" entry points of include/flowvpm_hip.h.
//
// Host-side orchestration only: argument checks, device buffers, and the
// order of launches on the field's stream.  The numerical work is in
// p2p_uj.hip (U/J), estr.hip (E_str), particle_ops.hip (per-particle FP64),
// p2p_ranges.hip (leaf-pair lists) and adjoint.hip.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "vpm_internal.h"

namespace vpm {

static thread_local std::string g_last_error;

int fail(int status, const std::string& msg)
{
    g_last_error = msg;
    return status;
}

int hip_fail(hipError_t e, const char* what)
{
    g_last_error = std::string(what) + ": " + hipGetErrorString(e);
    return e == hipErrorOutOfMemory ? VPM_ENOMEM : VPM_EDEVICE;
}

int ensure(DevBuf& b, size_t bytes)
{
    if (b.bytes >= bytes) return VPM_OK;
    if (b.p) {
        hipError_t e = hipFree(b.p);   // synchronizes with pending work
        b.p = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) return hip_fail(e, "hipFree");
    }
    size_t want = std::max<size_t>(bytes, 256);
    hipError_t e = hipMalloc(&b.p, want);
    if (e != hipSuccess) {
        b.p = nullptr;
        return hip_fail(e, "hipMalloc");
    }
    b.bytes = want;
    return VPM_OK;
}

int ensure_pinned(DevBuf& b, size_t bytes)
{
    if (b.bytes >= bytes) return VPM_OK;
    const size_t had = b.bytes;
    if (b.p) {
        hipError_t e = hipHostFree(b.p);
        b.p = nullptr;
        b.bytes = 0;
        if (e != hipSuccess) return hip_fail(e, "hipHostFree");
    }
    // grow by at least a half, so that lists growing step by step reallocate rarely
    size_t want = std::max<size_t>(bytes, 1 << 16);
    want = std::max(want, had + had / 2);
    hipError_t e = hipHostMalloc(&b.p, want, hipHostMallocDefault);
    if (e != hipSuccess) {
        b.p = nullptr;
        return hip_fail(e, "hipHostMalloc");
    }
    b.bytes = want;
    return VPM_OK;
}

static void release(DevBuf& b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.bytes = 0;
}

void timing_begin(vpm_field f)
{
    if (f->profiling) (void)hipEventRecord(f->ev0, f->stream);
}

void timing_end(vpm_field f, int which)
{
    if (!f->profiling) return;
    (void)hipEventRecord(f->ev1, f->stream);
    if (hipEventSynchronize(f->ev1) == hipSuccess) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, f->ev0, f->ev1) == hipSuccess) {
            f->timing[which].ms += ms;
            f->timing[which].count += 1;
        }
    }
}

double zeta0_of(int kernel)
{
    // kernel.zeta(0) of src/FLOWVPM_kernel.jl:45,51,60,69-74
    switch (kernel) {
    case VPM_KERNEL_SINGULAR: return 1.0;
    case VPM_KERNEL_GAUSSIAN: return kConst3;
    case VPM_KERNEL_GAUSSIANERF: return kConst1;
    case VPM_KERNEL_WINCKELMANS: return kConst4 * 7.5;
    }
    return NAN;
}

static int check_cfg(const vpm_config* cfg)
{
    if (!cfg) return fail(VPM_EINVAL, "null vpm_config");
    if (cfg->kernel < VPM_KERNEL_SINGULAR || cfg->kernel > VPM_KERNEL_WINCKELMANS)
        return fail(VPM_EINVAL, "unknown kernel id " + std::to_string(cfg->kernel));
    if (cfg->precision != VPM_PREC_F32C && cfg->precision != VPM_PREC_F64)
        return fail(VPM_EINVAL, "unknown precision " + std::to_string(cfg->precision));
    return VPM_OK;
}

static int set_device(vpm_field f)
{
    hipError_t e = hipSetDevice(f->device);
    return e == hipSuccess ? VPM_OK : hip_fail(e, "hipSetDevice");
}

static int sync(vpm_field f)
{
    hipError_t e = hipStreamSynchronize(f->stream);
    return e == hipSuccess ? VPM_OK : hip_fail(e, "hipStreamSynchronize");
}

// UJ_direct body on a locked field (src/FLOWVPM_UJ.jl:21-37)
static int uj_direct_locked(vpm_field f, const vpm_config* cfg, uint32_t flags)
{
    int rc;
    const bool far64 = flags & VPM_UJ_F64_FAR;
    if (far64 && cfg->precision != VPM_PREC_F64)
        return fail(VPM_EINVAL, "VPM_UJ_F64_FAR needs VPM_PREC_F64");
    if (flags & VPM_UJ_RESET) if ((rc = launch_reset(f, 0))) return rc;
    if (flags & VPM_UJ_RESET_SFS) if ((rc = launch_reset(f, 1))) return rc;
    const int prec = far64 ? kPrecF64Far : cfg->precision;
    const bool sort = (cfg->precision == VPM_PREC_F32C || far64) && !(flags & VPM_UJ_NO_SORT);
    Staged st;
    if ((rc = stage_field(f, cfg->kernel, prec, sort, st))) return rc;
    if ((rc = launch_uj(f, st, 0, f->np, st, 0, f->np, cfg->kernel, prec))) return rc;
    if (flags & VPM_UJ_SFS) {
        if (far64) return launch_estr_f64far(f, cfg->kernel, cfg->transposed);
        const uint32_t* perm = nullptr;
        if (cfg->precision == VPM_PREC_F32C && (rc = sort_field(f, perm))) return rc;
        if ((rc = launch_estr_impl(f, cfg->kernel, cfg->transposed, cfg->precision, perm))) return rc;
    }
    return VPM_OK;
}

// uj_flags: VPM_UJ_F64_FAR or 0, as the step's own U/J evaluations
static int sfs_before_locked(vpm_field f, const vpm_config* cfg, double alpha, double minC, double maxC,
                             uint32_t uj_flags = 0)
{
    if (minC < 0 || maxC < 0 || minC > maxC)                         // subfilterscale.jl:455-461
        return fail(VPM_EINVAL, "Invalid C bounds");
    int rc;
    if ((rc = launch_scale_sigma(f, alpha, 0))) return rc;
    if ((rc = uj_direct_locked(f, cfg, VPM_UJ_RESET | VPM_UJ_RESET_SFS | VPM_UJ_SFS | uj_flags))) return rc;
    if ((rc = launch_sfs_store_test(f, cfg->transposed))) return rc;
    return launch_scale_sigma(f, alpha, 1);
}

static int sfs_after_locked(vpm_field f, const vpm_config* cfg, double alpha, double rlxf, double minC,
                            double maxC, int force_positive, int clip)
{
    if (minC < 0 || maxC < 0 || minC > maxC) return fail(VPM_EINVAL, "Invalid C bounds");
    int nan_flag = 0;
    int rc = launch_sfs_coefficient(f, cfg->transposed, alpha, rlxf, minC, maxC, zeta0_of(cfg->kernel),
                                    force_positive, clip, &nan_flag);
    if (rc) return rc;
    if (nan_flag) return fail(VPM_ENAN, "NaN in dynamicprocedure_pseudo3level_afterUJ");
    return VPM_OK;
}

}  // namespace vpm

using namespace vpm;

#define VPM_FIELD_GUARD(f)                                                     \
    if (!(f)) return fail(VPM_EINVAL, "null vpm_field");                       \
    std::lock_guard<std::mutex> lock_((f)->mu);                                \
    if (int rc_ = set_device(f)) return rc_;

extern "C" {

const char* vpm_version(void) { return "flowvpm_hip 0.1.0 (gfx950, abi 1)"; }

const char* vpm_last_error(void) { return g_last_error.c_str(); }

int vpm_device_count(int* count)
{
    if (!count) return fail(VPM_EINVAL, "null count");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) {
        *count = 0;
        return hip_fail(e, "hipGetDeviceCount");
    }
    *count = n;
    return VPM_OK;
}

int vpm_create(int device, int64_t maxparticles, vpm_field* out)
{
    if (!out) return fail(VPM_EINVAL, "null output handle");
    *out = nullptr;
    // indices are 32-bit on the device (sort permutation, range lists);
    // 2^31 records would take 790 GB, beyond one GPU's memory
    if (maxparticles < 1 || maxparticles > (int64_t)INT32_MAX)
        return fail(VPM_EINVAL, "maxparticles must be in [1, 2^31)");
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) return hip_fail(e, "hipGetDeviceCount");
    if (device < 0 || device >= n) return fail(VPM_EDEVICE, "no such device " + std::to_string(device));
    vpm_field f = new (std::nothrow) vpm_field_s();
    if (!f) return fail(VPM_ENOMEM, "host allocation failed");
    f->device = device;
    f->cap = maxparticles;
    if ((e = hipSetDevice(device)) != hipSuccess) { delete f; return hip_fail(e, "hipSetDevice"); }
    if ((e = hipStreamCreateWithFlags(&f->stream, hipStreamNonBlocking)) != hipSuccess) {
        delete f;
        return hip_fail(e, "hipStreamCreateWithFlags");
    }
    if ((e = hipMalloc(&f->P, sizeof(double) * NF * (size_t)maxparticles)) != hipSuccess) {
        (void)hipStreamDestroy(f->stream);
        delete f;
        return hip_fail(e, "hipMalloc(particles)");
    }
    // the zero fill goes on the field's own stream and is waited for: the
    // stream is non-blocking, so a fill on the null stream would not be
    // ordered before the first upload and could land on top of it
    if ((e = hipMemsetAsync(f->P, 0, sizeof(double) * NF * (size_t)maxparticles, f->stream)) != hipSuccess ||
        (e = hipStreamSynchronize(f->stream)) != hipSuccess || (e = hipEventCreate(&f->ev0)) != hipSuccess ||
        (e = hipEventCreate(&f->ev1)) != hipSuccess) {
        int rc = hip_fail(e, "field setup");
        (void)hipFree(f->P);
        (void)hipStreamDestroy(f->stream);
        delete f;
        return rc;
    }
    *out = f;
    return VPM_OK;
}

int vpm_destroy(vpm_field f)
{
    if (!f) return VPM_OK;
    {
        std::lock_guard<std::mutex> lock(f->mu);
        (void)hipSetDevice(f->device);
        (void)hipStreamSynchronize(f->stream);
        DevBuf* bufs[] = {&f->pos, &f->gs, &f->cut, &f->g64, &f->part, &f->keys, &f->keys2, &f->perm,
                          &f->iota, &f->sorttmp, &f->bbox, &f->flag, &f->adj, &f->rng, &f->bar, &f->xfer,
                          &f->lrec};
        for (DevBuf* b : bufs) release(*b);
        if (f->hstage.p) (void)hipHostFree(f->hstage.p);
        f->hstage.p = nullptr;
        f->hstage.bytes = 0;
        for (vpm_tape t : f->tapes) {   // detach: the tape stays valid only for vpm_tape_destroy
            release(t->rec);
            t->f = nullptr;
            t->recorded = false;
        }
        f->tapes.clear();
        for (vpm_plan pl : f->plans) {   // detach: the plan stays valid only for vpm_plan_destroy
            release(pl->mem);
            pl->f = nullptr;
            pl->dp = DevPlan();
        }
        f->plans.clear();
        if (f->P) (void)hipFree(f->P);
        if (f->ev0) (void)hipEventDestroy(f->ev0);
        if (f->ev1) (void)hipEventDestroy(f->ev1);
        if (f->stream) (void)hipStreamDestroy(f->stream);
    }
    delete f;
    return VPM_OK;
}

int vpm_set_particles(vpm_field f, const double* particles, int64_t np)
{
    VPM_FIELD_GUARD(f);
    if (np < 0 || (np > 0 && !particles)) return fail(VPM_EINVAL, "bad particle array");
    if (np > f->cap) return fail(VPM_ERANGE, "PARTICLE OVERFLOW: " + std::to_string(np) + " > maxparticles " +
                                                  std::to_string(f->cap));
    for (int64_t i = 0; i < np; ++i) {
        const double* p = particles + i * NF;
        for (int k = 0; k < 7; ++k)
            if (!std::isfinite(p[k]))
                return fail(VPM_EINVAL, "non-finite X/Gamma/sigma at particle " + std::to_string(i));
        if (!(p[RSIG] > 0)) return fail(VPM_EINVAL, "sigma <= 0 at particle " + std::to_string(i));
    }
    if (np > 0) {
        hipError_t e = hipMemcpyAsync(f->P, particles, sizeof(double) * NF * (size_t)np, hipMemcpyHostToDevice,
                                      f->stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(H2D)");
    }
    f->np = np;
    f->perm_valid = false;
    return sync(f);
}

int vpm_add_particles(vpm_field f, const double* particles, int64_t n)
{
    VPM_FIELD_GUARD(f);
    if (n < 0 || (n > 0 && !particles)) return fail(VPM_EINVAL, "bad particle array");
    if (f->np + n > f->cap)
        return fail(VPM_ERANGE, "PARTICLE OVERFLOW. Max number of particles " + std::to_string(f->cap) +
                                    " has been reached");
    for (int64_t i = 0; i < n; ++i) {
        const double* p = particles + i * NF;
        for (int k = 0; k < 7; ++k)
            if (!std::isfinite(p[k]))
                return fail(VPM_EINVAL, "non-finite X/Gamma/sigma at particle " + std::to_string(i));
        if (!(p[RSIG] > 0)) return fail(VPM_EINVAL, "sigma <= 0 at particle " + std::to_string(i));
    }
    if (n == 0) return VPM_OK;
    hipError_t e = hipMemcpyAsync(f->P + f->np * NF, particles, sizeof(double) * NF * (size_t)n,
                                  hipMemcpyHostToDevice, f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(H2D)");
    f->np += n;
    f->perm_valid = false;
    return sync(f);
}

int vpm_remove_particle(vpm_field f, int64_t i)
{
    VPM_FIELD_GUARD(f);
    if (i < 0) return fail(VPM_EINVAL, "Requested removal of invalid particle index " + std::to_string(i));
    if (i >= f->np)
        return fail(VPM_EINVAL, "Requested removal of particle " + std::to_string(i) + ", but there is only " +
                                    std::to_string(f->np) + " particles in the field.");
    const int64_t last = f->np - 1;
    hipError_t e;
    if (i != last) {   // particlefield.jl:409-412
        e = hipMemcpyAsync(f->P + i * NF, f->P + last * NF, sizeof(double) * NF, hipMemcpyDeviceToDevice,
                           f->stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(D2D)");
    }
    // _reset_particle (particlefield.jl:485-490): U, vorticity, J, PSE rows
    e = hipMemsetAsync(f->P + last * NF + RU, 0, sizeof(double) * (RM - RU), f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync");
    f->np = last;
    f->perm_valid = false;
    return sync(f);
}

int vpm_get_particles(vpm_field f, double* particles, int64_t np, int32_t row0, int32_t row1)
{
    VPM_FIELD_GUARD(f);
    if (np < 0 || np > f->np) return fail(VPM_ERANGE, "np exceeds the field's particle count");
    if (row0 < 0 || row1 > NF || row0 >= row1) return fail(VPM_EINVAL, "bad row range");
    if (np == 0) return VPM_OK;
    if (!particles) return fail(VPM_EINVAL, "null output array");
    hipError_t e;
    if (row0 == 0 && row1 == NF) {   // whole records: one contiguous copy
        e = hipMemcpyAsync(particles, f->P, sizeof(double) * NF * (size_t)np, hipMemcpyDeviceToHost, f->stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(D2H)");
        return sync(f);
    }
    // a row range: gather on the device, one contiguous copy, scatter on the
    // host (a strided 2D copy of short rows is orders of magnitude slower)
    const int nr = row1 - row0;
    const size_t n = (size_t)np * (size_t)nr;
    if (int rc = ensure(f->xfer, sizeof(double) * n)) return rc;
    if (int rc = launch_gather_rows(f, row0, nr, (double*)f->xfer.p)) return rc;
    std::vector<double> host;
    try {
        host.resize(n);
    } catch (const std::exception&) {
        return fail(VPM_ENOMEM, "host buffer of vpm_get_particles");
    }
    e = hipMemcpyAsync(host.data(), f->xfer.p, sizeof(double) * n, hipMemcpyDeviceToHost, f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(D2H)");
    if (int rc = sync(f)) return rc;
    for (int64_t i = 0; i < np; ++i)
        std::memcpy(particles + i * NF + row0, host.data() + (size_t)i * nr, sizeof(double) * nr);
    return VPM_OK;
}

int vpm_get_np(vpm_field f, int64_t* np)
{
    if (!f || !np) return fail(VPM_EINVAL, "null argument");
    *np = f->np;
    return VPM_OK;
}

int vpm_synchronize(vpm_field f)
{
    VPM_FIELD_GUARD(f);
    return sync(f);
}

int vpm_uj_direct(vpm_field f, const vpm_config* cfg, uint32_t flags)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (int rc = uj_direct_locked(f, cfg, flags)) return rc;
    return sync(f);
}

int vpm_uj_direct_fields(vpm_field target, vpm_field source, const vpm_config* cfg)
{
    if (!target || !source) return fail(VPM_EINVAL, "null vpm_field");
    if (int rc = check_cfg(cfg)) return rc;
    if (target == source) {
        std::lock_guard<std::mutex> lock(target->mu);
        if (int rc = set_device(target)) return rc;
        Staged st;
        const bool sort = cfg->precision == VPM_PREC_F32C;
        if (int rc = stage_field(target, cfg->kernel, cfg->precision, sort, st)) return rc;
        if (int rc = launch_uj(target, st, 0, target->np, st, 0, target->np, cfg->kernel, cfg->precision)) return rc;
        return sync(target);
    }
    if (target->device != source->device) return fail(VPM_EINVAL, "fields live on different devices");
    // lock in address order to avoid lock inversion between two callers
    vpm_field a = target < source ? target : source, b = target < source ? source : target;
    std::lock_guard<std::mutex> la(a->mu), lb(b->mu);
    if (int rc = set_device(target)) return rc;
    const bool sort = cfg->precision == VPM_PREC_F32C;
    Staged ss, ts;
    if (int rc = stage_field(source, cfg->kernel, cfg->precision, sort, ss)) return rc;
    if (int rc = sync(source)) return rc;          // source staging is on the source stream
    if (int rc = stage_field(target, cfg->kernel, cfg->precision, sort, ts)) return rc;
    if (int rc = launch_uj(target, ts, 0, target->np, ss, 0, source->np, cfg->kernel, cfg->precision)) return rc;
    return sync(target);
}

int vpm_p2p_ranges(vpm_field f, const vpm_config* cfg, const int64_t* tb, const int64_t* te, const int64_t* sb,
                   const int64_t* se, int64_t npairs)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (npairs < 0) return fail(VPM_EINVAL, "npairs < 0");
    if (npairs == 0) return VPM_OK;
    if (!tb || !te || !sb || !se) return fail(VPM_EINVAL, "null range array");
    // each pair is checked against [0, np) while its plan is built
    // synchronized on every path: the plan's pinned staging buffer may be the
    // source of a copy still in flight
    const int rc = launch_p2p_ranges(f, cfg->kernel, cfg->precision, tb, te, sb, se, npairs);
    const int rs = sync(f);
    return rc ? rc : rs;
}

int vpm_estr_ranges(vpm_field f, const vpm_config* cfg, const int64_t* tb, const int64_t* te, const int64_t* sb,
                    const int64_t* se, int64_t npairs)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (npairs < 0) return fail(VPM_EINVAL, "npairs < 0");
    if (npairs == 0) return VPM_OK;
    if (!tb || !te || !sb || !se) return fail(VPM_EINVAL, "null range array");
    // each pair is checked against [0, np) while its plan is built
    const int rc = launch_estr_ranges(f, cfg->kernel, cfg->precision, cfg->transposed, tb, te, sb, se, npairs);
    const int rs = sync(f);
    return rc ? rc : rs;
}

int vpm_zeta_ranges(vpm_field f, const vpm_config* cfg, const int64_t* tb, const int64_t* te, const int64_t* sb,
                    const int64_t* se, int64_t npairs)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (npairs < 0) return fail(VPM_EINVAL, "npairs < 0");
    if (npairs == 0) return VPM_OK;
    if (!tb || !te || !sb || !se) return fail(VPM_EINVAL, "null range array");
    // each pair is checked against [0, np) while its plan is built
    const int rc = launch_zeta_ranges(f, cfg->kernel, cfg->precision, tb, te, sb, se, npairs);
    const int rs = sync(f);
    return rc ? rc : rs;
}

// ---- planned direct lists ----
int vpm_plan_ranges(vpm_field f, const int64_t* tb, const int64_t* te, const int64_t* sb, const int64_t* se,
                    int64_t npairs, vpm_plan* out)
{
    if (!out) return fail(VPM_EINVAL, "null output pointer");
    *out = nullptr;
    VPM_FIELD_GUARD(f);
    if (npairs < 0) return fail(VPM_EINVAL, "npairs < 0");
    if (npairs > 0 && (!tb || !te || !sb || !se)) return fail(VPM_EINVAL, "null range array");
    vpm_plan p = new (std::nothrow) vpm_plan_s;
    if (!p) return fail(VPM_ENOMEM, "vpm_plan_ranges");
    p->f = f;
    p->np = f->np;
    // the F32C leaf form (wave items, lane sharing up to 8); the FP64 block
    // kernels run over the same items
    int rc = npairs > 0 ? make_dev_plan(f, tb, te, sb, se, npairs, true, 8, p->mem, p->dp) : VPM_OK;
    const int rs = sync(f);   // the pinned staging buffer is free again
    if (!rc) rc = rs;
    if (!rc) {
        try {
            f->plans.push_back(p);
        } catch (const std::exception&) {
            rc = fail(VPM_ENOMEM, "vpm_plan_ranges");
        }
    }
    if (rc) {
        release(p->mem);
        delete p;
        return rc;
    }
    *out = p;
    return VPM_OK;
}

int vpm_plan_destroy(vpm_plan p)
{
    if (!p) return VPM_OK;
    if (vpm_field f = p->f) {   // null once the field was destroyed (mem already released)
        std::lock_guard<std::mutex> lock(f->mu);
        if (p->mem.p) {
            if (int rc = set_device(f)) return rc;
            (void)hipStreamSynchronize(f->stream);
            release(p->mem);
        }
        f->plans.erase(std::remove(f->plans.begin(), f->plans.end(), p), f->plans.end());
    }
    delete p;
    return VPM_OK;
}

static int check_plan(vpm_field f, vpm_plan p)
{
    if (!p) return fail(VPM_EINVAL, "null vpm_plan");
    if (p->f != f) return fail(VPM_EINVAL, p->f ? "plan belongs to another field" : "the plan's field was destroyed");
    if (p->np != f->np)
        return fail(VPM_EINVAL, "plan was made for np = " + std::to_string(p->np) + ", the field holds " +
                                    std::to_string(f->np));
    return VPM_OK;
}

int vpm_p2p_plan(vpm_field f, const vpm_config* cfg, vpm_plan p)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (int rc = check_plan(f, p)) return rc;
    const int rc = launch_p2p_plan(f, cfg->kernel, cfg->precision, p->dp);
    const int rs = sync(f);
    return rc ? rc : rs;
}

int vpm_estr_plan(vpm_field f, const vpm_config* cfg, vpm_plan p)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (int rc = check_plan(f, p)) return rc;
    const int rc = launch_estr_plan(f, cfg->kernel, cfg->precision, cfg->transposed, p->dp);
    const int rs = sync(f);
    return rc ? rc : rs;
}

int vpm_zeta_plan(vpm_field f, const vpm_config* cfg, vpm_plan p)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (int rc = check_plan(f, p)) return rc;
    const int rc = launch_zeta_plan(f, cfg->kernel, cfg->precision, p->dp);
    const int rs = sync(f);
    return rc ? rc : rs;
}

int vpm_estr_direct(vpm_field f, const vpm_config* cfg)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    const uint32_t* perm = nullptr;
    if (cfg->precision == VPM_PREC_F32C)
        if (int rc = sort_field(f, perm)) return rc;
    if (int rc = launch_estr_impl(f, cfg->kernel, cfg->transposed, cfg->precision, perm)) return rc;
    return sync(f);
}

int vpm_reset_particles(vpm_field f)
{
    VPM_FIELD_GUARD(f);
    if (int rc = launch_reset(f, 0)) return rc;
    return sync(f);
}

int vpm_reset_particles_sfs(vpm_field f)
{
    VPM_FIELD_GUARD(f);
    if (int rc = launch_reset(f, 1)) return rc;
    return sync(f);
}

int vpm_update_particle_states(vpm_field f, double a, double b, double dt, const double Uinf[3], double fr,
                               double gr, double zeta0, int32_t transposed)
{
    VPM_FIELD_GUARD(f);
    if (!Uinf) return fail(VPM_EINVAL, "null Uinf");
    if (int rc = launch_update_states(f, a, b, dt, Uinf, fr, gr, zeta0, transposed)) return rc;
    return sync(f);
}

int vpm_relax(vpm_field f, double rlxf, int32_t corrected)
{
    VPM_FIELD_GUARD(f);
    if (int rc = launch_relax(f, rlxf, corrected)) return rc;
    return sync(f);
}

int vpm_core_spreading(vpm_field f, double nu, double a, double b, double dt)
{
    VPM_FIELD_GUARD(f);
    if (int rc = launch_core_spreading(f, nu, a, b, dt)) return rc;
    return sync(f);
}

int vpm_sfs_dynamic_before(vpm_field f, const vpm_config* cfg, double alpha, double minC, double maxC)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (int rc = sfs_before_locked(f, cfg, alpha, minC, maxC)) return rc;
    return sync(f);
}

int vpm_sfs_dynamic_after(vpm_field f, const vpm_config* cfg, double alpha, double rlxf, double minC, double maxC,
                          int32_t force_positive, int32_t clip)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    return sfs_after_locked(f, cfg, alpha, rlxf, minC, maxC, force_positive, clip);
}

}  // extern "C"

// One rk3 step; with a tape, the state after each stage's UJ evaluation is
// copied into the tape (the per-stage checkpoints of the step adjoint).
static int rk3_step_locked(vpm_field f, vpm_step_config* s, double dt, int32_t relax_now, vpm_tape tape)
{
    const vpm_config* cfg = &s->uj;
    if (int rc = check_cfg(cfg)) return rc;
    if ((s->uj_flags & ~(int32_t)VPM_UJ_F64_FAR) != 0)
        return fail(VPM_EINVAL, "vpm_step_config.uj_flags: only VPM_UJ_F64_FAR is taken");
    if (s->uj_flags && cfg->precision != VPM_PREC_F64) return fail(VPM_EINVAL, "VPM_UJ_F64_FAR needs VPM_PREC_F64");
    if (s->viscous == 1 && cfg->kernel != VPM_KERNEL_GAUSSIANERF)   // FLOWVPM.jl:265-267, utils.jl:58-64
        return fail(VPM_EINVAL, "CoreSpreading requires the Gaussian-erf kernel");
    const double zeta0 = zeta0_of(cfg->kernel);                    // timeintegration.jl:398
    const bool sfs = s->sfs != 0;
    int rc;
    bool reset_needed = false, unconverged = false;
    const size_t rec = sizeof(double) * NF * (size_t)f->np;
    if (tape) {
        if (sfs || s->viscous || (relax_now && s->relax))
            return fail(VPM_EINVAL, "taped rk3 step: SFS, viscous diffusion and relaxation are not differentiated");
        if ((rc = ensure(tape->rec, 3 * rec + 1))) return rc;
        tape->recorded = false;
    }
    if ((rc = launch_reset(f, 2))) return rc;                      // :401-414
    const double ab[3][2] = {{0.0, 1.0 / 3.0}, {-5.0 / 9.0, 15.0 / 16.0}, {-153.0 / 128.0, 8.0 / 15.0}};
    for (int st = 0; st < 3; ++st) {
        const double a = ab[st][0], b = ab[st][1];
        const bool first = (a == 1 || a == 0);                      // subfilterscale.jl:207,219
        // F32C with a two-level test filter: the test-minus-domain
        // differences come from the difference kernels (sfs_delta.hip)
        // instead of a second U/J + E_str pass at alpha sigma
        const bool sfs_diff = sfs && first && cfg->precision == VPM_PREC_F32C &&
                              cfg->kernel != VPM_KERNEL_SINGULAR && 1.0 - s->sfs_alpha <= 0.01 &&
                              s->sfs_alpha < 1.0;
        if (sfs && first && !sfs_diff)
            if ((rc = sfs_before_locked(f, cfg, s->sfs_alpha, s->sfs_minC, s->sfs_maxC, (uint32_t)s->uj_flags)))
                return rc;
        const uint32_t fl = VPM_UJ_RESET | (sfs ? (VPM_UJ_RESET_SFS | VPM_UJ_SFS) : 0u) | (uint32_t)s->uj_flags;
        if ((rc = uj_direct_locked(f, cfg, fl))) return rc;        // :422
        if (tape && f->np > 0) {
            hipError_t e = hipMemcpyAsync((char*)tape->rec.p + st * rec, f->P, rec, hipMemcpyDeviceToDevice,
                                          f->stream);
            if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(tape)");
        }
        if (sfs_diff) {
            if (s->sfs_minC < 0 || s->sfs_maxC < 0 || s->sfs_minC > s->sfs_maxC)   // subfilterscale.jl:455-461
                return fail(VPM_EINVAL, "Invalid C bounds");
            if ((rc = launch_sfs_delta(f, cfg->kernel, cfg->transposed, s->sfs_alpha))) return rc;
        }
        if (sfs && first)
            if ((rc = sfs_after_locked(f, cfg, s->sfs_alpha, s->sfs_rlxf, s->sfs_minC, s->sfs_maxC,
                                       s->sfs_force_positive, s->sfs_clip)))
                return rc;
        if ((rc = launch_update_states(f, a, b, dt, s->Uinf, s->f, s->g, zeta0, cfg->transposed))) return rc;
        if (s->viscous == 1) {                                      // viscous.jl:167-221
            if ((rc = launch_core_spreading(f, s->nu, a, b, dt))) return rc;
            if (std::fabs(b - 8.0 / 15.0) <= 1e-7) {
                s->t_sgm += dt;
                const double beta_cur = std::sqrt(2 * s->nu * s->t_sgm / (s->sgm0 * s->sgm0) + 1);
                if (beta_cur >= s->beta) {                          // viscous.jl:202-220
                    if (s->rbf_itmax > 0) {
                        int iters = 0;
                        double res[3];
                        rc = launch_core_spreading_reset(f, cfg->kernel, cfg->precision, s->sgm0, s->rbf_itmax,
                                                         s->rbf_tol, s->rbf_iterror, &iters, res);
                        if (rc == VPM_ECONVERGE) unconverged = true;
                        else if (rc) return rc;
                        s->t_sgm = 0;
                    } else {
                        reset_needed = true;
                    }
                }
            }
        }
    }
    if (relax_now && s->relax) {                                    // :437-458
        if ((rc = launch_reset(f, 0))) return rc;
        if ((rc = uj_direct_locked(f, cfg, VPM_UJ_RESET | (uint32_t)s->uj_flags))) return rc;
        if ((rc = launch_relax(f, s->rlxf, s->relax == 2))) return rc;
    }
    if ((rc = launch_update_uprev(f))) return rc;                   // particlefield.jl:443-455
    if ((rc = sync(f))) return rc;
    if (tape) {
        tape->np = f->np;
        tape->s = *s;
        tape->dt = dt;
        tape->recorded = true;
    }
    if (reset_needed)
        return fail(VPM_ERESET, "core-spreading reset required (sigma growth >= beta) and rbf_itmax is 0");
    if (unconverged)
        return fail(VPM_ECONVERGE, "RBF interpolation: maximum number of iterations reached before convergence");
    return VPM_OK;
}

extern "C" {

int vpm_rk3_step(vpm_field f, vpm_step_config* s, double dt, int32_t relax_now)
{
    VPM_FIELD_GUARD(f);
    if (!s) return fail(VPM_EINVAL, "null vpm_step_config");
    return rk3_step_locked(f, s, dt, relax_now, nullptr);
}

int vpm_zeta_direct(vpm_field f, const vpm_config* cfg)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (int rc = launch_zeta(f, cfg->kernel, cfg->precision)) return rc;
    return sync(f);
}

int vpm_core_spreading_reset(vpm_field f, const vpm_config* cfg, double sgm0, int32_t itmax, double tol,
                             int32_t iterror, int32_t* iters, double* res)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (!(sgm0 > 0) || !std::isfinite(sgm0)) return fail(VPM_EINVAL, "sgm0 must be positive and finite");
    if (itmax < 0) return fail(VPM_EINVAL, "itmax must be >= 0");
    int it = 0;
    double r[3] = {0, 0, 0};
    const int rc = launch_core_spreading_reset(f, cfg->kernel, cfg->precision, sgm0, itmax, tol, iterror, &it, r);
    if (iters) *iters = it;
    if (res) for (int c = 0; c < 3; ++c) res[c] = r[c];
    return rc;
}

int vpm_rbf_conjugategradient(vpm_field f, const vpm_config* cfg, int32_t itmax, double tol, int32_t iterror,
                              int32_t* iters, double* res)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (itmax < 0) return fail(VPM_EINVAL, "itmax must be >= 0");
    int it = 0;
    double r[3] = {0, 0, 0};
    const int rc = launch_rbf_cg(f, cfg->kernel, cfg->precision, itmax, tol, iterror, &it, r);
    if (iters) *iters = it;
    if (res) for (int c = 0; c < 3; ++c) res[c] = r[c];
    return rc;
}

int vpm_tape_create(vpm_field f, vpm_tape* out)
{
    if (!out) return fail(VPM_EINVAL, "null output pointer");
    *out = nullptr;
    if (!f) return fail(VPM_EINVAL, "null vpm_field");
    vpm_tape t = new (std::nothrow) vpm_tape_s;
    if (!t) return fail(VPM_ENOMEM, "vpm_tape_create");
    t->f = f;
    try {
        std::lock_guard<std::mutex> lock(f->mu);
        f->tapes.push_back(t);
    } catch (const std::exception&) {
        delete t;
        return fail(VPM_ENOMEM, "vpm_tape_create");
    }
    *out = t;
    return VPM_OK;
}

int vpm_tape_destroy(vpm_tape t)
{
    if (!t) return VPM_OK;
    if (vpm_field f = t->f) {   // null once the field was destroyed (rec already released)
        std::lock_guard<std::mutex> lock(f->mu);
        if (t->rec.p) {
            if (int rc = set_device(f)) return rc;
            release(t->rec);
        }
        f->tapes.erase(std::remove(f->tapes.begin(), f->tapes.end(), t), f->tapes.end());
    }
    delete t;
    return VPM_OK;
}

int vpm_rk3_step_taped(vpm_field f, vpm_step_config* s, double dt, vpm_tape t)
{
    VPM_FIELD_GUARD(f);
    if (!s || !t) return fail(VPM_EINVAL, "null vpm_step_config or vpm_tape");
    if (t->f != f) return fail(VPM_EINVAL, "tape belongs to another field");
    return rk3_step_locked(f, s, dt, 0, t);
}

// Reverse pass over the three recorded stages: restore the stage state,
// pull the cotangents back through update_particle_states, then through the
// stage's U/J evaluation (adjoint kernel), and accumulate.
int vpm_rk3_step_adjoint(vpm_tape t, double* bar)
{
    if (!t) return fail(VPM_EINVAL, "null vpm_tape");
    if (!t->f) return fail(VPM_EINVAL, "the tape's field was destroyed");
    vpm_field f = t->f;
    VPM_FIELD_GUARD(f);
    if (!t->recorded) return fail(VPM_EINVAL, "tape holds no recorded step");
    const int64_t np = t->np;
    if (np == 0) return VPM_OK;
    if (!bar) return fail(VPM_EINVAL, "null cotangent array");
    if (np > f->cap) return fail(VPM_ERANGE, "recorded particle count exceeds the field capacity");
    const size_t rec = sizeof(double) * NF * (size_t)np;
    int rc;
    if ((rc = ensure(f->bar, rec))) return rc;
    double* B = (double*)f->bar.p;
    hipError_t e = hipMemcpyAsync(B, bar, rec, hipMemcpyHostToDevice, f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(bar)");
    f->np = np;
    AdjBufs ab;
    if ((rc = adj_buffers(f, ab))) return rc;
    const vpm_step_config& s = t->s;
    const double coef[3][2] = {{0.0, 1.0 / 3.0}, {-5.0 / 9.0, 15.0 / 16.0}, {-153.0 / 128.0, 8.0 / 15.0}};
    for (int st = 2; st >= 0; --st) {
        e = hipMemcpyAsync(f->P, (const char*)t->rec.p + st * rec, rec, hipMemcpyDeviceToDevice, f->stream);
        if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(restore stage)");
        f->perm_valid = false;
        if ((rc = launch_update_states_adjoint(f, B, coef[st][0], coef[st][1], t->dt, s.f, s.g, s.uj.transposed,
                                               ab.U, ab.J)))
            return rc;
        if ((rc = launch_uj_adjoint_dev(f, s.uj.kernel, s.uj.precision, ab))) return rc;
        if ((rc = launch_adj_add_bar(f, B, ab.X, ab.G, ab.S))) return rc;
    }
    if ((rc = launch_adj_step_start(f, B))) return rc;
    e = hipMemcpyAsync(bar, B, rec, hipMemcpyDeviceToHost, f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(bar out)");
    return sync(f);
}

int vpm_uj_adjoint(vpm_field f, const vpm_config* cfg, const double* Ubar, const double* Jbar, double* Xbar,
                   double* Gbar, double* Sbar)
{
    VPM_FIELD_GUARD(f);
    if (int rc = check_cfg(cfg)) return rc;
    if (f->np > 0 && (!Ubar || !Jbar || !Xbar || !Gbar || !Sbar)) return fail(VPM_EINVAL, "null adjoint array");
    if (int rc = launch_uj_adjoint(f, cfg->kernel, cfg->precision, Ubar, Jbar, Xbar, Gbar, Sbar)) return rc;
    return sync(f);
}

int vpm_profile(vpm_field f, int32_t enable)
{
    VPM_FIELD_GUARD(f);
    f->profiling = enable != 0;
    for (auto& t : f->timing) t = Timing();
    return VPM_OK;
}

int vpm_timing(vpm_field f, int32_t which, double* total_ms, int64_t* count)
{
    if (!f || !total_ms || !count) return fail(VPM_EINVAL, "null argument");
    if (which < 0 || which > 4) return fail(VPM_EINVAL, "timing class must be 0..4");
    std::lock_guard<std::mutex> lock(f->mu);
    *total_ms = f->timing[which].ms;
    *count = f->timing[which].count;
    return VPM_OK;
}

}  // extern "C"
