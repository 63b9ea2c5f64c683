// noalloc.hip -- vpm_uj_direct without device allocation.
//
// vpm_uj_direct grows the field's staging and partial-sum buffers with
// hipMalloc on first use.  Callers that cannot allocate (code under HIP graph
// capture, tight serving loops, harnesses that forbid it) call
// vpm_uj_direct_noalloc instead: the same kernels, with every scratch buffer
// (Morton keys and permutation, the radix sort's temp storage, the staged
// sources, the K x 14 (or 12) FP64 partial sums per target) taken from one
// file-scope device array of kNoAllocBytes.  Each GPU holds its own copy of
// the array until the process ends; its address on a device is looked up
// once, with that device current, and kept.
//
// The array is shared by every call on one GPU and keeps the last call's
// contents (every byte a call reads it first writes).  Calls that overlapped
// on one GPU would race on it, so a call holds that GPU's lock until its
// work on the field's stream has finished.
#include <mutex>
#include <string>

#include "vpm_internal.h"

namespace vpm {

constexpr size_t kNoAllocBytes = (size_t)64 << 20;   // 64 MiB per GPU

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

namespace {

constexpr int kMaxDevices = 64;

struct DeviceScratch {
    std::mutex mu;          // held by a call for its whole use of the array
    void* ptr = nullptr;    // the array's device address on this GPU
};

DeviceScratch g_dev[kMaxDevices];
std::mutex g_lookup;

// The array's address on the current device f->device (looked up on the
// first call there, with the device current).
int scratch_ptr(vpm_field f, void*& p)
{
    if (f->device < 0 || f->device >= kMaxDevices) return fail(VPM_EDEVICE, "device index beyond the no-alloc table");
    std::lock_guard<std::mutex> lk(g_lookup);
    DeviceScratch& d = g_dev[f->device];
    if (!d.ptr) {
        void* q = nullptr;
        hipError_t e = hipGetSymbolAddress(&q, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !q) return e != hipSuccess ? hip_fail(e, "hipGetSymbolAddress(scratch)")
                                                          : fail(VPM_EDEVICE, "no address for the no-alloc scratch");
        d.ptr = q;
    }
    p = d.ptr;
    return VPM_OK;
}

size_t al(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// Bytes of the array one evaluation takes with K source splits (the order of
// stage_field, sort_field and launch_uj; every piece 256-byte aligned).
static int noalloc_bytes(vpm_field f, int precision, bool sort, int64_t K, size_t& out)
{
    const size_t np = (size_t)f->np;
    if (precision == VPM_PREC_F64) {
        out = al(sizeof(double) * 3 * np) + al(sizeof(double4) * np) + al(sizeof(double) * kAccF64 * K * np);
        return VPM_OK;
    }
    size_t sb = 0;
    if (sort && np > 1)
        if (int rc = sort_scratch_bytes(f, sb)) return rc;
    out = al(sizeof(double) * 3 * np) + al(sizeof(float4) * np) + al(sizeof(float) * np) + sb +
          al(sizeof(double) * kAccF32 * K * np);
    return VPM_OK;
}

}  // namespace vpm

using namespace vpm;

extern "C" int vpm_uj_direct_noalloc(vpm_field f, const vpm_config* cfg, uint32_t flags)
{
    if (!f) return fail(VPM_EINVAL, "null vpm_field");
    if (!cfg) return fail(VPM_EINVAL, "null vpm_config");
    if (cfg->kernel < VPM_KERNEL_SINGULAR || cfg->kernel > VPM_KERNEL_WINCKELMANS)
        return fail(VPM_EINVAL, "unknown kernel id " + std::to_string(cfg->kernel));
    if (cfg->precision != VPM_PREC_F32C && cfg->precision != VPM_PREC_F64)
        return fail(VPM_EINVAL, "unknown precision " + std::to_string(cfg->precision));
    if (flags & (VPM_UJ_SFS | VPM_UJ_F64_FAR))
        return fail(VPM_EINVAL, "vpm_uj_direct_noalloc: VPM_UJ_SFS (E_str) and VPM_UJ_F64_FAR are not served "
                                "without allocation; call vpm_uj_direct");
    std::lock_guard<std::mutex> lock(f->mu);
    hipError_t e = hipSetDevice(f->device);
    if (e != hipSuccess) return hip_fail(e, "hipSetDevice");
    void* base = nullptr;
    int rc;
    if ((rc = scratch_ptr(f, base))) return rc;
    const bool sort = cfg->precision == VPM_PREC_F32C && !(flags & VPM_UJ_NO_SORT);
    const int64_t np = f->np;
    // the splits vpm_uj_direct would take, fewer when they do not fit
    int64_t K = 1, chunk = 0;
    size_t need = 0;
    if (np > 0) {
        choose_split(np, np, 0, K, chunk);
        size_t one = 0;
        if ((rc = noalloc_bytes(f, cfg->precision, sort, 1, one))) return rc;
        if (one > kNoAllocBytes)
            return fail(VPM_ERANGE, "vpm_uj_direct_noalloc: " + std::to_string(np) + " particles need " +
                                        std::to_string(one) + " bytes of scratch, above the " +
                                        std::to_string(kNoAllocBytes) + "-byte maximum");
        if ((rc = noalloc_bytes(f, cfg->precision, sort, K, need))) return rc;
        if (need > kNoAllocBytes) {
            const size_t per_split = sizeof(double) * (cfg->precision == VPM_PREC_F64 ? kAccF64 : kAccF32) *
                                     (size_t)np;
            K = std::min<int64_t>(K, 1 + (int64_t)((kNoAllocBytes - one) / al(per_split)));
            if ((rc = noalloc_bytes(f, cfg->precision, sort, K, need))) return rc;
            while (K > 1 && need > kNoAllocBytes)
                if ((rc = noalloc_bytes(f, cfg->precision, sort, --K, need))) return rc;
        }
    }
    DeviceScratch& d = g_dev[f->device];
    std::lock_guard<std::mutex> dlk(d.mu);
    Arena ar{(char*)base, kNoAllocBytes, 0};
    if (flags & VPM_UJ_RESET) if ((rc = launch_reset(f, 0))) return rc;
    if (flags & VPM_UJ_RESET_SFS) if ((rc = launch_reset(f, 1))) return rc;
    Staged st;
    if ((rc = stage_field(f, cfg->kernel, cfg->precision, sort, st, &ar)) ||
        (rc = launch_uj(f, st, 0, np, st, 0, np, cfg->kernel, cfg->precision, &ar, K))) {
        (void)hipStreamSynchronize(f->stream);   // no launch of this call still reads the array
        return rc;
    }
    e = hipStreamSynchronize(f->stream);
    return e == hipSuccess ? VPM_OK : hip_fail(e, "hipStreamSynchronize");
}
