// p2p_ranges.hip -- U/J over a list of (target range, source range) pairs:
// FastMultipole's leaf-pair P2P (fmm.direct! per direct-list entry,
// src/FLOWVPM_fmm.jl:102-168), batched into one launch on one device copy.
//
// Target ranges of different pairs may overlap, so the host first cuts the
// target index space at every range end into disjoint elementary intervals
// and gives each interval the list of source ranges of every pair covering
// it, in pair order.  A block then owns up to 256 targets of one interval,
// streams its source ranges through LDS, and adds the result onto its own
// targets' records: no two blocks write the same particle, and each target's
// sum runs in pair order.  The pair arithmetic is p2p_pair.h's.
#include <algorithm>
#include <vector>

#include "p2p_pair.h"

namespace vpm {

struct RangeItem {
    int64_t t_begin, t_end;    // targets of this block (<= 256, one interval)
    int64_t l_begin, l_end;    // its source ranges in the list arrays
};

template <int K>
__global__ __launch_bounds__(kBlock) void ranges_f32c_kernel(const RangeItem* __restrict__ items,
                                                             const int64_t* __restrict__ lsb,
                                                             const int64_t* __restrict__ lse,
                                                             const double* __restrict__ pos, int64_t ld,
                                                             const float4* __restrict__ gs,
                                                             const float* __restrict__ cut,
                                                             double* __restrict__ P)
{
    __shared__ float4 lpos[kTile];
    __shared__ float4 lgs[kTile];
    __shared__ float4 llo[kTile];
    const RangeItem it = items[blockIdx.x];
    const int tid = threadIdx.x;
    const int64_t ib = it.t_begin;
    const double ox = pos[ib], oy = pos[ld + ib], oz = pos[2 * ld + ib];
    const int64_t t = it.t_begin + tid;
    const bool active = t < it.t_end;
    const int64_t tt = active ? t : ib;
    float tx, ty, tz, tlx, tly, tlz;
    split_rel(pos[tt], ox, tx, tlx);
    split_rel(pos[ld + tt], oy, ty, tly);
    split_rel(pos[2 * ld + tt], oz, tz, tlz);
    const WaveBox box = wave_box(tx, ty, tz);
    float sum[kAccF32], err[kAccF32];
#pragma unroll
    for (int c = 0; c < kAccF32; ++c) { sum[c] = 0.f; err[c] = 0.f; }
    for (int64_t l = it.l_begin; l < it.l_end; ++l) {
        const int64_t sb = lsb[l], se = lse[l];
        for (int64_t base = sb; base < se; base += kTile) {
            const int n = (int)((se - base) < kTile ? (se - base) : kTile);
            __syncthreads();
            if (tid < n) {
                const int64_t j = base + tid;
                float4 p, q;
                split_rel(pos[j], ox, p.x, q.x);
                split_rel(pos[ld + j], oy, p.y, q.y);
                split_rel(pos[2 * ld + j], oz, p.z, q.z);
                p.w = cut[j];
                q.w = 0.f;
                lpos[tid] = p;
                llo[tid] = q;
                lgs[tid] = gs[j];
            }
            __syncthreads();
            float acc[kAccF32];
#pragma unroll
            for (int c = 0; c < kAccF32; ++c) acc[c] = 0.f;
            tile_f32c<K>(n, lpos, lgs, llo, tx, ty, tz, tlx, tly, tlz, box, acc);
            fold_twosum(acc, sum, err);
        }
    }
    if (active) {
        double a[kAccF32];
#pragma unroll
        for (int c = 0; c < kAccF32; ++c) a[c] = (double)sum[c] + (double)err[c];
        add_f32c_to_record(P + t * NF, a);
    }
}

#pragma clang fp contract(off)
template <int K>
__global__ __launch_bounds__(kBlock) void ranges_f64_kernel(const RangeItem* __restrict__ items,
                                                            const int64_t* __restrict__ lsb,
                                                            const int64_t* __restrict__ lse,
                                                            const double* __restrict__ pos, int64_t ld,
                                                            const double4* __restrict__ g64,
                                                            double* __restrict__ P)
{
    __shared__ double4 lpos[kTile];
    __shared__ double4 lgam[kTile];
    const RangeItem it = items[blockIdx.x];
    const int tid = threadIdx.x;
    const int64_t t = it.t_begin + tid;
    const bool active = t < it.t_end;
    const int64_t tt = active ? t : it.t_begin;
    const double tx = pos[tt], ty = pos[ld + tt], tz = pos[2 * ld + tt];
    double acc[kAccF64];
#pragma unroll
    for (int c = 0; c < kAccF64; ++c) acc[c] = 0.0;
    for (int64_t l = it.l_begin; l < it.l_end; ++l) {
        const int64_t sb = lsb[l], se = lse[l];
        for (int64_t base = sb; base < se; base += kTile) {
            const int n = (int)((se - base) < kTile ? (se - base) : kTile);
            __syncthreads();
            if (tid < n) {
                const int64_t j = base + tid;
                const double4 g = g64[j];
                lpos[tid] = make_double4(pos[j], pos[ld + j], pos[2 * ld + j], g.w);
                lgam[tid] = make_double4(g.x, g.y, g.z, 0.0);
            }
            __syncthreads();
            for (int s = 0; s < n; ++s) pair_f64<K>(tx, ty, tz, lpos[s], lgam[s], acc);
        }
    }
    if (active) {
        double* rec = P + t * NF;
#pragma unroll
        for (int c = 0; c < 3; ++c) rec[RU + c] += acc[c];
#pragma unroll
        for (int c = 0; c < 9; ++c) rec[RJ + c] += acc[3 + c];
    }
}
#pragma clang fp contract(on)

int launch_p2p_ranges(vpm_field f, int kernel, int precision, const int64_t* tb, const int64_t* te,
                      const int64_t* sb, const int64_t* se, int64_t npairs)
{
    // elementary target intervals
    std::vector<int64_t> cuts;
    cuts.reserve(2 * (size_t)npairs);
    for (int64_t k = 0; k < npairs; ++k)
        if (tb[k] < te[k] && sb[k] < se[k]) { cuts.push_back(tb[k]); cuts.push_back(te[k]); }
    if (cuts.empty()) return VPM_OK;
    std::sort(cuts.begin(), cuts.end());
    cuts.erase(std::unique(cuts.begin(), cuts.end()), cuts.end());
    const size_t nint = cuts.size() - 1;
    std::vector<int64_t> count(nint + 1, 0);
    for (int64_t k = 0; k < npairs; ++k) {
        if (!(tb[k] < te[k] && sb[k] < se[k])) continue;
        const size_t lo = std::lower_bound(cuts.begin(), cuts.end(), tb[k]) - cuts.begin();
        const size_t hi = std::lower_bound(cuts.begin(), cuts.end(), te[k]) - cuts.begin();
        for (size_t m = lo; m < hi; ++m) count[m + 1]++;
    }
    for (size_t m = 0; m < nint; ++m) count[m + 1] += count[m];
    std::vector<int64_t> lsb((size_t)count[nint]), lse((size_t)count[nint]);
    std::vector<int64_t> fill(count.begin(), count.end() - 1);
    for (int64_t k = 0; k < npairs; ++k) {
        if (!(tb[k] < te[k] && sb[k] < se[k])) continue;
        const size_t lo = std::lower_bound(cuts.begin(), cuts.end(), tb[k]) - cuts.begin();
        const size_t hi = std::lower_bound(cuts.begin(), cuts.end(), te[k]) - cuts.begin();
        for (size_t m = lo; m < hi; ++m) {
            lsb[(size_t)fill[m]] = sb[k];
            lse[(size_t)fill[m]] = se[k];
            fill[m]++;
        }
    }
    std::vector<RangeItem> items;
    for (size_t m = 0; m < nint; ++m) {
        if (count[m + 1] == count[m]) continue;
        for (int64_t t = cuts[m]; t < cuts[m + 1]; t += kBlock)
            items.push_back({t, std::min<int64_t>(t + kBlock, cuts[m + 1]), count[m], count[m + 1]});
    }
    if (items.empty()) return VPM_OK;

    Staged st;
    int rc;
    if ((rc = stage_field(f, kernel, precision, false, st))) return rc;
    const size_t bytes_items = sizeof(RangeItem) * items.size();
    const size_t bytes_list = sizeof(int64_t) * lsb.size();
    if ((rc = ensure(f->rng, bytes_items + 2 * bytes_list + 64))) return rc;
    char* base = (char*)f->rng.p;
    RangeItem* d_items = (RangeItem*)base;
    int64_t* d_lsb = (int64_t*)(base + bytes_items);
    int64_t* d_lse = d_lsb + lsb.size();
    hipError_t e;
    if ((e = hipMemcpyAsync(d_items, items.data(), bytes_items, hipMemcpyHostToDevice, f->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(d_lsb, lsb.data(), bytes_list, hipMemcpyHostToDevice, f->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(d_lse, lse.data(), bytes_list, hipMemcpyHostToDevice, f->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(range lists)");
    const dim3 grid((unsigned)items.size());
    timing_begin(f);
    if (precision == VPM_PREC_F64) {
        switch (kernel) {
#define VPM_R64(KID) case KID: hipLaunchKernelGGL(ranges_f64_kernel<KID>, grid, dim3(kBlock), 0, f->stream, d_items, d_lsb, d_lse, st.pos, st.ld, st.g64, f->P); break;
            VPM_R64(VPM_KERNEL_SINGULAR) VPM_R64(VPM_KERNEL_GAUSSIAN) VPM_R64(VPM_KERNEL_GAUSSIANERF) VPM_R64(VPM_KERNEL_WINCKELMANS)
#undef VPM_R64
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    } else {
        switch (kernel) {
#define VPM_R32(KID) case KID: hipLaunchKernelGGL(ranges_f32c_kernel<KID>, grid, dim3(kBlock), 0, f->stream, d_items, d_lsb, d_lse, st.pos, st.ld, st.gs, st.cut, f->P); break;
            VPM_R32(VPM_KERNEL_SINGULAR) VPM_R32(VPM_KERNEL_GAUSSIAN) VPM_R32(VPM_KERNEL_GAUSSIANERF) VPM_R32(VPM_KERNEL_WINCKELMANS)
#undef VPM_R32
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "p2p_ranges kernel");
    timing_end(f, precision == VPM_PREC_F64 ? 1 : 0);
    // the host vectors must outlive the asynchronous copies
    e = hipStreamSynchronize(f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return VPM_OK;
}

}  // namespace vpm
