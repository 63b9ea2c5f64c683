// p2p_uj.hip -- all-pairs U/J kernels for gfx950 (src/FLOWVPM_fmm.jl:102-176).
//
// Two precisions share one tiling: a block of 256 lanes (4 waves of 64) owns
// 256 consecutive targets, one per lane; the sources of the block's source
// range stream through LDS in tiles of 256, loaded with one coalesced record
// per lane; every lane then reads each source as an LDS broadcast.  Large
// source counts per target are split across blockIdx.y into K chunks whose
// partial sums are reduced in a fixed order (deterministic, no atomics).
//
// F32C (FP32 pair arithmetic, compensated sums):
//  * positions are differenced relative to the block origin O (the block's
//    first target, FP64): t' = fl32(t - O), s' = fl32(s - O).  After a
//    spatial (Morton) sort, |t'|,|s'| are small for every pair that matters,
//    so dx = t' - s' keeps FP32 relative accuracy of the distance itself
//    rather than of the absolute coordinate;
//  * with Gamma' = -Gamma/(4 pi) per source, a pair contributes
//      U  += F' (d x Gamma'),   W += F' Gamma',   Jhv += H'' (d x Gamma') d^T
//    with F' = g/r^3 and H'' = -(q g' - 3 g)/(3 r^5); at the end
//      J = -3 Jhv + eps . W      (the epsilon term of fmm.jl:146-158)
//    and Jhv[2][2] = -Jhv[0][0] - Jhv[1][1] (every pair's Jhv is traceless);
//  * far field: where g == 1 and q g' - 3 g == -3 to FP32 rounding
//    (q >= q_far), F' = r^-3, H'' = r^-5.  The test r^2 < (q_far sigma)^2 is
//    voted per wave per source; only if some lane is near does the wave take
//    the full kernel evaluation (which is valid at every q);
//  * near field: reduced forms (kernel_fits.h) below the cancellation
//    threshold, closed forms above it;
//  * every tile's 256 contributions are summed in FP32, and tile sums are
//    folded into a TwoSum (value, error) pair per accumulator.
//
// F64: the reference's FP64 pair arithmetic statement by statement
// (fmm.jl:113-161, kernel.jl:45-84, custom_erf64 of gpu_erf.jl:159-191),
// contraction off, sources in storage order.  This is the GPU FP64 oracle.
#include <cfloat>
#include <cstring>

#include <rocprim/device/device_radix_sort.hpp>

#include "p2p_pair.h"


namespace vpm {

// ------------------------------------------------------------ F32C kernel
template <int K>
__global__ __launch_bounds__(kBlock) void uj_f32c_kernel(
    const double* __restrict__ tpos, int64_t tld, int64_t t0, int64_t nt,
    const double* __restrict__ spos, int64_t sld, const float4* __restrict__ sgs,
    const float* __restrict__ scut, int64_t s0, int64_t s1, int64_t chunk,
    double* __restrict__ part, int64_t pld)
{
    __shared__ float4 lpos[kTile];   // s'_hi x, y, z, far-field r^2 threshold
    __shared__ float4 lgs[kTile];    // Gamma' x, y, z, 1/sigma
    __shared__ float4 llo[kTile];    // s'_lo x, y, z

    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const int64_t ib = t0 + (int64_t)blockIdx.x * kBlock;
    const double ox = tpos[ib], oy = tpos[tld + ib], oz = tpos[2 * tld + ib];
    const bool active = i < nt;
    const int64_t it = active ? t0 + i : ib;
    float tx, ty, tz, tlx, tly, tlz;
    split_rel(tpos[it], ox, tx, tlx);
    split_rel(tpos[tld + it], oy, ty, tly);
    split_rel(tpos[2 * tld + it], oz, tz, tlz);

    const WaveBox box = wave_box(tx, ty, tz);
    const int64_t sb = s0 + (int64_t)blockIdx.y * chunk;
    const int64_t se = min(sb + chunk, s1);

    float sum[kAccF32], err[kAccF32];
#pragma unroll
    for (int c = 0; c < kAccF32; ++c) { sum[c] = 0.f; err[c] = 0.f; }

    for (int64_t base = sb; base < se; base += kTile) {
        const int64_t j = base + tid;
        float4 p, g, l;
        if (j < se) {
            split_rel(spos[j], ox, p.x, l.x);
            split_rel(spos[sld + j], oy, p.y, l.y);
            split_rel(spos[2 * sld + j], oz, p.z, l.z);
            p.w = scut[j];
            l.w = 0.f;
            g = sgs[j];
        } else {   // padding: zero strength, far away, never near
            p = make_float4(1e15f, 1e15f, 1e15f, 0.f);
            g = make_float4(0.f, 0.f, 0.f, 0.f);
            l = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();
        lpos[tid] = p;
        lgs[tid] = g;
        llo[tid] = l;
        __syncthreads();

        float acc[kAccF32];
#pragma unroll
        for (int c = 0; c < kAccF32; ++c) acc[c] = 0.f;
        tile_f32c<K>((int)min((int64_t)kTile, se - base), lpos, lgs, llo, tx, ty, tz, tlx, tly, tlz, box, acc, sum,
                     err);
        fold_twosum(acc, sum, err);
    }
    if (active) {
        double* out = part + (int64_t)blockIdx.y * kAccF32 * pld + i;
#pragma unroll
        for (int c = 0; c < kAccF32; ++c) out[c * pld] = (double)sum[c] + (double)err[c];
    }
}

// ------------------------------------------------------ F64 oracle kernel
#pragma clang fp contract(off)
template <int K>
__global__ __launch_bounds__(kBlock) void uj_f64_kernel(
    const double* __restrict__ tpos, int64_t tld, int64_t t0, int64_t nt,
    const double* __restrict__ spos, int64_t sld, const double4* __restrict__ sg,
    int64_t s0, int64_t s1, int64_t chunk, double* __restrict__ part, int64_t pld)
{
    __shared__ double4 lpos[kTile];   // x, y, z, sigma
    __shared__ double4 lgam[kTile];   // Gamma x, y, z

    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const bool active = i < nt;
    const int64_t it = t0 + (active ? i : 0);
    const double tx = tpos[it], ty = tpos[tld + it], tz = tpos[2 * tld + it];
    const int64_t sb = s0 + (int64_t)blockIdx.y * chunk;
    const int64_t se = min(sb + chunk, s1);

    double acc[kAccF64];
#pragma unroll
    for (int c = 0; c < kAccF64; ++c) acc[c] = 0.0;

    for (int64_t base = sb; base < se; base += kTile) {
        const int64_t j = base + tid;
        const int64_t n = min((int64_t)kTile, se - base);
        __syncthreads();
        if (tid < n) {
            const double4 gg = sg[j];
            lpos[tid] = make_double4(spos[j], spos[sld + j], spos[2 * sld + j], gg.w);
            lgam[tid] = make_double4(gg.x, gg.y, gg.z, 0.0);
        }
        __syncthreads();
        for (int s = 0; s < n; ++s) pair_f64<K>(tx, ty, tz, lpos[s], lgam[s], acc);
    }
    if (active) {
        double* out = part + (int64_t)blockIdx.y * kAccF64 * pld + i;
#pragma unroll
        for (int c = 0; c < kAccF64; ++c) out[c * pld] = acc[c];
    }
}
#pragma clang fp contract(on)

// ----------------------------------------------- F64 far-form kernel
// FP64 pair arithmetic with the far form of Far64<K> (VPM_UJ_F64_FAR): the
// sources come in the F32C kernel's spatial order, staged as (x, y, z, far
// r^2 threshold) and (Gamma', 1/sigma); each wave splits a 64-source group
// into near and far sources against its target box.  The 14 sums are those
// of the F32C kernel (add_f32c_to_record), summed in FP64.
template <int K>
__global__ __launch_bounds__(kBlock, 5) void uj_f64far_kernel(
    const double* __restrict__ tpos, int64_t tld, int64_t t0, int64_t nt,
    const double* __restrict__ spos, int64_t sld, const double4* __restrict__ sg,
    int64_t s0, int64_t s1, int64_t chunk, double* __restrict__ part, int64_t pld)
{
    __shared__ double4 lpos[kTile];   // x, y, z, far-field r^2 threshold
    __shared__ double4 lgs[kTile];    // Gamma' x, y, z, 1/sigma

    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const bool active = i < nt;
    const int64_t it = t0 + (active ? i : (int64_t)blockIdx.x * kBlock);
    const double tx = tpos[it], ty = tpos[tld + it], tz = tpos[2 * tld + it];
    const WaveBox64 box = wave_box64(tx, ty, tz);
    const int64_t sb = s0 + (int64_t)blockIdx.y * chunk;
    const int64_t se = min(sb + chunk, s1);

    double acc[kAccF32];
#pragma unroll
    for (int c = 0; c < kAccF32; ++c) acc[c] = 0.0;

    for (int64_t base = sb; base < se; base += kTile) {
        const int64_t j = base + tid;
        double4 p, g;
        if (j < se) {
            const double4 r = sg[j];   // Gamma', sigma
            p = make_double4(spos[j], spos[sld + j], spos[2 * sld + j], far_cut64<K>(r.w));
            g = make_double4(r.x, r.y, r.z, 1.0 / r.w);
        } else {   // not visited (s >= n in the tile loop)
            p = make_double4(0.0, 0.0, 0.0, 0.0);
            g = make_double4(0.0, 0.0, 0.0, 0.0);
        }
        __syncthreads();
        lpos[tid] = p;
        lgs[tid] = g;
        __syncthreads();
        tile_f64far<K>((int)min((int64_t)kTile, se - base), lpos, lgs, tx, ty, tz, box, acc);
    }
    if (active) {
        double* out = part + (int64_t)blockIdx.y * kAccF32 * pld + i;
#pragma unroll
        for (int c = 0; c < kAccF32; ++c) out[c * pld] = acc[c];
    }
}

// ------------------------------------------------------------- finalize
// Sums the K partials of each target in split order and adds the total onto
// U and J of the particle record (buffer_to_target_system!, fmm.jl:170-176).
__global__ __launch_bounds__(kBlock) void finalize_f32c_kernel(
    const double* __restrict__ part, int64_t pld, int K, int64_t nt,
    const uint32_t* __restrict__ perm, int64_t t0, double* __restrict__ P)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nt) return;
    double a[kAccF32];
#pragma unroll
    for (int c = 0; c < kAccF32; ++c) a[c] = 0.0;
    for (int k = 0; k < K; ++k) {
        const double* p = part + (int64_t)k * kAccF32 * pld + i;
#pragma unroll
        for (int c = 0; c < kAccF32; ++c) a[c] += p[c * pld];
    }
    const int64_t idx = perm ? (int64_t)perm[t0 + i] : t0 + i;
    add_f32c_to_record(P + idx * NF, a);
}

__global__ __launch_bounds__(kBlock) void finalize_f64_kernel(
    const double* __restrict__ part, int64_t pld, int K, int64_t nt,
    const uint32_t* __restrict__ perm, int64_t t0, double* __restrict__ P)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= nt) return;
    double a[kAccF64];
#pragma unroll
    for (int c = 0; c < kAccF64; ++c) a[c] = 0.0;
    for (int k = 0; k < K; ++k) {
        const double* p = part + (int64_t)k * kAccF64 * pld + i;
#pragma unroll
        for (int c = 0; c < kAccF64; ++c) a[c] += p[c * pld];
    }
    const int64_t idx = perm ? (int64_t)perm[t0 + i] : t0 + i;
    double* rec = P + idx * NF;
#pragma unroll
    for (int c = 0; c < 3; ++c) rec[RU + c] += a[c];
#pragma unroll
    for (int c = 0; c < 9; ++c) rec[RJ + c] += a[3 + c];
}

// ------------------------------------------------------------- staging
// bounding box of the positions: per-block partial min/max, then one block
__global__ __launch_bounds__(kBlock) void bbox_partial_kernel(const double* __restrict__ P, int64_t np,
                                                              double* __restrict__ out)
{
    __shared__ double red[6][kBlock];
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x; i < np; i += (int64_t)gridDim.x * kBlock) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const double v = P[i * NF + RX + c];
            mn[c] = fmin(mn[c], v);
            mx[c] = fmax(mx[c], v);
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) { red[c][threadIdx.x] = mn[c]; red[3 + c][threadIdx.x] = mx[c]; }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                red[c][threadIdx.x] = fmin(red[c][threadIdx.x], red[c][threadIdx.x + w]);
                red[3 + c][threadIdx.x] = fmax(red[3 + c][threadIdx.x], red[3 + c][threadIdx.x + w]);
            }
        }
        __syncthreads();
    }
    if (threadIdx.x < 6) out[6 + blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(64) void bbox_final_kernel(double* __restrict__ bb, int nblocks)
{
    const int c = threadIdx.x;
    if (c >= 6) return;
    double v = c < 3 ? DBL_MAX : -DBL_MAX;
    for (int b = 0; b < nblocks; ++b) {
        const double x = bb[6 + b * 6 + c];
        v = c < 3 ? fmin(v, x) : fmax(v, x);
    }
    bb[c] = v;
}

__device__ __forceinline__ uint64_t spread21(uint64_t v)
{
    v &= 0x1fffff;
    v = (v | v << 32) & 0x1f00000000ffffull;
    v = (v | v << 16) & 0x1f0000ff0000ffull;
    v = (v | v << 8) & 0x100f00f00f00f00full;
    v = (v | v << 4) & 0x10c30c30c30c30c3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}

__global__ __launch_bounds__(kBlock) void morton_kernel(const double* __restrict__ P, int64_t np,
                                                        const double* __restrict__ bb,
                                                        uint64_t* __restrict__ keys,
                                                        uint32_t* __restrict__ iota)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    const double ext = fmax(fmax(bb[3] - bb[0], bb[4] - bb[1]), fmax(bb[5] - bb[2], 1e-300));
    const double sc = 2097151.0 / ext;
    uint64_t k = 0;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double u = (P[i * NF + RX + c] - bb[c]) * sc;
        u = fmin(fmax(u, 0.0), 2097151.0);
        k |= spread21((uint64_t)u) << c;
    }
    keys[i] = k;
    iota[i] = (uint32_t)i;
}

template <int K>
__global__ __launch_bounds__(kBlock) void pack_f32c_kernel(const double* __restrict__ P, int64_t np,
                                                           const uint32_t* __restrict__ perm,
                                                           double* __restrict__ pos, int64_t ld,
                                                           float4* __restrict__ gs, float* __restrict__ cut)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= np) return;
    const int64_t p = perm ? (int64_t)perm[k] : k;
    const double* rec = P + p * NF;
    pos[k] = rec[RX];
    pos[ld + k] = rec[RX + 1];
    pos[2 * ld + k] = rec[RX + 2];
    const double sg = rec[RSIG];
    gs[k] = make_float4((float)(-kConst4 * rec[RG]), (float)(-kConst4 * rec[RG + 1]),
                        (float)(-kConst4 * rec[RG + 2]), (float)(1.0 / sg));
    cut[k] = far_cut<K>(sg);
}

__global__ __launch_bounds__(kBlock) void pack_f64_kernel(const double* __restrict__ P, int64_t np,
                                                          double* __restrict__ pos, int64_t ld,
                                                          double4* __restrict__ g64)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= np) return;
    const double* rec = P + k * NF;
    pos[k] = rec[RX];
    pos[ld + k] = rec[RX + 1];
    pos[2 * ld + k] = rec[RX + 2];
    g64[k] = make_double4(rec[RG], rec[RG + 1], rec[RG + 2], rec[RSIG]);
}

// sources of the FP64 far-form kernel: positions in the order of perm (or
// storage order), (-Gamma / 4 pi, sigma)
__global__ __launch_bounds__(kBlock) void pack_f64far_kernel(const double* __restrict__ P, int64_t np,
                                                             const uint32_t* __restrict__ perm,
                                                             double* __restrict__ pos, int64_t ld,
                                                             double4* __restrict__ g64)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= np) return;
    const int64_t p = perm ? (int64_t)perm[k] : k;
    const double* rec = P + p * NF;
    pos[k] = rec[RX];
    pos[ld + k] = rec[RX + 1];
    pos[2 * ld + k] = rec[RX + 2];
    g64[k] = make_double4(-kConst4 * rec[RG], -kConst4 * rec[RG + 1], -kConst4 * rec[RG + 2], rec[RSIG]);
}

static inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

#define VPM_LAUNCH_CHECK(what)                                                 \
    do {                                                                       \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess) return hip_fail(e_, what);                       \
    } while (0)

// A grow-only field buffer, or the next piece of the arena when there is one.
static int take_buffer(vpm_field, DevBuf& b, size_t bytes, Arena* ar, void*& p, const char* what)
{
    if (!ar) {
        if (int rc = ensure(b, bytes)) return rc;
        p = b.p;
        return VPM_OK;
    }
    p = ar->take(bytes);
    if (!p) return fail(VPM_ERANGE, std::string("no-alloc scratch exhausted at ") + what);
    return VPM_OK;
}

static constexpr int kSortBoxes = 256;   // blocks of the bounding-box pass

int sort_scratch_bytes(vpm_field f, size_t& bytes)
{
    const size_t np = (size_t)f->np;
    size_t tmp = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, tmp, (const uint64_t*)nullptr, (uint64_t*)nullptr,
                                             (const uint32_t*)nullptr, (uint32_t*)nullptr, (unsigned)np, 0, 63,
                                             f->stream);
    if (e != hipSuccess) return hip_fail(e, "radix_sort_pairs (size query)");
    const auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
    bytes = al(sizeof(double) * (6 + 6 * kSortBoxes)) + 2 * al(sizeof(uint64_t) * np) +
            2 * al(sizeof(uint32_t) * np) + al(tmp);
    return VPM_OK;
}

// Morton (Z-order) permutation of the field's particles: perm[k] is the
// particle at sorted position k.  Reused until positions change (not when
// it lives in an arena).
int sort_field(vpm_field f, const uint32_t*& perm, Arena* ar)
{
    const int64_t np = f->np;
    const int64_t ld = ar ? np : f->cap;
    int rc;
    perm = nullptr;
    if (np <= 1) return VPM_OK;
    if (f->perm_valid && !ar) {
        perm = (const uint32_t*)f->perm.p;
        return VPM_OK;
    }
    void *bbp, *kp, *k2p, *pp, *ip;
    if ((rc = take_buffer(f, f->bbox, sizeof(double) * (6 + 6 * kSortBoxes), ar, bbp, "bbox"))) return rc;
    if ((rc = take_buffer(f, f->keys, sizeof(uint64_t) * (size_t)ld, ar, kp, "keys"))) return rc;
    if ((rc = take_buffer(f, f->keys2, sizeof(uint64_t) * (size_t)ld, ar, k2p, "keys"))) return rc;
    if ((rc = take_buffer(f, f->perm, sizeof(uint32_t) * (size_t)ld, ar, pp, "perm"))) return rc;
    if ((rc = take_buffer(f, f->iota, sizeof(uint32_t) * (size_t)ld, ar, ip, "perm"))) return rc;
    double* bb = (double*)bbp;
    hipLaunchKernelGGL(bbox_partial_kernel, dim3(kSortBoxes), dim3(kBlock), 0, f->stream, f->P, np, bb);
    VPM_LAUNCH_CHECK("bbox_partial");
    hipLaunchKernelGGL(bbox_final_kernel, dim3(1), dim3(64), 0, f->stream, bb, kSortBoxes);
    VPM_LAUNCH_CHECK("bbox_final");
    hipLaunchKernelGGL(morton_kernel, dim3(nblocks(np)), dim3(kBlock), 0, f->stream, f->P, np, bb,
                       (uint64_t*)kp, (uint32_t*)ip);
    VPM_LAUNCH_CHECK("morton");
    size_t tmp = 0;
    hipError_t e = rocprim::radix_sort_pairs(nullptr, tmp, (const uint64_t*)kp, (uint64_t*)k2p,
                                             (const uint32_t*)ip, (uint32_t*)pp, (unsigned)np, 0,
                                             63, f->stream);
    if (e != hipSuccess) return hip_fail(e, "radix_sort_pairs (size query)");
    void* tp;
    if ((rc = take_buffer(f, f->sorttmp, tmp, ar, tp, "sort"))) return rc;
    e = rocprim::radix_sort_pairs(tp, tmp, (const uint64_t*)kp, (uint64_t*)k2p,
                                  (const uint32_t*)ip, (uint32_t*)pp, (unsigned)np, 0, 63,
                                  f->stream);
    if (e != hipSuccess) return hip_fail(e, "radix_sort_pairs");
    if (!ar) f->perm_valid = true;
    perm = (const uint32_t*)pp;
    return VPM_OK;
}

int stage_field(vpm_field f, int kernel, int precision, bool sort, Staged& out, Arena* ar)
{
    const int64_t np = f->np;
    const int64_t ld = ar ? np : f->cap;
    int rc;
    void *posp, *gp, *cp;
    if ((rc = take_buffer(f, f->pos, sizeof(double) * 3 * (size_t)ld, ar, posp, "positions"))) return rc;
    out.pos = (const double*)posp;
    out.ld = ld;
    out.gs = nullptr; out.cut = nullptr; out.g64 = nullptr; out.perm = nullptr;
    if (np == 0) return VPM_OK;
    if (precision == VPM_PREC_F64) {
        if ((rc = take_buffer(f, f->g64, sizeof(double4) * (size_t)ld, ar, gp, "sources"))) return rc;
        hipLaunchKernelGGL(pack_f64_kernel, dim3(nblocks(np)), dim3(kBlock), 0, f->stream,
                           f->P, np, (double*)posp, ld, (double4*)gp);
        VPM_LAUNCH_CHECK("pack_f64");
        out.g64 = (const double4*)gp;
        return VPM_OK;
    }
    if (precision == kPrecF64Far) {
        if ((rc = take_buffer(f, f->g64, sizeof(double4) * (size_t)ld, ar, gp, "sources"))) return rc;
        const uint32_t* perm = nullptr;
        if (sort && np > 1 && (rc = sort_field(f, perm, ar))) return rc;
        hipLaunchKernelGGL(pack_f64far_kernel, dim3(nblocks(np)), dim3(kBlock), 0, f->stream,
                           f->P, np, perm, (double*)posp, ld, (double4*)gp);
        VPM_LAUNCH_CHECK("pack_f64far");
        out.g64 = (const double4*)gp;
        out.perm = perm;
        return VPM_OK;
    }
    if ((rc = take_buffer(f, f->gs, sizeof(float4) * (size_t)ld, ar, gp, "sources"))) return rc;
    if ((rc = take_buffer(f, f->cut, sizeof(float) * (size_t)ld, ar, cp, "sources"))) return rc;
    const uint32_t* perm = nullptr;
    if (sort && np > 1 && (rc = sort_field(f, perm, ar))) return rc;
    switch (kernel) {
#define VPM_PACK(KID)                                                                                  \
    case KID:                                                                                          \
        hipLaunchKernelGGL(pack_f32c_kernel<KID>, dim3(nblocks(np)), dim3(kBlock), 0, f->stream, f->P,  \
                           np, perm, (double*)posp, ld, (float4*)gp, (float*)cp);                      \
        break;
        VPM_PACK(VPM_KERNEL_SINGULAR)
        VPM_PACK(VPM_KERNEL_GAUSSIAN)
        VPM_PACK(VPM_KERNEL_GAUSSIANERF)
        VPM_PACK(VPM_KERNEL_WINCKELMANS)
#undef VPM_PACK
    default:
        return fail(VPM_EINVAL, "unknown kernel id");
    }
    VPM_LAUNCH_CHECK("pack_f32c");
    out.gs = (const float4*)gp;
    out.cut = (const float*)cp;
    out.perm = perm;
    return VPM_OK;
}

// Number of source splits: about 32,768 blocks in all, chunks a multiple of
// the tile.  Blocks differ in near-field work, and many short blocks balance
// the CUs better than a few long ones: against 2,560 blocks, 32,768 run
// 11-25 % faster from N = 32,768 to 262,144, and at N = 2^20 (K = 8 instead
// of 1) 3-5 % faster; beyond that the gain is flat.  The FP64 partial sums
// then take at most ~1 GB (K x 14 x 8 B per target).
void choose_split(int64_t nt, int64_t ns, int64_t kmax, int64_t& K, int64_t& chunk)
{
    const int64_t nblk = (nt + kBlock - 1) / kBlock;
    const int64_t want = 32768;
    K = std::max<int64_t>(1, (want + nblk - 1) / nblk);
    if (kmax > 0) K = std::min(K, kmax);
    const int64_t tiles = std::max<int64_t>(1, (ns + kTile - 1) / kTile);
    K = std::min(K, tiles);
    const int64_t tiles_per = (tiles + K - 1) / K;
    chunk = tiles_per * kTile;
    K = (ns + chunk - 1) / chunk;
    if (K < 1) K = 1;
}

int launch_uj(vpm_field ft, const Staged& T, int64_t t0, int64_t nt, const Staged& S, int64_t s0,
              int64_t s1, int kernel, int precision, Arena* ar, int64_t kmax)
{
    if (nt <= 0 || s1 <= s0) return VPM_OK;
    int64_t K, chunk;
    choose_split(nt, s1 - s0, kmax, K, chunk);
    const int acc = precision == VPM_PREC_F64 ? kAccF64 : kAccF32;
    const int64_t pld = nt;
    int rc;
    void* pp;
    if ((rc = take_buffer(ft, ft->part, sizeof(double) * (size_t)acc * (size_t)K * (size_t)pld, ar, pp, "partial sums")))
        return rc;
    double* part = (double*)pp;
    const dim3 grid(nblocks(nt), (unsigned)K);
    timing_begin(ft);
    if (precision == VPM_PREC_F64) {
        switch (kernel) {
#define VPM_UJ64(KID)                                                                                 \
    case KID:                                                                                         \
        hipLaunchKernelGGL(uj_f64_kernel<KID>, grid, dim3(kBlock), 0, ft->stream, T.pos, T.ld, t0, nt, \
                           S.pos, S.ld, S.g64, s0, s1, chunk, part, pld);                             \
        break;
            VPM_UJ64(VPM_KERNEL_SINGULAR)
            VPM_UJ64(VPM_KERNEL_GAUSSIAN)
            VPM_UJ64(VPM_KERNEL_GAUSSIANERF)
            VPM_UJ64(VPM_KERNEL_WINCKELMANS)
#undef VPM_UJ64
        default:
            return fail(VPM_EINVAL, "unknown kernel id");
        }
        VPM_LAUNCH_CHECK("uj_f64");
        timing_end(ft, 1);
        hipLaunchKernelGGL(finalize_f64_kernel, dim3(nblocks(nt)), dim3(kBlock), 0, ft->stream, part, pld,
                           (int)K, nt, T.perm, t0, ft->P);
        VPM_LAUNCH_CHECK("finalize_f64");
        return VPM_OK;
    }
    if (precision == kPrecF64Far) {
        switch (kernel) {
#define VPM_UJ64F(KID)                                                                                    \
    case KID:                                                                                             \
        hipLaunchKernelGGL(uj_f64far_kernel<KID>, grid, dim3(kBlock), 0, ft->stream, T.pos, T.ld, t0, nt, \
                           S.pos, S.ld, S.g64, s0, s1, chunk, part, pld);                                 \
        break;
            VPM_UJ64F(VPM_KERNEL_SINGULAR)
            VPM_UJ64F(VPM_KERNEL_GAUSSIAN)
            VPM_UJ64F(VPM_KERNEL_GAUSSIANERF)
            VPM_UJ64F(VPM_KERNEL_WINCKELMANS)
#undef VPM_UJ64F
        default:
            return fail(VPM_EINVAL, "unknown kernel id");
        }
        VPM_LAUNCH_CHECK("uj_f64far");
        timing_end(ft, 1);
        hipLaunchKernelGGL(finalize_f32c_kernel, dim3(nblocks(nt)), dim3(kBlock), 0, ft->stream, part, pld,
                           (int)K, nt, T.perm, t0, ft->P);
        VPM_LAUNCH_CHECK("finalize_f32c");
        return VPM_OK;
    }
    switch (kernel) {
#define VPM_UJ32(KID)                                                                                  \
    case KID:                                                                                          \
        hipLaunchKernelGGL(uj_f32c_kernel<KID>, grid, dim3(kBlock), 0, ft->stream, T.pos, T.ld, t0, nt, \
                           S.pos, S.ld, S.gs, S.cut, s0, s1, chunk, part, pld);                        \
        break;
        VPM_UJ32(VPM_KERNEL_SINGULAR)
        VPM_UJ32(VPM_KERNEL_GAUSSIAN)
        VPM_UJ32(VPM_KERNEL_GAUSSIANERF)
        VPM_UJ32(VPM_KERNEL_WINCKELMANS)
#undef VPM_UJ32
    default:
        return fail(VPM_EINVAL, "unknown kernel id");
    }
    VPM_LAUNCH_CHECK("uj_f32c");
    timing_end(ft, 0);
    hipLaunchKernelGGL(finalize_f32c_kernel, dim3(nblocks(nt)), dim3(kBlock), 0, ft->stream, part, pld,
                       (int)K, nt, T.perm, t0, ft->P);
    VPM_LAUNCH_CHECK("finalize_f32c");
    return VPM_OK;
}

}  // namespace vpm
