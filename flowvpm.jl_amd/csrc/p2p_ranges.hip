// p2p_ranges.hip -- U/J over a list of (target range, source range) pairs:
// FastMultipole's leaf-pair P2P (fmm.direct! per direct-list entry,
// src/FLOWVPM_fmm.jl:102-168), batched into one launch on one device copy.
//
// The host planner (range_plan.cpp) cuts the target index space into
// disjoint elementary intervals, each with the source ranges of every pair
// covering it, in pair order.  A block (or, for FMM leaves, a wave) owns the
// targets of one item, streams its source ranges through LDS, and adds the
// result onto its own targets' records: no two blocks write the same
// particle, and each target's sum runs in pair order.  The pair arithmetic
// is p2p_pair.h's.  A plan can be made once and served to U/J, E_str and
// zeta (vpm_plan_ranges).
#include <algorithm>
#include <string>

#include "p2p_pair.h"
#include "range_plan.h"

namespace vpm {


// Wave-wide inclusive scans over the 64 lanes by DPP: row shifts by 1, 2, 4
// and 8 within each 16-lane row, then the last lane of row 0 and of rows 0-1
// broadcast into the rows above (row_bcast:15 into rows 1 and 3, row_bcast:31
// into rows 2 and 3).  No LDS round trip; lanes a shift leaves without a
// source take 0, which neither sum nor max (of values >= 0) sees.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_src(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xf, false);
}

__device__ __forceinline__ int wave_scan_add(int v)
{
    v += dpp_src<0x111, 0xf>(v);
    v += dpp_src<0x112, 0xf>(v);
    v += dpp_src<0x114, 0xf>(v);
    v += dpp_src<0x118, 0xf>(v);
    v += dpp_src<0x142, 0xa>(v);
    v += dpp_src<0x143, 0xc>(v);
    return v;
}

__device__ __forceinline__ int wave_scan_max(int v)   // v >= 0
{
    v = max(v, dpp_src<0x111, 0xf>(v));
    v = max(v, dpp_src<0x112, 0xf>(v));
    v = max(v, dpp_src<0x114, 0xf>(v));
    v = max(v, dpp_src<0x118, 0xf>(v));
    v = max(v, dpp_src<0x142, 0xa>(v));
    v = max(v, dpp_src<0x143, 0xc>(v));
    return v;
}

// The next <= 64 sources of an item's source stream (its ranges in list
// order, concatenated): lane m gets stream slot m in j (0 past `filled`),
// and (l, off) -- the current range and the offset in it, wave-uniform --
// move past them.  Up to 64 ranges are read at once, one per lane; an
// inclusive wave scan of their lengths (each capped at 64) gives every
// range its first slot B_i.  Slot m belongs to the last range with B_i <=
// m: each nonempty range marks its first slot with its index (an LDS max,
// so of the ranges that start at one slot the last, nonempty one wins) and
// a max-scan over the slots carries the marks forward.  `mark` is 64 ints
// of the wave's own LDS, free at the call.  Returns filled, which is 0 only
// when the ranges read were all empty (then l has moved past them).  The
// lane-shared kernels use it at R = 8, where a 64-source chunk spans ~8
// ranges of ~8 sources: against the range-at-a-time walk (a dependent scalar
// load and ~30 scalar instructions per range) it took the 2^20 jet's
// 8-particle leaf lists from 1.58 to 1.22 ms (E_str), 1.20 to 0.88 ms
// (zeta) and 2.33 to 2.20 ms (U/J); at R = 4 (~4 ranges per chunk) U/J was
// 3 % slower with it, so R < 8 keeps the walk.
__device__ __forceinline__ int stream_fill(const int2* __restrict__ lsrc, int64_t& l, int64_t l_end, int64_t& off,
                                           int lane, int64_t& j, int* mark)
{
    const int nr = (int)min<int64_t>(64, l_end - l);   // wave-uniform, >= 1
    int2 rs = make_int2(0, 0);
    if (lane < nr) rs = lsrc[l + lane];
    const int b = rs.x + (lane == 0 ? (int)off : 0);
    const int len = lane < nr ? rs.y - b : 0;
    const int lc = min(len, 64);
    const int S = wave_scan_add(lc);
    const int B = S - lc;
    const int total = __builtin_amdgcn_readlane(S, 63);
    const int filled = min(total, 64);
    // `mark` aliases the staging rows the previous steps read, under
    // another type: wave-scope fences keep every access to it in order
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    mark[lane] = 0;
    if (lc > 0 && B < 64) atomicMax(&mark[B], lane);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int o = wave_scan_max(mark[lane]);
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int Bo = __shfl(B, o), bo = __shfl(b, o);
    j = lane < filled ? (int64_t)bo + (lane - Bo) : 0;
    if (total < 64) {   // no length was capped: every range read is used up
        l += nr;
        off = 0;
    } else {            // slot 63's range ends in this chunk or continues
        const int o63 = __builtin_amdgcn_readlane(o, 63);
        const int B63 = __builtin_amdgcn_readlane(B, o63);
        if (B63 + __builtin_amdgcn_readlane(len, o63) == 64) {
            l += o63 + 1;
            off = 0;
        } else {
            off = (o63 == 0 ? off : 0) + (64 - B63);
            l += o63;
        }
    }
    return filled;
}

// Component-wise select of two float4 values (kept in registers: a ?: on
// the vector type itself can go through private memory).
__device__ __forceinline__ float4 sel4(bool c, float4 a, float4 b)
{
    return make_float4(c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z, c ? a.w : b.w);
}

template <int K>
__global__ __launch_bounds__(kBlock) void ranges_f32c_kernel(const RangeItem* __restrict__ items,
                                                             const int2* __restrict__ lsrc,
                                                             const double* __restrict__ pos, int64_t ld,
                                                             const float4* __restrict__ gs,
                                                             const float* __restrict__ cut,
                                                             double* __restrict__ P)
{
    __shared__ float4 lpos[kTile];
    __shared__ float4 lgs[kTile];
    __shared__ float4 llo[kTile];
    const RangeItem it = items[xcd_block(blockIdx.x, gridDim.x)];
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
        const int2 rs = lsrc[l];
        const int64_t sb = rs.x, se = rs.y;
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
            tile_f32c<K>(n, lpos, lgs, llo, tx, ty, tz, tlx, tly, tlz, box, acc, sum, err);
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

// F32C over short target intervals (FMM leaves): each wave owns one item of
// at most 64 / R targets and stages sources through its own 64-row LDS slot,
// with no block-wide barrier (within a wave LDS accesses complete in issue
// order; the scheduling barriers keep the compiler from moving the next
// slot's stores above this slot's loads).  R = 1: one lane per target, each
// source range in turn through the all-pairs tile code.  R > 1 (leaves of
// <= 32 particles): each target is replicated over R lane groups, and group
// g takes every R-th source (g, g + R, ...) of the item's source stream --
// its ranges in list order, concatenated and staged 64 at a time -- so small
// leaves still keep all 64 lanes busy.
// A step of R sources is far when all R are beyond the wave's target box by
// the far radius (then every pair has r > 0), near otherwise.  Per-lane FP32
// sums fold into TwoSum pairs, the U partials after every 64-source chunk
// (64 / R steps), the rest every ~64 pairs; at the end the R groups' FP64
// values are added in a fixed butterfly and group 0 writes.
// R > 1: six waves per SIMD (80 VGPRs; R = 4 and 8 spill 12 bytes per lane,
// touched only around the per-chunk folds) instead of the five that 81-84
// VGPRs allow, 3 % faster on 16-particle leaves.  R = 1 would spill at five
// and keeps four.
template <int K, int R>
__global__ __launch_bounds__(kBlock, R > 1 ? 6 : 4) void ranges_wave_f32c_kernel(const RangeItem* __restrict__ items,
                                                                  int64_t nitems,
                                                                  const int2* __restrict__ lsrc,
                                                                  const double* __restrict__ pos, int64_t ld,
                                                                  const float4* __restrict__ gs,
                                                                  const float* __restrict__ cut,
                                                                  double* __restrict__ P)
{
    static_assert(R == 1 || R == 2 || R == 4 || R == 8, "replication");
    constexpr int TW = 64 / R;                  // targets per wave
    __shared__ float4 lpos_all[kBlock / 64][64];
    __shared__ float4 lgs_all[kBlock / 64][64];
    __shared__ float4 llo_all[kBlock / 64][64];
    const int lane = threadIdx.x & 63;
    // wave-uniform (kept in scalar registers, so the item and its source
    // ranges are read with scalar loads and the stream bookkeeping is SALU)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t w = xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + wave;
    if (w >= nitems) return;   // no block barrier below
    float4* lpos = lpos_all[wave];
    float4* lgs = lgs_all[wave];
    float4* llo = llo_all[wave];
    const RangeItem it = items[w];
    const int64_t ib = it.t_begin;
    const double ox = pos[ib], oy = pos[ld + ib], oz = pos[2 * ld + ib];
    const int g = lane / TW;
    const int64_t t = ib + (lane % TW);
    const bool active = t < it.t_end;
    const int64_t tt = active ? t : ib;
    float tx, ty, tz, tlx, tly, tlz;
    split_rel(pos[tt], ox, tx, tlx);
    split_rel(pos[ld + tt], oy, ty, tly);
    split_rel(pos[2 * ld + tt], oz, tz, tlz);
    const WaveBox box = wave_box(tx, ty, tz);
    if constexpr (R == 1) {
        // one lane per target: each source range in turn, 64 sources at a
        // time, through the all-pairs tile code; a fold per chunk
        float sum[kAccF32], err[kAccF32];
#pragma unroll
        for (int c = 0; c < kAccF32; ++c) { sum[c] = 0.f; err[c] = 0.f; }
        for (int64_t l = it.l_begin; l < it.l_end; ++l) {
            const int2 rs = lsrc[l];
            const int64_t sb = rs.x, se = rs.y;
            for (int64_t base = sb; base < se; base += 64) {
                const int n = (int)((se - base) < 64 ? (se - base) : 64);
                __builtin_amdgcn_wave_barrier();
                if (lane < n) {
                    const int64_t j = base + lane;
                    float4 p, q;
                    split_rel(pos[j], ox, p.x, q.x);
                    split_rel(pos[ld + j], oy, p.y, q.y);
                    split_rel(pos[2 * ld + j], oz, p.z, q.z);
                    p.w = cut[j];
                    q.w = 0.f;
                    lpos[lane] = p;
                    llo[lane] = q;
                    lgs[lane] = gs[j];
                }
                __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                __builtin_amdgcn_wave_barrier();
                __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                float acc[kAccF32];
#pragma unroll
                for (int c = 0; c < kAccF32; ++c) acc[c] = 0.f;
                tile_f32c<K, 1>(n, lpos, lgs, llo, tx, ty, tz, tlx, tly, tlz, box, acc, sum, err);
                fold_twosum(acc, sum, err);
            }
        }
        if (active) {
            double a[kAccF32];
#pragma unroll
            for (int c = 0; c < kAccF32; ++c) a[c] = (double)sum[c] + (double)err[c];
            add_f32c_to_record(P + t * NF, a);
        }
        return;
    }
    float sum[kAccF32], err[kAccF32], acc[kAccF32];
#pragma unroll
    for (int c = 0; c < kAccF32; ++c) { sum[c] = 0.f; err[c] = 0.f; acc[c] = 0.f; }
    int since = 0;                              // pairs per lane since the last fold (wave-uniform)
    int64_t l = it.l_begin, off = 0;            // stream position: range l, offset in it
    while (l < it.l_end) {
        // the next <= 64 sources of the stream; lane m takes stream slot m
        int filled = 0;
        int64_t j = 0;
        if constexpr (R == 8) {
            filled = stream_fill(lsrc, l, it.l_end, off, lane, j, (int*)lpos);
            if (filled == 0) continue;          // only empty ranges were read
        } else {
            while (filled < 64 && l < it.l_end) {
                const int2 rs = lsrc[l];
                const int64_t len = (int64_t)rs.y - rs.x;
                const int take = (int)min<int64_t>(len - off, (int64_t)(64 - filled));
                if (lane >= filled && lane < filled + take) j = rs.x + off + (lane - filled);
                filled += take;
                off += take;
                if (off == len) { ++l; off = 0; }
            }
            if (filled == 0) break;                 // only empty ranges were left
        }
        __builtin_amdgcn_wave_barrier();
        {
            // slots past `filled` hold a zero-strength source far away, so
            // every lane evaluates its step's pair unconditionally: it adds
            // exact zeros (F' = H'' = 0 at 1/sigma = 0, and d x 0 = 0), and
            // no lane-dependent branch splits the accumulator updates.  Those
            // lanes load source j = 0 (a valid index) and select the padding.
            const bool real = lane < filled;
            float4 p, q;
            split_rel(pos[j], ox, p.x, q.x);
            split_rel(pos[ld + j], oy, p.y, q.y);
            split_rel(pos[2 * ld + j], oz, p.z, q.z);
            const float c = cut[j];
            const float4 gr = gs[j];
            p = sel4(real, make_float4(p.x, p.y, p.z, c), make_float4(1e15f, 1e15f, 1e15f, 0.f));
            q = sel4(real, make_float4(q.x, q.y, q.z, 0.f), make_float4(0.f, 0.f, 0.f, 0.f));
            lpos[lane] = p;
            llo[lane] = q;
            lgs[lane] = sel4(real, gr, make_float4(0.f, 0.f, 0.f, 0.f));
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        {
            bool nr = false;
            if (lane < filled) {
                const float4 a = lpos[lane];
                const float ex = fmaxf(fmaxf(box.lo[0] - a.x, a.x - box.hi[0]), 0.f);
                const float ey = fmaxf(fmaxf(box.lo[1] - a.y, a.y - box.hi[1]), 0.f);
                const float ez = fmaxf(fmaxf(box.lo[2] - a.z, a.z - box.hi[2]), 0.f);
                // 1e-4 margin covers the FP32 rounding of the box test
                nr = fmaf(ez, ez, fmaf(ey, ey, ex * ex)) < a.w * 1.0001f;
            }
            const uint64_t nearm = __ballot(nr);
            // k R < filled <= 64 and 64 is a multiple of R, so sidx < 64
            for (int k = 0; k * R < filled; ++k) {
                const int sidx = k * R + g;
                if ((nearm >> (k * R)) & ((1ull << R) - 1))
                    near_pair<K>(tx, ty, tz, tlx, tly, tlz, lpos[sidx], lgs[sidx], llo[sidx], acc);
                else
                    far_pair<K>(tx, ty, tz, lpos[sidx], lgs[sidx], acc);
            }
            // a leaf's near-field U nearly cancels (its neighbourhood
            // balances): the U partials fold after every chunk (64 / R
            // steps), the rest after ~64.  With all 14 folded every 64
            // steps, the 2^20 jet's 16-particle leaf list lost up to 2.5e-5
            // of U; folded per chunk 8.0e-6, at the same speed
            fold_u(acc, sum, err);
        }
        since += (filled + R - 1) / R;
        if (since >= 64) {
            fold_twosum(acc, sum, err);
#pragma unroll
            for (int c = 0; c < kAccF32; ++c) acc[c] = 0.f;
            since = 0;
        }
    }
    fold_twosum(acc, sum, err);
    double a[kAccF32];
#pragma unroll
    for (int c = 0; c < kAccF32; ++c) a[c] = (double)sum[c] + (double)err[c];
    if constexpr (R > 1) {
#pragma unroll
        for (int m = TW; m < 64; m <<= 1) {
#pragma unroll
            for (int c = 0; c < kAccF32; ++c) a[c] += __shfl_xor(a[c], m);
        }
    }
    if (g == 0 && active) add_f32c_to_record(P + t * NF, a);
}

#pragma clang fp contract(off)
// NT lanes per item: kBlock, or 64 (one wave) when every item of the plan
// has at most 64 targets; each target sums its sources in list order either way.
template <int K, int NT>
__global__ __launch_bounds__(NT) void ranges_f64_kernel(const RangeItem* __restrict__ items,
                                                            const int2* __restrict__ lsrc,
                                                            const double* __restrict__ pos, int64_t ld,
                                                            const double4* __restrict__ g64,
                                                            double* __restrict__ P)
{
    __shared__ double4 lpos[NT];
    __shared__ double4 lgam[NT];
    const RangeItem it = items[xcd_block(blockIdx.x, gridDim.x)];
    const int tid = threadIdx.x;
    const int64_t t = it.t_begin + tid;
    const bool active = t < it.t_end;
    const int64_t tt = active ? t : it.t_begin;
    const double tx = pos[tt], ty = pos[ld + tt], tz = pos[2 * ld + tt];
    double acc[kAccF64];
#pragma unroll
    for (int c = 0; c < kAccF64; ++c) acc[c] = 0.0;
    for (int64_t l = it.l_begin; l < it.l_end; ++l) {
        const int2 rs = lsrc[l];
        const int64_t sb = rs.x, se = rs.y;
        for (int64_t base = sb; base < se; base += NT) {
            const int n = (int)((se - base) < NT ? (se - base) : NT);
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
// FP64 U/J over a wave plan whose items hold at most 64 / G targets: G items
// per wave, each in its own group of 64 / G lanes with its own LDS rows,
// streaming its own source ranges (groups run their loops under the exec
// mask).  Every target sums its sources in list order, the order of the
// block kernel, so the results are the same bit for bit.
template <int K, int G>
__global__ __launch_bounds__(kBlock) void ranges_f64_group_kernel(const RangeItem* __restrict__ items,
                                                                  int64_t nitems, const int2* __restrict__ lsrc,
                                                                  const double* __restrict__ pos, int64_t ld,
                                                                  const double4* __restrict__ g64,
                                                                  double* __restrict__ P)
{
    constexpr int NG = 64 / G;   // lanes and LDS rows per item
    __shared__ double4 lpos_all[kBlock / 64][64];
    __shared__ double4 lgam_all[kBlock / 64][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane / NG, k = lane % NG;
    double4* lpos = lpos_all[wave] + g * NG;
    double4* lgam = lgam_all[wave] + g * NG;
    const int64_t item = (xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + wave) * G + g;
    const RangeItem it = item < nitems ? items[item] : RangeItem{0, 0, 0, 0};
    const int64_t t = it.t_begin + k;
    const bool active = t < it.t_end;
    const int64_t tt = active ? t : 0;
    const double tx = pos[tt], ty = pos[ld + tt], tz = pos[2 * ld + tt];
    double acc[kAccF64];
#pragma unroll
    for (int c = 0; c < kAccF64; ++c) acc[c] = 0.0;
    int64_t l = it.l_begin, base = 0, se = 0;
    if (l < it.l_end) { const int2 rs = lsrc[l]; base = rs.x; se = rs.y; }
    for (;;) {
        while (l < it.l_end && base >= se) {   // the group's next nonempty range
            if (++l < it.l_end) { const int2 rs = lsrc[l]; base = rs.x; se = rs.y; }
        }
        const bool work = l < it.l_end;
        if (!__any(work)) break;
        const int n = work ? (int)((se - base) < NG ? (se - base) : NG) : 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (k < n) {
            const int64_t j = base + k;
            const double4 gg = g64[j];
            lpos[k] = make_double4(pos[j], pos[ld + j], pos[2 * ld + j], gg.w);
            lgam[k] = make_double4(gg.x, gg.y, gg.z, 0.0);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int s = 0; s < n; ++s) pair_f64<K>(tx, ty, tz, lpos[s], lgam[s], acc);
        base += n;
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

// ------------------------------------------------------------- device plans
// A plan made by the host planner: items written straight into the field's
// pinned staging buffer (items first, then source ranges), copied into a
// device buffer with the same layout.  On leaf lists the packed source
// ranges are copied in pieces as the planner's workers write them, so the
// upload runs behind the packing.  The pinned source stays in use until the
// field's stream is synchronized, which every entry point does before it
// returns (also on failure: copies of a list that then fails its checks land
// in the device buffer only).
namespace {

struct FieldSink final : PlanSink {
    vpm_field f;
    DevBuf* dst;
    int64_t nitems = 0, nsrc = 0;

    FieldSink(vpm_field f_, DevBuf* d) : f(f_), dst(d) {}

    size_t items_bytes() const { return sizeof(RangeItem) * (size_t)nitems; }
    size_t bytes() const { return items_bytes() + sizeof(SrcRange) * (size_t)nsrc; }

    int stage(int64_t ni, int64_t ns, RangeItem*& items, SrcRange*& src) override
    {
        nitems = ni;
        nsrc = ns;
        if (int rc = ensure_pinned(f->hstage, bytes())) return rc;
        items = (RangeItem*)f->hstage.p;
        src = (SrcRange*)(items + ni);
        return VPM_OK;
    }
    int prepare_copies(int64_t, int64_t) override { return ensure(*dst, bytes() + 64); }
    int copy_piece(int64_t a, int64_t b) override
    {
        const char* hsrc = (const char*)f->hstage.p + items_bytes();
        char* dsrc = (char*)dst->p + items_bytes();
        const size_t off = sizeof(SrcRange) * (size_t)a;
        const hipError_t e = hipMemcpyAsync(dsrc + off, hsrc + off, sizeof(SrcRange) * (size_t)(b - a),
                                            hipMemcpyHostToDevice, f->stream);
        return e == hipSuccess ? 0 : (int)e;
    }
    int report(int code) override { return hip_fail((hipError_t)code, "hipMemcpyAsync(range lists)"); }
};

}  // namespace

int make_dev_plan(vpm_field f, const int64_t* tb, const int64_t* te, const int64_t* sb, const int64_t* se,
                  int64_t npairs, bool allow_wave, int rmax, DevBuf& dst, DevPlan& out)
{
    out = DevPlan();
    FieldSink sink(f, &dst);
    PlanShape shape;
    if (int rc = plan_ranges(f->np, tb, te, sb, se, npairs, allow_wave, rmax, sink, shape)) return rc;
    if (shape.nitems == 0) return VPM_OK;
    if (int rc = ensure(dst, sink.bytes() + 64)) return rc;   // the size the copies already ensured, if any
    char* base = (char*)dst.p;
    const hipError_t e = hipMemcpyAsync(base, f->hstage.p, shape.src_copied ? sink.items_bytes() : sink.bytes(),
                                        hipMemcpyHostToDevice, f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(range lists)");
    out.items = (const RangeItem*)base;
    out.src = (const int2*)(base + sink.items_bytes());
    out.nitems = shape.nitems;
    out.wave = shape.wave;
    out.R = shape.R;
    return VPM_OK;
}

// U/J of a planned list (F32C wave kernels for wave plans, the block kernels
// otherwise; F64 takes its block kernel, with one wave per item on wave plans,
// whose items have at most 64 targets).
int launch_p2p_plan(vpm_field f, int kernel, int precision, const DevPlan& dp)
{
    if (dp.nitems == 0) return VPM_OK;
    Staged st;
    if (int rc = stage_field(f, kernel, precision, false, st)) return rc;
    const RangeItem* d_items = dp.items;
    const int2* d_src = dp.src;
    const dim3 grid((unsigned)dp.nitems);
    timing_begin(f);
    if (precision == VPM_PREC_F64) {
        switch (kernel) {
#define VPM_RG64(KID, G)                                                                                          \
    hipLaunchKernelGGL((ranges_f64_group_kernel<KID, G>),                                                         \
                       dim3((unsigned)((dp.nitems + G * (kBlock / 64) - 1) / (G * (kBlock / 64)))), dim3(kBlock), \
                       0, f->stream, d_items, dp.nitems, d_src, st.pos, st.ld, st.g64, f->P)
#define VPM_R64(KID)                                                                                              \
    case KID:                                                                                                     \
        if (dp.wave && dp.R > 1) {                                                                                \
            if (dp.R == 8) VPM_RG64(KID, 8);                                                                      \
            else if (dp.R == 4) VPM_RG64(KID, 4);                                                                 \
            else VPM_RG64(KID, 2);                                                                                \
        } else if (dp.wave) {                                                                                     \
            hipLaunchKernelGGL((ranges_f64_kernel<KID, 64>), grid, dim3(64), 0, f->stream, d_items, d_src,        \
                               st.pos, st.ld, st.g64, f->P);                                                      \
        } else {                                                                                                  \
            hipLaunchKernelGGL((ranges_f64_kernel<KID, kBlock>), grid, dim3(kBlock), 0, f->stream, d_items,       \
                               d_src, st.pos, st.ld, st.g64, f->P);                                               \
        }                                                                                                         \
        break;
            VPM_R64(VPM_KERNEL_SINGULAR) VPM_R64(VPM_KERNEL_GAUSSIAN) VPM_R64(VPM_KERNEL_GAUSSIANERF) VPM_R64(VPM_KERNEL_WINCKELMANS)
#undef VPM_R64
#undef VPM_RG64
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    } else if (dp.wave) {
        const int64_t ni = dp.nitems;
        const dim3 wgrid((unsigned)((ni + kBlock / 64 - 1) / (kBlock / 64)));
        switch (kernel) {
#define VPM_RWR(KID, R) hipLaunchKernelGGL((ranges_wave_f32c_kernel<KID, R>), wgrid, dim3(kBlock), 0, f->stream, d_items, ni, d_src, st.pos, st.ld, st.gs, st.cut, f->P)
#define VPM_RW(KID)                                                                   \
    case KID:                                                                         \
        switch (dp.R) {                                                               \
        case 8: VPM_RWR(KID, 8); break;                                               \
        case 4: VPM_RWR(KID, 4); break;                                               \
        case 2: VPM_RWR(KID, 2); break;                                               \
        default: VPM_RWR(KID, 1); break;                                              \
        }                                                                             \
        break;
            VPM_RW(VPM_KERNEL_SINGULAR) VPM_RW(VPM_KERNEL_GAUSSIAN) VPM_RW(VPM_KERNEL_GAUSSIANERF) VPM_RW(VPM_KERNEL_WINCKELMANS)
#undef VPM_RW
#undef VPM_RWR
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    } else {
        switch (kernel) {
#define VPM_R32(KID) case KID: hipLaunchKernelGGL(ranges_f32c_kernel<KID>, grid, dim3(kBlock), 0, f->stream, d_items, d_src, st.pos, st.ld, st.gs, st.cut, f->P); break;
            VPM_R32(VPM_KERNEL_SINGULAR) VPM_R32(VPM_KERNEL_GAUSSIAN) VPM_R32(VPM_KERNEL_GAUSSIANERF) VPM_R32(VPM_KERNEL_WINCKELMANS)
#undef VPM_R32
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "p2p_ranges kernel");
    timing_end(f, precision == VPM_PREC_F64 ? 1 : 0);
    return VPM_OK;
}

// Short target intervals (FMM leaves, mean <= 128 targets): one wave per
// <= 64 / R targets (F32C) or <= 64 targets (FP64); otherwise one block per
// <= 256 targets.
int launch_p2p_ranges(vpm_field f, int kernel, int precision, const int64_t* tb, const int64_t* te,
                      const int64_t* sb, const int64_t* se, int64_t npairs)
{
    DevPlan dp;
    if (int rc = make_dev_plan(f, tb, te, sb, se, npairs, true, 8, f->rng, dp)) return rc;
    return launch_p2p_plan(f, kernel, precision, dp);
}

// Source records of the F32C list kernels for E_str and zeta, packed once
// per call.  The list kernels read each source once for every target leaf
// that lists it (about 140 times on FMM leaf lists); gathering 48 (zeta) or
// 72 (E_str) contiguous bytes per source instead of parts of its 368-byte
// particle record cuts that traffic threefold.  Per particle: the position
// relative to particle 0 as an FP32 hi + lo pair, with the cut and a beside
// it, (Gamma, b) in FP32, and for E_str J^T Gamma (transposed) or J Gamma
// in FP64, formed with Gamma already rounded to FP32, so that the kernel's
// J_ref^T Gamma subtracts the same Gamma and (J_j - J_ref)^T Gamma keeps
// FP64 accuracy whatever |J_ref|.  With the positions split once, staging
// a source (or a target) is a copy: no FP64 work per item.
struct ListSrc {
    const float4* hi = nullptr;    // x - x_0 in FP32 (hi part), cut
    const float4* lo = nullptr;    // the FP32 remainder of x - x_0 (lo part), a
    const float4* gb = nullptr;    // Gamma, b
    const double* jg = nullptr;    // [3][ld] (E_str only)
    int64_t ld = 0;
};

// E_str's zeta_sigma = b exp2(a r^2) (Gaussian-erf), b exp2(a r^3)
// (Gaussian), b (1 + a r^2)^-5/2 (Winckelmans), and the r^2 past which it is
// below FP32 resolution of the sum (q = 7, 3; none).
template <int K>
__device__ __forceinline__ void estr_params(double sg, double& a, double& b, double& cut)
{
    const double l2e = 1.4426950408889634;
    const double si = 1.0 / sg, s2 = si * si, s3 = s2 * si;
    if (K == VPM_KERNEL_GAUSSIANERF) { a = -0.5 * l2e * s2; b = kConst1 * s3; cut = 49.0 * sg * sg; }
    else if (K == VPM_KERNEL_GAUSSIAN) { a = -l2e * s3; b = kConst3 * s3; cut = 9.0 * sg * sg; }
    else { a = s2; b = kConst4 * 7.5 * s3; cut = (double)FLT_MAX; }
}

// MODE 0: zeta; 1: E_str with J^T Gamma; 2: E_str with J Gamma.
template <int K, int MODE>
__global__ __launch_bounds__(kBlock) void pack_list_src_kernel(const double* __restrict__ P, int64_t np,
                                                               float4* __restrict__ hi, float4* __restrict__ lo,
                                                               float4* __restrict__ gb, double* __restrict__ jg,
                                                               int64_t ld)
{
    const int64_t j = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (j >= np) return;
    const double* sr = P + j * NF;
    const double sg = sr[RSIG];
    double a, b;
    float cut;
    if constexpr (MODE == 0) {
        zeta_params<K>(sg, a, b);
        cut = zeta_cut<K>(sg);
    } else {
        double c;
        estr_params<K>(sg, a, b, c);
        cut = (float)c;
    }
    // positions relative to particle 0, split into FP32 hi + lo: for a pair
    // (t, s) the difference of the hi parts is exact whenever |t - s| is
    // below the hi parts themselves (Sterbenz), so d = (t_hi - s_hi) +
    // (t_lo - s_lo) keeps FP32 relative accuracy at any distance, as with
    // an origin per item
    float4 h, l;
    split_rel(sr[RX], P[RX], h.x, l.x);
    split_rel(sr[RX + 1], P[RX + 1], h.y, l.y);
    split_rel(sr[RX + 2], P[RX + 2], h.z, l.z);
    h.w = cut;
    l.w = (float)a;
    hi[j] = h;
    lo[j] = l;
    const float G0 = (float)sr[RG], G1 = (float)sr[RG + 1], G2 = (float)sr[RG + 2];
    gb[j] = make_float4(G0, G1, G2, (float)b);
    if constexpr (MODE != 0) {
        const double* J = sr + RJ;
        const double g0 = G0, g1 = G1, g2 = G2;
#pragma unroll
        for (int c = 0; c < 3; ++c)
            jg[c * ld + j] = MODE == 1 ? J[3 * c] * g0 + J[3 * c + 1] * g1 + J[3 * c + 2] * g2
                                       : J[c] * g0 + J[3 + c] * g1 + J[6 + c] * g2;
    }
}

// Packs the source records of the F32C E_str (mode 1, 2) and zeta (mode 0)
// list kernels on the field's stream.
int pack_list_src(vpm_field f, int kernel, int mode, ListSrc& out)
{
    const int64_t np = f->np, ld = f->cap;
    const size_t f4 = sizeof(float4) * (size_t)ld;
    const size_t jg = mode ? sizeof(double) * 3 * (size_t)ld : 0;
    if (int rc = ensure(f->lrec, 3 * f4 + jg)) return rc;
    char* base = (char*)f->lrec.p;
    float4* h4 = (float4*)base;
    float4* l4 = (float4*)(base + f4);
    float4* g4 = (float4*)(base + 2 * f4);
    double* j3 = mode ? (double*)(base + 3 * f4) : nullptr;
    out.hi = h4;
    out.lo = l4;
    out.gb = g4;
    out.jg = j3;
    out.ld = ld;
    if (np == 0) return VPM_OK;
    const dim3 grid((unsigned)((np + kBlock - 1) / kBlock));
#define VPM_PK(KID, M) hipLaunchKernelGGL((pack_list_src_kernel<KID, M>), grid, dim3(kBlock), 0, f->stream, f->P, np, h4, l4, g4, j3, ld)
#define VPM_PKK(KID)                         \
    case KID:                                \
        if (mode == 1) VPM_PK(KID, 1);       \
        else if (mode == 2) VPM_PK(KID, 2);  \
        else VPM_PK(KID, 0);                 \
        break;
    switch (kernel) {
        VPM_PKK(VPM_KERNEL_GAUSSIAN) VPM_PKK(VPM_KERNEL_GAUSSIANERF) VPM_PKK(VPM_KERNEL_WINCKELMANS)
    default: return fail(VPM_EINVAL, "unknown kernel id");
    }
#undef VPM_PKK
#undef VPM_PK
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "pack_list_src kernel");
    return VPM_OK;
}

// ------------------------------------------------------------ E_str ranges
// Estr_fmm! over a direct list (src/FLOWVPM_subfilterscale_models.jl:157-188):
// E_str pairs (:16-41) for every (target range, source range), no static skip,
// accumulated onto SFS.
//
// F32C: within a block (<= 256 targets of one interval), J' = J - J_ref with
// J_ref the J of the block's first target (FP64), and
//     sum_j zeta_j (J'_i - J'_j)^T Gamma_j = J'_i^T A - B,
//     A = sum zeta Gamma, B = sum zeta J'_j^T Gamma_j,
// the packed sources of pack_list_src (hi/lo positions relative to particle
// 0, J^T Gamma in FP64), sources beyond q_far skipped per wave (Gaussian-erf,
// Gaussian), sums folded into TwoSum every 32 sources.
template <int K>
__global__ __launch_bounds__(kBlock) void estr_ranges_f32c_kernel(const RangeItem* __restrict__ items,
                                                                  const int2* __restrict__ lsrc,
                                                                  const ListSrc S,
                                                                  double* __restrict__ P, int transposed)
{
    __shared__ float4 lp[kTile];   // s_hi, cut
    __shared__ float4 lg[kTile];   // Gamma, b
    __shared__ float4 lq[kTile];   // J'^T Gamma (or J' Gamma), a
    __shared__ float4 ll[kTile];   // s_lo
    const RangeItem it = items[xcd_block(blockIdx.x, gridDim.x)];
    const int tid = threadIdx.x;
    const double* r0 = P + it.t_begin * NF;
    double Jr[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) Jr[c] = r0[RJ + c];
    const int64_t t = it.t_begin + tid;
    const bool active = t < it.t_end;
    const int64_t tt = active ? t : it.t_begin;
    const float4 th = S.hi[tt], tl = S.lo[tt];
    const float tx = th.x, ty = th.y, tz = th.z, tlx = tl.x, tly = tl.y, tlz = tl.z;
    float sum[6], err[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) { sum[c] = 0.f; err[c] = 0.f; }
    for (int64_t l = it.l_begin; l < it.l_end; ++l) {
        const int2 rs = lsrc[l];
        const int64_t sb = rs.x, se = rs.y;
        for (int64_t base = sb; base < se; base += kTile) {
            const int n = (int)((se - base) < kTile ? (se - base) : kTile);
            __syncthreads();
            if (tid < n) {
                const int64_t j = base + tid;
                const float4 lo = S.lo[j];    // s_lo, a
                const float4 gq = S.gb[j];    // Gamma, b
                // (J_j - J_ref)^T Gamma_j (or (J_j - J_ref) Gamma_j)
                const double G0 = gq.x, G1 = gq.y, G2 = gq.z;
                double q3[3];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const double ref = transposed ? Jr[3 * c] * G0 + Jr[3 * c + 1] * G1 + Jr[3 * c + 2] * G2
                                                  : Jr[c] * G0 + Jr[3 + c] * G1 + Jr[6 + c] * G2;
                    q3[c] = S.jg[c * S.ld + j] - ref;
                }
                lp[tid] = S.hi[j];            // s_hi, cut
                ll[tid] = lo;
                lq[tid] = make_float4((float)q3[0], (float)q3[1], (float)q3[2], lo.w);
                lg[tid] = gq;
            }
            __syncthreads();
            for (int s0 = 0; s0 < n; s0 += 32) {
                const int s1 = min(s0 + 32, n);
                float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
                for (int s = s0; s < s1; ++s) {
                    const float4 a = lp[s];
                    float dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
                    float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    if (__any(r2 < a.w)) {
                        const float4 c = lq[s];
                        const float4 lo = ll[s];
                        const float4 g = lg[s];
                        dx += tlx - lo.x;
                        dy += tly - lo.y;
                        dz += tlz - lo.z;
                        r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        float z;
                        if (base + s == t) {
                            z = 0.f;   // self pair: exactly zero (see estr.hip)
                        } else if constexpr (K == VPM_KERNEL_GAUSSIANERF) {
                            z = g.w * ex2(c.w * r2);
                        } else if constexpr (K == VPM_KERNEL_GAUSSIAN) {
                            const float r = r2 > 0.f ? r2 * rsq(r2) : 0.f;
                            z = g.w * ex2(c.w * r2 * r);
                        } else {
                            const float w = rsq(fmaf(c.w, r2, 1.f));
                            const float w2 = w * w;
                            z = g.w * (w2 * w2) * (w2 * w);
                        }
                        acc[0] = fmaf(z, g.x, acc[0]);
                        acc[1] = fmaf(z, g.y, acc[1]);
                        acc[2] = fmaf(z, g.z, acc[2]);
                        acc[3] = fmaf(z, c.x, acc[3]);
                        acc[4] = fmaf(z, c.y, acc[4]);
                        acc[5] = fmaf(z, c.z, acc[5]);
                    }
                }
#pragma unroll
                for (int c = 0; c < 6; ++c) {
                    const float x = acc[c];
                    const float tt = sum[c] + x;
                    const float bp = tt - sum[c];
                    err[c] += (sum[c] - (tt - bp)) + (x - bp);
                    sum[c] = tt;
                }
            }
        }
    }
    if (active) {
        double* rec = P + t * NF;
        double A[6];
#pragma unroll
        for (int c = 0; c < 6; ++c) A[c] = (double)sum[c] + (double)err[c];
        double Jp[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) Jp[c] = rec[RJ + c] - Jr[c];
        for (int q = 0; q < 3; ++q) {
            const double S = transposed ? Jp[3 * q] * A[0] + Jp[3 * q + 1] * A[1] + Jp[3 * q + 2] * A[2] - A[3 + q]
                                        : Jp[q] * A[0] + Jp[3 + q] * A[1] + Jp[6 + q] * A[2] - A[3 + q];
            rec[RSFS + q] += S;
        }
    }
}

// E_str over short target intervals (FMM leaves), as ranges_wave_f32c_kernel:
// one wave per item of <= 64 / R targets, each target replicated over R lane
// groups that split the item's source stream (its ranges in list order,
// staged 64 at a time through the wave's own LDS slot).  A step of R sources
// is evaluated when some lane's pair is inside its source's cut; the lanes
// whose own pair is outside add a zeta below FP32 resolution, as the block
// kernel does for such lanes.  The per-lane (A, B) sums are combined over the
// R groups in a fixed butterfly before J'^T A - B is formed (it is linear in
// A and B).  Each lane adds its FP32 sums into FP64 accumulators after
// kEstrFold of its own sources (a conversion and an FP64 add per component).
// Chains of 4 keep the error against the summation scale sum |zeta (J_i -
// J_j)^T Gamma_j| at the block kernel's on 5-particle leaves and below it on
// longer ones; chains of 16 doubled it on 5-particle leaves for 4-8 % less time.
constexpr int kEstrFold = 4;

template <int K, int R>
__global__ __launch_bounds__(kBlock) void estr_ranges_wave_f32c_kernel(const RangeItem* __restrict__ items,
                                                                       int64_t nitems,
                                                                       const int2* __restrict__ lsrc,
                                                                       const ListSrc S,
                                                                       double* __restrict__ P, int transposed)
{
    static_assert(R == 1 || R == 2 || R == 4 || R == 8, "replication");
    constexpr int TW = 64 / R;
    __shared__ float4 lp_all[kBlock / 64][64];   // s'_hi, cut
    __shared__ float4 lg_all[kBlock / 64][64];   // Gamma, b
    __shared__ float4 lq_all[kBlock / 64][64];   // J'^T Gamma (or J' Gamma), a
    __shared__ float4 ll_all[kBlock / 64][64];   // s'_lo
    const int lane = threadIdx.x & 63;
    // wave-uniform (kept in scalar registers, so the item and its source
    // ranges are read with scalar loads and the stream bookkeeping is SALU)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t w = xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + wave;
    if (w >= nitems) return;   // no block barrier below
    float4* lp = lp_all[wave];
    float4* lg = lg_all[wave];
    float4* lq = lq_all[wave];
    float4* ll = ll_all[wave];
    const RangeItem it = items[w];
    const double* r0 = P + it.t_begin * NF;
    double Jr[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) Jr[c] = r0[RJ + c];
    const int g = lane / TW;
    const int64_t t = it.t_begin + (lane % TW);
    const bool active = t < it.t_end;
    const int64_t tt = active ? t : it.t_begin;
    const float4 th = S.hi[tt], tl = S.lo[tt];
    const float tx = th.x, ty = th.y, tz = th.z, tlx = tl.x, tly = tl.y, tlz = tl.z;
    float acc[6];
    double A[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) { acc[c] = 0.f; A[c] = 0.0; }
    int since = 0;
    int64_t l = it.l_begin, off = 0;
    while (l < it.l_end) {
        int filled = 0;
        int64_t j = 0;
        if constexpr (R == 8) {
            filled = stream_fill(lsrc, l, it.l_end, off, lane, j, (int*)lp);
            if (filled == 0) continue;          // only empty ranges were read
        } else {
            while (filled < 64 && l < it.l_end) {
                const int2 rs = lsrc[l];
                const int64_t len = (int64_t)rs.y - rs.x;
                const int take = (int)min<int64_t>(len - off, (int64_t)(64 - filled));
                if (lane >= filled && lane < filled + take) j = rs.x + off + (lane - filled);
                filled += take;
                off += take;
                if (off == len) { ++l; off = 0; }
            }
            if (filled == 0) break;
        }
        __builtin_amdgcn_wave_barrier();
        // slots past `filled`: a zero-strength source far away with a zero
        // cut (it never starts a step, and adds exact zeros to one); those
        // lanes read source j = 0 (a valid index) and select the padding
        float4 p, lo, q, gq;
        {
            p = S.hi[j];                                // s_hi, cut
            lo = S.lo[j];                               // s_lo, a
            gq = S.gb[j];
            const float a = lo.w;
            lo.w = __builtin_bit_cast(float, (int)j);   // the source, for the self-pair test
            // (J_j - J_ref)^T Gamma_j (or (J_j - J_ref) Gamma_j)
            const double G0 = gq.x, G1 = gq.y, G2 = gq.z;
            double q3[3];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                const double ref = transposed ? Jr[3 * c] * G0 + Jr[3 * c + 1] * G1 + Jr[3 * c + 2] * G2
                                              : Jr[c] * G0 + Jr[3 + c] * G1 + Jr[6 + c] * G2;
                q3[c] = S.jg[c * S.ld + j] - ref;
            }
            q = make_float4((float)q3[0], (float)q3[1], (float)q3[2], a);
        }
        const bool real = lane < filled;
        const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
        lp[lane] = sel4(real, p, make_float4(1e15f, 1e15f, 1e15f, 0.f));
        ll[lane] = sel4(real, lo, make_float4(0.f, 0.f, 0.f, __builtin_bit_cast(float, -1)));
        lq[lane] = sel4(real, q, z4);
        lg[lane] = sel4(real, gq, z4);
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if constexpr (R == 1) {
            // whole-wave leaves: two steps per iteration, so the two pairs'
            // LDS reads and exp2 overlap (10 % faster on 64-particle leaves;
            // 4 % slower with lane sharing, which keeps one step)
            // steps k and k + 1: past the stream the slot is padding, and
            // k + 1 < 64 / R whenever k R < filled <= 64 with k even
            for (int k = 0; k * R < filled; k += 2) {
                const int s0 = k * R + g, s1 = s0 + R;
                const float4 a0 = lp[s0], a1 = lp[s1];
                float dx0 = tx - a0.x, dy0 = ty - a0.y, dz0 = tz - a0.z;
                float dx1 = tx - a1.x, dy1 = ty - a1.y, dz1 = tz - a1.z;
                float r20 = fmaf(dz0, dz0, fmaf(dy0, dy0, dx0 * dx0));
                float r21 = fmaf(dz1, dz1, fmaf(dy1, dy1, dx1 * dx1));
                if (__any(r20 < a0.w || r21 < a1.w)) {
                    const float4 c0 = lq[s0], c1 = lq[s1];
                    const float4 lo0 = ll[s0], lo1 = ll[s1];
                    const float4 g0 = lg[s0], g1 = lg[s1];
                    dx0 += tlx - lo0.x; dy0 += tly - lo0.y; dz0 += tlz - lo0.z;
                    dx1 += tlx - lo1.x; dy1 += tly - lo1.y; dz1 += tlz - lo1.z;
                    r20 = fmaf(dz0, dz0, fmaf(dy0, dy0, dx0 * dx0));
                    r21 = fmaf(dz1, dz1, fmaf(dy1, dy1, dx1 * dx1));
                    float z0, z1;
                    if constexpr (K == VPM_KERNEL_GAUSSIANERF) {
                        z0 = g0.w * ex2(c0.w * r20);
                        z1 = g1.w * ex2(c1.w * r21);
                    } else if constexpr (K == VPM_KERNEL_GAUSSIAN) {
                        const float q0 = r20 > 0.f ? r20 * rsq(r20) : 0.f;
                        const float q1 = r21 > 0.f ? r21 * rsq(r21) : 0.f;
                        z0 = g0.w * ex2(c0.w * r20 * q0);
                        z1 = g1.w * ex2(c1.w * r21 * q1);
                    } else {
                        const float w0 = rsq(fmaf(c0.w, r20, 1.f)), w1 = rsq(fmaf(c1.w, r21, 1.f));
                        const float v0 = w0 * w0, v1 = w1 * w1;
                        z0 = g0.w * (v0 * v0) * (v0 * w0);
                        z1 = g1.w * (v1 * v1) * (v1 * w1);
                    }
                    if (__builtin_bit_cast(int, lo0.w) == (int)t) z0 = 0.f;   // self pair
                    if (__builtin_bit_cast(int, lo1.w) == (int)t) z1 = 0.f;
                    acc[0] = fmaf(z0, g0.x, acc[0]);
                    acc[1] = fmaf(z0, g0.y, acc[1]);
                    acc[2] = fmaf(z0, g0.z, acc[2]);
                    acc[3] = fmaf(z0, c0.x, acc[3]);
                    acc[4] = fmaf(z0, c0.y, acc[4]);
                    acc[5] = fmaf(z0, c0.z, acc[5]);
                    acc[0] = fmaf(z1, g1.x, acc[0]);
                    acc[1] = fmaf(z1, g1.y, acc[1]);
                    acc[2] = fmaf(z1, g1.z, acc[2]);
                    acc[3] = fmaf(z1, c1.x, acc[3]);
                    acc[4] = fmaf(z1, c1.y, acc[4]);
                    acc[5] = fmaf(z1, c1.z, acc[5]);
                }
                since = __builtin_amdgcn_readfirstlane(since + 2);   // wave-uniform
                if (since >= kEstrFold) {
    #pragma unroll
                    for (int c = 0; c < 6; ++c) { A[c] += (double)acc[c]; acc[c] = 0.f; }
                    since = 0;
                }
            }
        } else {
            // groups of kEstrFold steps, each chain folded after its group:
            // 64 / R is a multiple of kEstrFold, so with k0 R < filled <= 64
            // every step k < k0 + kEstrFold has sidx < 64 (past `filled` the
            // slot is padding, whose zero cut skips the step)
            static_assert((64 / R) % kEstrFold == 0, "fold groups must tile the slot");
            for (int k0 = 0; k0 * R < filled; k0 += kEstrFold) {
#pragma unroll
                for (int kk = 0; kk < kEstrFold; ++kk) {
                    const int sidx = (k0 + kk) * R + g;
                    const float4 a = lp[sidx];
                    float dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
                    float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    if (__any(r2 < a.w)) {
                        const float4 c = lq[sidx];
                        const float4 lo = ll[sidx];
                        const float4 gg = lg[sidx];
                        dx += tlx - lo.x;
                        dy += tly - lo.y;
                        dz += tlz - lo.z;
                        r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        float z;
                        if (__builtin_bit_cast(int, lo.w) == (int)t) {
                            z = 0.f;   // self pair: exactly zero (see estr.hip)
                        } else if constexpr (K == VPM_KERNEL_GAUSSIANERF) {
                            z = gg.w * ex2(c.w * r2);
                        } else if constexpr (K == VPM_KERNEL_GAUSSIAN) {
                            const float r = r2 > 0.f ? r2 * rsq(r2) : 0.f;
                            z = gg.w * ex2(c.w * r2 * r);
                        } else {
                            const float ww = rsq(fmaf(c.w, r2, 1.f));
                            const float w2 = ww * ww;
                            z = gg.w * (w2 * w2) * (w2 * ww);
                        }
                        acc[0] = fmaf(z, gg.x, acc[0]);
                        acc[1] = fmaf(z, gg.y, acc[1]);
                        acc[2] = fmaf(z, gg.z, acc[2]);
                        acc[3] = fmaf(z, c.x, acc[3]);
                        acc[4] = fmaf(z, c.y, acc[4]);
                        acc[5] = fmaf(z, c.z, acc[5]);
                    }
                }
#pragma unroll
                for (int c = 0; c < 6; ++c) { A[c] += (double)acc[c]; acc[c] = 0.f; }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) A[c] += (double)acc[c];
    if constexpr (R > 1) {
#pragma unroll
        for (int m = TW; m < 64; m <<= 1) {
#pragma unroll
            for (int c = 0; c < 6; ++c) A[c] += __shfl_xor(A[c], m);
        }
    }
    if (g == 0 && active) {
        double* rec = P + t * NF;
        double Jp[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) Jp[c] = rec[RJ + c] - Jr[c];
        for (int q = 0; q < 3; ++q) {
            const double S = transposed ? Jp[3 * q] * A[0] + Jp[3 * q + 1] * A[1] + Jp[3 * q + 2] * A[2] - A[3 + q]
                                        : Jp[q] * A[0] + Jp[3 + q] * A[1] + Jp[6 + q] * A[2] - A[3 + q];
            rec[RSFS + q] += S;
        }
    }
}

#pragma clang fp contract(off)
template <int K>
__device__ __forceinline__ double zeta_ref_r(double r)
{
    if constexpr (K == VPM_KERNEL_SINGULAR) return r == 0 ? 1.0 : 0.0;
    else if constexpr (K == VPM_KERNEL_GAUSSIAN) return kConst3 * exp(-r * r * r);
    else if constexpr (K == VPM_KERNEL_GAUSSIANERF) return kConst1 * exp(-r * r / 2);
    else {
        double temp = r * r + 1;
        const double temp2 = temp * temp * temp;
        temp = temp2 * temp2 * temp;
        return kConst4 * 7.5 / sqrt(temp);
    }
}

// FP64: subfilterscale_models.jl:16-41 per pair, list order, accumulating
// onto SFS directly as the reference does.
template <int K, int NT>
__global__ __launch_bounds__(NT) void estr_ranges_f64_kernel(const RangeItem* __restrict__ items,
                                                                 const int2* __restrict__ lsrc,
                                                                 double* __restrict__ P, int transposed)
{
    __shared__ double lx[NT][4];
    __shared__ double lg[NT][3];
    __shared__ double lj[NT][9];
    const RangeItem it = items[xcd_block(blockIdx.x, gridDim.x)];
    const int tid = threadIdx.x;
    const int64_t t = it.t_begin + tid;
    const bool active = t < it.t_end;
    const double* tr = P + (active ? t : it.t_begin) * NF;
    const double tx = tr[RX], ty = tr[RX + 1], tz = tr[RX + 2];
    double JT[9];
    for (int c = 0; c < 9; ++c) JT[c] = tr[RJ + c];
    double s1 = tr[RSFS], s2 = tr[RSFS + 1], s3 = tr[RSFS + 2];
    for (int64_t l = it.l_begin; l < it.l_end; ++l) {
        const int2 rs = lsrc[l];
        const int64_t sb = rs.x, se = rs.y;
        for (int64_t base = sb; base < se; base += NT) {
            const int n = (int)((se - base) < NT ? (se - base) : NT);
            __syncthreads();
            if (tid < n) {
                const double* sr = P + (base + tid) * NF;
                lx[tid][0] = sr[RX]; lx[tid][1] = sr[RX + 1]; lx[tid][2] = sr[RX + 2]; lx[tid][3] = sr[RSIG];
                lg[tid][0] = sr[RG]; lg[tid][1] = sr[RG + 1]; lg[tid][2] = sr[RG + 2];
                for (int c = 0; c < 9; ++c) lj[tid][c] = sr[RJ + c];
            }
            __syncthreads();
            for (int s = 0; s < n; ++s) {
                const double dx = lx[s][0] - tx, dy = lx[s][1] - ty, dz = lx[s][2] - tz;
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                const double* JS = lj[s];
                const double* GS = lg[s];
                double S1, S2, S3;
                if (transposed) {
                    S1 = (JT[0] - JS[0]) * GS[0] + (JT[1] - JS[1]) * GS[1] + (JT[2] - JS[2]) * GS[2];
                    S2 = (JT[3] - JS[3]) * GS[0] + (JT[4] - JS[4]) * GS[1] + (JT[5] - JS[5]) * GS[2];
                    S3 = (JT[6] - JS[6]) * GS[0] + (JT[7] - JS[7]) * GS[1] + (JT[8] - JS[8]) * GS[2];
                } else {
                    S1 = (JT[0] - JS[0]) * GS[0] + (JT[3] - JS[3]) * GS[1] + (JT[6] - JS[6]) * GS[2];
                    S2 = (JT[1] - JS[1]) * GS[0] + (JT[4] - JS[4]) * GS[1] + (JT[7] - JS[7]) * GS[2];
                    S3 = (JT[2] - JS[2]) * GS[0] + (JT[5] - JS[5]) * GS[1] + (JT[8] - JS[8]) * GS[2];
                }
                const double sigma_inv = 1.0 / lx[s][3];
                const double zeta_sgm = zeta_ref_r<K>(r * sigma_inv) * sigma_inv * sigma_inv * sigma_inv;
                s1 += zeta_sgm * S1;
                s2 += zeta_sgm * S2;
                s3 += zeta_sgm * S3;
            }
        }
    }
    if (active) {
        double* rec = P + t * NF;
        rec[RSFS] = s1;
        rec[RSFS + 1] = s2;
        rec[RSFS + 2] = s3;
    }
}
// E_str of estr_ranges_f64_kernel with G items per wave (see
// ranges_f64_group_kernel): the same pair expression in the same order.
template <int K, int G>
__global__ __launch_bounds__(kBlock) void estr_ranges_f64_group_kernel(const RangeItem* __restrict__ items,
                                                                       int64_t nitems, const int2* __restrict__ lsrc,
                                                                       double* __restrict__ P, int transposed)
{
    constexpr int NG = 64 / G;
    __shared__ double lx_all[kBlock / 64][64][4];
    __shared__ double lg_all[kBlock / 64][64][3];
    __shared__ double lj_all[kBlock / 64][64][9];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane / NG, k = lane % NG;
    double (*lx)[4] = lx_all[wave] + g * NG;
    double (*lg)[3] = lg_all[wave] + g * NG;
    double (*lj)[9] = lj_all[wave] + g * NG;
    const int64_t item = (xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + wave) * G + g;
    const RangeItem it = item < nitems ? items[item] : RangeItem{0, 0, 0, 0};
    const int64_t t = it.t_begin + k;
    const bool active = t < it.t_end;
    const double* tr = P + (active ? t : 0) * NF;
    const double tx = tr[RX], ty = tr[RX + 1], tz = tr[RX + 2];
    double JT[9];
    for (int c = 0; c < 9; ++c) JT[c] = tr[RJ + c];
    double s1 = tr[RSFS], s2 = tr[RSFS + 1], s3 = tr[RSFS + 2];
    int64_t l = it.l_begin, base = 0, se = 0;
    if (l < it.l_end) { const int2 rs = lsrc[l]; base = rs.x; se = rs.y; }
    for (;;) {
        while (l < it.l_end && base >= se) {
            if (++l < it.l_end) { const int2 rs = lsrc[l]; base = rs.x; se = rs.y; }
        }
        const bool work = l < it.l_end;
        if (!__any(work)) break;
        const int n = work ? (int)((se - base) < NG ? (se - base) : NG) : 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (k < n) {
            const double* sr = P + (base + k) * NF;
            lx[k][0] = sr[RX]; lx[k][1] = sr[RX + 1]; lx[k][2] = sr[RX + 2]; lx[k][3] = sr[RSIG];
            lg[k][0] = sr[RG]; lg[k][1] = sr[RG + 1]; lg[k][2] = sr[RG + 2];
            for (int c = 0; c < 9; ++c) lj[k][c] = sr[RJ + c];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int s = 0; s < n; ++s) {
            const double dx = lx[s][0] - tx, dy = lx[s][1] - ty, dz = lx[s][2] - tz;
            const double r = sqrt(dx * dx + dy * dy + dz * dz);
            const double* JS = lj[s];
            const double* GS = lg[s];
            double S1, S2, S3;
            if (transposed) {
                S1 = (JT[0] - JS[0]) * GS[0] + (JT[1] - JS[1]) * GS[1] + (JT[2] - JS[2]) * GS[2];
                S2 = (JT[3] - JS[3]) * GS[0] + (JT[4] - JS[4]) * GS[1] + (JT[5] - JS[5]) * GS[2];
                S3 = (JT[6] - JS[6]) * GS[0] + (JT[7] - JS[7]) * GS[1] + (JT[8] - JS[8]) * GS[2];
            } else {
                S1 = (JT[0] - JS[0]) * GS[0] + (JT[3] - JS[3]) * GS[1] + (JT[6] - JS[6]) * GS[2];
                S2 = (JT[1] - JS[1]) * GS[0] + (JT[4] - JS[4]) * GS[1] + (JT[7] - JS[7]) * GS[2];
                S3 = (JT[2] - JS[2]) * GS[0] + (JT[5] - JS[5]) * GS[1] + (JT[8] - JS[8]) * GS[2];
            }
            const double sigma_inv = 1.0 / lx[s][3];
            const double zeta_sgm = zeta_ref_r<K>(r * sigma_inv) * sigma_inv * sigma_inv * sigma_inv;
            s1 += zeta_sgm * S1;
            s2 += zeta_sgm * S2;
            s3 += zeta_sgm * S3;
        }
        base += n;
    }
    if (active) {
        double* rec = P + t * NF;
        rec[RSFS] = s1;
        rec[RSFS + 1] = s2;
        rec[RSFS + 2] = s3;
    }
}
#pragma clang fp contract(on)

// E_str of a planned list: F32C wave kernels for wave plans, block kernels
// otherwise; FP64 (and the singular kernel, whose zeta is a delta) takes the
// FP64 block kernel, with one wave per item on wave plans.
int launch_estr_plan(vpm_field f, int kernel, int precision, int transposed, const DevPlan& dp)
{
    if (dp.nitems == 0) return VPM_OK;
    const bool f64 = precision == VPM_PREC_F64 || kernel == VPM_KERNEL_SINGULAR;
    const RangeItem* d_items = dp.items;
    const int2* d_src = dp.src;
    const dim3 grid((unsigned)dp.nitems);
    timing_begin(f);
    ListSrc S;
    if (!f64) {
        if (int rc = pack_list_src(f, kernel, transposed ? 1 : 2, S)) return rc;
    }
    if (!f64 && dp.wave) {
        const int64_t ni = dp.nitems;
        const dim3 wgrid((unsigned)((ni + kBlock / 64 - 1) / (kBlock / 64)));
        switch (kernel) {
#define VPM_EWR(KID, R) hipLaunchKernelGGL((estr_ranges_wave_f32c_kernel<KID, R>), wgrid, dim3(kBlock), 0, f->stream, d_items, ni, d_src, S, f->P, transposed)
#define VPM_EW(KID)                                                                   \
    case KID:                                                                         \
        switch (dp.R) {                                                               \
        case 8: VPM_EWR(KID, 8); break;                                               \
        case 4: VPM_EWR(KID, 4); break;                                               \
        case 2: VPM_EWR(KID, 2); break;                                               \
        default: VPM_EWR(KID, 1); break;                                              \
        }                                                                             \
        break;
            VPM_EW(VPM_KERNEL_GAUSSIAN) VPM_EW(VPM_KERNEL_GAUSSIANERF) VPM_EW(VPM_KERNEL_WINCKELMANS)
#undef VPM_EW
#undef VPM_EWR
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    } else if (f64) {
        switch (kernel) {
#define VPM_EG64(KID, G)                                                                                        \
    hipLaunchKernelGGL((estr_ranges_f64_group_kernel<KID, G>),                                                  \
                       dim3((unsigned)((dp.nitems + G * (kBlock / 64) - 1) / (G * (kBlock / 64)))),             \
                       dim3(kBlock), 0, f->stream, d_items, dp.nitems, d_src, f->P, transposed)
#define VPM_E64(KID)                                                                                            \
    case KID:                                                                                                   \
        if (dp.wave && dp.R > 1) {                                                                              \
            if (dp.R == 8) VPM_EG64(KID, 8);                                                                    \
            else if (dp.R == 4) VPM_EG64(KID, 4);                                                               \
            else VPM_EG64(KID, 2);                                                                              \
        } else if (dp.wave) {                                                                                   \
            hipLaunchKernelGGL((estr_ranges_f64_kernel<KID, 64>), grid, dim3(64), 0, f->stream, d_items, d_src, \
                               f->P, transposed);                                                               \
        } else {                                                                                                \
            hipLaunchKernelGGL((estr_ranges_f64_kernel<KID, kBlock>), grid, dim3(kBlock), 0, f->stream, d_items, \
                               d_src, f->P, transposed);                                                        \
        }                                                                                                       \
        break;
            VPM_E64(VPM_KERNEL_SINGULAR) VPM_E64(VPM_KERNEL_GAUSSIAN) VPM_E64(VPM_KERNEL_GAUSSIANERF) VPM_E64(VPM_KERNEL_WINCKELMANS)
#undef VPM_E64
#undef VPM_EG64
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    } else {
        switch (kernel) {
#define VPM_E32(KID) case KID: hipLaunchKernelGGL(estr_ranges_f32c_kernel<KID>, grid, dim3(kBlock), 0, f->stream, d_items, d_src, S, f->P, transposed); break;
            VPM_E32(VPM_KERNEL_GAUSSIAN) VPM_E32(VPM_KERNEL_GAUSSIANERF) VPM_E32(VPM_KERNEL_WINCKELMANS)
#undef VPM_E32
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "estr_ranges kernel");
    timing_end(f, 2);
    return VPM_OK;
}

int launch_estr_ranges(vpm_field f, int kernel, int precision, int transposed, const int64_t* tb,
                       const int64_t* te, const int64_t* sb, const int64_t* se, int64_t npairs)
{
    const bool f64 = precision == VPM_PREC_F64 || kernel == VPM_KERNEL_SINGULAR;
    DevPlan dp;
    if (int rc = make_dev_plan(f, tb, te, sb, se, npairs, true, 8, f->rng, dp)) return rc;
    return launch_estr_plan(f, kernel, precision, transposed, dp);
}

// ------------------------------------------------------------- zeta ranges
// zeta_fmm's pair loop over a direct list (src/FLOWVPM_viscous.jl:523-558):
// for every (target range, source range), J[1:3] of each target += Gamma_j
// zeta(r/sigma_j)/sigma_j^3 over the sources.  No reset, no static skip.
//
// F32C: the packed sources of pack_list_src (hi/lo positions relative to
// particle 0, (a, b) per source as zeta_params, the cut as zeta_cut),
// sources beyond the cut skipped per wave, sums
// folded into TwoSum every 32 sources and added onto the FP64 record.  BS
// is the block size: 64 (one wave per FMM leaf) for short items, else 256.
// Leaves of <= 32 targets on average take zeta_ranges_wave_f32c_kernel.
template <int K, int BS>
__global__ __launch_bounds__(BS) void zeta_ranges_f32c_kernel(const RangeItem* __restrict__ items,
                                                              const int2* __restrict__ lsrc,
                                                              const ListSrc S,
                                                              double* __restrict__ P)
{
    __shared__ float4 lp[BS];   // s'_hi, cut
    __shared__ float4 lg[BS];   // Gamma, b
    __shared__ float4 ll[BS];   // s'_lo, a
    const RangeItem it = items[xcd_block(blockIdx.x, gridDim.x)];
    const int tid = threadIdx.x;
    const int64_t t = it.t_begin + tid;
    const bool active = t < it.t_end;
    const int64_t tt = active ? t : it.t_begin;
    const float4 th = S.hi[tt], tl = S.lo[tt];
    const float tx = th.x, ty = th.y, tz = th.z, tlx = tl.x, tly = tl.y, tlz = tl.z;
    float sum[3] = {0.f, 0.f, 0.f}, err[3] = {0.f, 0.f, 0.f};
    for (int64_t l = it.l_begin; l < it.l_end; ++l) {
        const int2 rs = lsrc[l];
        const int64_t sb = rs.x, se = rs.y;
        for (int64_t base = sb; base < se; base += BS) {
            const int n = (int)((se - base) < BS ? (se - base) : BS);
            __syncthreads();
            if (tid < n) {
                const int64_t j = base + tid;
                lp[tid] = S.hi[j];   // s_hi, cut
                ll[tid] = S.lo[j];   // s_lo, a
                lg[tid] = S.gb[j];
            }
            __syncthreads();
            for (int s0 = 0; s0 < n; s0 += 32) {
                const int s1 = min(s0 + 32, n);
                float acc[3] = {0.f, 0.f, 0.f};
                for (int s = s0; s < s1; ++s) {
                    const float4 a = lp[s];
                    float dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
                    float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    if (__any(r2 < a.w)) {
                        const float4 lo = ll[s];
                        const float4 g = lg[s];
                        dx += tlx - lo.x;
                        dy += tly - lo.y;
                        dz += tlz - lo.z;
                        r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                        const float z = zeta_ab<K>(r2, lo.w, g.w);
                        acc[0] = fmaf(g.x, z, acc[0]);
                        acc[1] = fmaf(g.y, z, acc[1]);
                        acc[2] = fmaf(g.z, z, acc[2]);
                    }
                }
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    const float x = acc[c];
                    const float tt = sum[c] + x;
                    const float bp = tt - sum[c];
                    err[c] += (sum[c] - (tt - bp)) + (x - bp);
                    sum[c] = tt;
                }
            }
        }
    }
    if (active) {
        double* rec = P + t * NF;
#pragma unroll
        for (int c = 0; c < 3; ++c) rec[RJ + c] += (double)sum[c] + (double)err[c];
    }
}

// zeta over FMM leaves shared R = 2, 4 or 8 ways per wave, as
// estr_ranges_wave_f32c_kernel: one wave per item of <= 64 / R targets, each
// target replicated over R lane groups that split the item's source stream
// (its ranges in list order, staged 64 at a time through the wave's own LDS
// slot).  FP32 chains of kZetaFold lane sources (8 when R = 8, whose slot
// holds 8 steps) are added into FP64 sums, which the R groups combine in a
// fixed butterfly.  The zeta sum has no
// cancelling J'^T A - B form, so its chains are longer than E_str's.  Leaves
// of a whole wave (R = 1) take zeta_ranges_f32c_kernel<K, 64>, which was 22 %
// faster there (6.35 against 8.13 ms on the 2^20 jet's 64-particle leaves).
constexpr int kZetaFold = 16;
template <int K, int R>
__global__ __launch_bounds__(kBlock) void zeta_ranges_wave_f32c_kernel(const RangeItem* __restrict__ items,
                                                                       int64_t nitems,
                                                                       const int2* __restrict__ lsrc,
                                                                       const ListSrc S,
                                                                       double* __restrict__ P)
{
    static_assert(R == 2 || R == 4 || R == 8, "replication");
    constexpr int TW = 64 / R;
    constexpr int ZG = kZetaFold < 64 / R ? kZetaFold : 64 / R;   // steps per FP32 chain
    static_assert((64 / R) % ZG == 0, "chain groups must tile the slot");
    __shared__ float4 lp_all[kBlock / 64][64];   // s'_hi, cut
    __shared__ float4 lg_all[kBlock / 64][64];   // Gamma, b
    __shared__ float4 ll_all[kBlock / 64][64];   // s'_lo, a
    const int lane = threadIdx.x & 63;
    // wave-uniform (kept in scalar registers, so the item and its source
    // ranges are read with scalar loads and the stream bookkeeping is SALU)
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t w = xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + wave;
    if (w >= nitems) return;   // no block barrier below
    float4* lp = lp_all[wave];
    float4* lg = lg_all[wave];
    float4* ll = ll_all[wave];
    const RangeItem it = items[w];
    const int g = lane / TW;
    const int64_t t = it.t_begin + (lane % TW);
    const bool active = t < it.t_end;
    const int64_t tt = active ? t : it.t_begin;
    const float4 th = S.hi[tt], tl = S.lo[tt];
    const float tx = th.x, ty = th.y, tz = th.z, tlx = tl.x, tly = tl.y, tlz = tl.z;
    float acc[3] = {0.f, 0.f, 0.f};
    double A[3] = {0.0, 0.0, 0.0};
    int64_t l = it.l_begin, off = 0;
    while (l < it.l_end) {
        int filled = 0;
        int64_t j = 0;
        if constexpr (R == 8) {
            filled = stream_fill(lsrc, l, it.l_end, off, lane, j, (int*)lp);
            if (filled == 0) continue;          // only empty ranges were read
        } else {
            while (filled < 64 && l < it.l_end) {
                const int2 rs = lsrc[l];
                const int64_t len = (int64_t)rs.y - rs.x;
                const int take = (int)min<int64_t>(len - off, (int64_t)(64 - filled));
                if (lane >= filled && lane < filled + take) j = rs.x + off + (lane - filled);
                filled += take;
                off += take;
                if (off == len) { ++l; off = 0; }
            }
            if (filled == 0) break;
        }
        __builtin_amdgcn_wave_barrier();
        {
            // slots past `filled`: a zero-strength source far away with a
            // zero cut (it never starts a step, and adds exact zeros to one);
            // those lanes read source j = 0 (a valid index) and select the
            // padding
            const float4 p = S.hi[j];    // s_hi, cut
            const float4 lo = S.lo[j];   // s_lo, a
            const float4 gq = S.gb[j];
            const bool real = lane < filled;
            const float4 z4 = make_float4(0.f, 0.f, 0.f, 0.f);
            lp[lane] = sel4(real, p, make_float4(1e15f, 1e15f, 1e15f, 0.f));
            ll[lane] = sel4(real, lo, z4);
            lg[lane] = sel4(real, gq, z4);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // groups of ZG steps, each chain folded after its group: 64 / R is a
        // multiple of ZG, so with k0 R < filled <= 64 every step k < k0 + ZG
        // has sidx < 64 (past `filled` the slot is padding, whose zero cut
        // skips the step)
        for (int k0 = 0; k0 * R < filled; k0 += ZG) {
#pragma unroll
            for (int kk = 0; kk < ZG; ++kk) {
                const int sidx = (k0 + kk) * R + g;
                const float4 a = lp[sidx];
                float dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
                float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                if (__any(r2 < a.w)) {
                    const float4 lo = ll[sidx];
                    const float4 gg = lg[sidx];
                    dx += tlx - lo.x;
                    dy += tly - lo.y;
                    dz += tlz - lo.z;
                    r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    const float z = zeta_ab<K>(r2, lo.w, gg.w);
                    acc[0] = fmaf(gg.x, z, acc[0]);
                    acc[1] = fmaf(gg.y, z, acc[1]);
                    acc[2] = fmaf(gg.z, z, acc[2]);
                }
            }
#pragma unroll
            for (int c = 0; c < 3; ++c) { A[c] += (double)acc[c]; acc[c] = 0.f; }
        }
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) A[c] += (double)acc[c];
    if constexpr (R > 1) {
#pragma unroll
        for (int m = TW; m < 64; m <<= 1) {
#pragma unroll
            for (int c = 0; c < 3; ++c) A[c] += __shfl_xor(A[c], m);
        }
    }
    if (g == 0 && active) {
        double* rec = P + t * NF;
#pragma unroll
        for (int c = 0; c < 3; ++c) rec[RJ + c] += A[c];
    }
}

#pragma clang fp contract(off)
// FP64: viscous.jl:545-554 per pair, list order, accumulating onto J[1:3]
// as the reference does.
template <int K, int NT>
__global__ __launch_bounds__(NT) void zeta_ranges_f64_kernel(const RangeItem* __restrict__ items,
                                                                 const int2* __restrict__ lsrc,
                                                                 double* __restrict__ P)
{
    __shared__ double4 ls[NT];   // x, y, z, sigma
    __shared__ double4 lg[NT];   // Gamma
    const RangeItem it = items[xcd_block(blockIdx.x, gridDim.x)];
    const int tid = threadIdx.x;
    const int64_t t = it.t_begin + tid;
    const bool active = t < it.t_end;
    const double* tr = P + (active ? t : it.t_begin) * NF;
    const double tx = tr[RX], ty = tr[RX + 1], tz = tr[RX + 2];
    double acc[3] = {tr[RJ], tr[RJ + 1], tr[RJ + 2]};
    for (int64_t l = it.l_begin; l < it.l_end; ++l) {
        const int2 rs = lsrc[l];
        const int64_t sb = rs.x, se = rs.y;
        for (int64_t base = sb; base < se; base += NT) {
            const int n = (int)((se - base) < NT ? (se - base) : NT);
            __syncthreads();
            if (tid < n) {
                const double* sr = P + (base + tid) * NF;
                ls[tid] = make_double4(sr[RX], sr[RX + 1], sr[RX + 2], sr[RSIG]);
                lg[tid] = make_double4(sr[RG], sr[RG + 1], sr[RG + 2], 0.0);
            }
            __syncthreads();
            for (int s = 0; s < n; ++s) {
                const double4 a = ls[s];
                const double dX1 = tx - a.x, dX2 = ty - a.y, dX3 = tz - a.z;
                const double r = sqrt(dX1 * dX1 + dX2 * dX2 + dX3 * dX3);
                const double zeta_sgm = 1 / (a.w * a.w * a.w) * zeta_ref_r<K>(r / a.w);
                acc[0] += lg[s].x * zeta_sgm;
                acc[1] += lg[s].y * zeta_sgm;
                acc[2] += lg[s].z * zeta_sgm;
            }
        }
    }
    if (active) {
        double* rec = P + t * NF;
        for (int c = 0; c < 3; ++c) rec[RJ + c] = acc[c];
    }
}
// zeta_ranges_f64_kernel with G items per wave (see ranges_f64_group_kernel).
template <int K, int G>
__global__ __launch_bounds__(kBlock) void zeta_ranges_f64_group_kernel(const RangeItem* __restrict__ items,
                                                                       int64_t nitems, const int2* __restrict__ lsrc,
                                                                       double* __restrict__ P)
{
    constexpr int NG = 64 / G;
    __shared__ double4 ls_all[kBlock / 64][64];
    __shared__ double4 lg_all[kBlock / 64][64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane / NG, k = lane % NG;
    double4* ls = ls_all[wave] + g * NG;
    double4* lg = lg_all[wave] + g * NG;
    const int64_t item = (xcd_block(blockIdx.x, gridDim.x) * (kBlock / 64) + wave) * G + g;
    const RangeItem it = item < nitems ? items[item] : RangeItem{0, 0, 0, 0};
    const int64_t t = it.t_begin + k;
    const bool active = t < it.t_end;
    const double* tr = P + (active ? t : 0) * NF;
    const double tx = tr[RX], ty = tr[RX + 1], tz = tr[RX + 2];
    double acc[3] = {tr[RJ], tr[RJ + 1], tr[RJ + 2]};
    int64_t l = it.l_begin, base = 0, se = 0;
    if (l < it.l_end) { const int2 rs = lsrc[l]; base = rs.x; se = rs.y; }
    for (;;) {
        while (l < it.l_end && base >= se) {
            if (++l < it.l_end) { const int2 rs = lsrc[l]; base = rs.x; se = rs.y; }
        }
        const bool work = l < it.l_end;
        if (!__any(work)) break;
        const int n = work ? (int)((se - base) < NG ? (se - base) : NG) : 0;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        if (k < n) {
            const double* sr = P + (base + k) * NF;
            ls[k] = make_double4(sr[RX], sr[RX + 1], sr[RX + 2], sr[RSIG]);
            lg[k] = make_double4(sr[RG], sr[RG + 1], sr[RG + 2], 0.0);
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        for (int s = 0; s < n; ++s) {
            const double4 a = ls[s];
            const double dX1 = tx - a.x, dX2 = ty - a.y, dX3 = tz - a.z;
            const double r = sqrt(dX1 * dX1 + dX2 * dX2 + dX3 * dX3);
            const double zeta_sgm = 1 / (a.w * a.w * a.w) * zeta_ref_r<K>(r / a.w);
            acc[0] += lg[s].x * zeta_sgm;
            acc[1] += lg[s].y * zeta_sgm;
            acc[2] += lg[s].z * zeta_sgm;
        }
        base += n;
    }
    if (active) {
        double* rec = P + t * NF;
        for (int c = 0; c < 3; ++c) rec[RJ + c] = acc[c];
    }
}
#pragma clang fp contract(on)

// zeta of a planned list: the F32C wave kernels (R > 1) or the 64-lane block
// kernel for wave plans, 256-lane blocks otherwise; FP64 (and the singular
// kernel) takes the FP64 block kernel, with one wave per item on wave plans.
int launch_zeta_plan(vpm_field f, int kernel, int precision, const DevPlan& dp)
{
    if (dp.nitems == 0) return VPM_OK;
    const bool f64 = precision == VPM_PREC_F64 || kernel == VPM_KERNEL_SINGULAR;
    const bool wave = dp.wave;
    const RangeItem* d_items = dp.items;
    const int2* d_src = dp.src;
    const dim3 grid((unsigned)dp.nitems);
    timing_begin(f);
    ListSrc S;
    if (!f64) {
        if (int rc = pack_list_src(f, kernel, 0, S)) return rc;
    }
    if (f64) {
        switch (kernel) {
#define VPM_ZG64(KID, G)                                                                                        \
    hipLaunchKernelGGL((zeta_ranges_f64_group_kernel<KID, G>),                                                  \
                       dim3((unsigned)((dp.nitems + G * (kBlock / 64) - 1) / (G * (kBlock / 64)))),             \
                       dim3(kBlock), 0, f->stream, d_items, dp.nitems, d_src, f->P)
#define VPM_Z64(KID)                                                                                            \
    case KID:                                                                                                   \
        if (wave && dp.R > 1) {                                                                                 \
            if (dp.R == 8) VPM_ZG64(KID, 8);                                                                    \
            else if (dp.R == 4) VPM_ZG64(KID, 4);                                                               \
            else VPM_ZG64(KID, 2);                                                                              \
        } else if (wave) {                                                                                      \
            hipLaunchKernelGGL((zeta_ranges_f64_kernel<KID, 64>), grid, dim3(64), 0, f->stream, d_items, d_src,  \
                               f->P);                                                                           \
        } else {                                                                                                \
            hipLaunchKernelGGL((zeta_ranges_f64_kernel<KID, kBlock>), grid, dim3(kBlock), 0, f->stream, d_items, \
                               d_src, f->P);                                                                    \
        }                                                                                                       \
        break;
            VPM_Z64(VPM_KERNEL_SINGULAR) VPM_Z64(VPM_KERNEL_GAUSSIAN) VPM_Z64(VPM_KERNEL_GAUSSIANERF) VPM_Z64(VPM_KERNEL_WINCKELMANS)
#undef VPM_Z64
#undef VPM_ZG64
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    } else if (wave) {
        const int64_t ni = dp.nitems;
        const dim3 wgrid((unsigned)((ni + kBlock / 64 - 1) / (kBlock / 64)));
        switch (kernel) {
#define VPM_ZWR(KID, R) hipLaunchKernelGGL((zeta_ranges_wave_f32c_kernel<KID, R>), wgrid, dim3(kBlock), 0, f->stream, d_items, ni, d_src, S, f->P)
#define VPM_ZW(KID)                                                                                                   \
    case KID:                                                                                                         \
        switch (dp.R) {                                                                                               \
        case 8: VPM_ZWR(KID, 8); break;                                                                               \
        case 4: VPM_ZWR(KID, 4); break;                                                                               \
        case 2: VPM_ZWR(KID, 2); break;                                                                               \
        default: hipLaunchKernelGGL((zeta_ranges_f32c_kernel<KID, 64>), grid, dim3(64), 0, f->stream, d_items, d_src, S, f->P); break; \
        }                                                                                                             \
        break;
            VPM_ZW(VPM_KERNEL_GAUSSIAN) VPM_ZW(VPM_KERNEL_GAUSSIANERF) VPM_ZW(VPM_KERNEL_WINCKELMANS)
#undef VPM_ZW
#undef VPM_ZWR
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    } else {
        switch (kernel) {
#define VPM_Z32(KID) case KID: hipLaunchKernelGGL((zeta_ranges_f32c_kernel<KID, kBlock>), grid, dim3(kBlock), 0, f->stream, d_items, d_src, S, f->P); break;
            VPM_Z32(VPM_KERNEL_GAUSSIAN) VPM_Z32(VPM_KERNEL_GAUSSIANERF) VPM_Z32(VPM_KERNEL_WINCKELMANS)
#undef VPM_Z32
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "zeta_ranges kernel");
    timing_end(f, 4);
    return VPM_OK;
}

int launch_zeta_ranges(vpm_field f, int kernel, int precision, const int64_t* tb, const int64_t* te,
                       const int64_t* sb, const int64_t* se, int64_t npairs)
{
    const bool f64 = precision == VPM_PREC_F64 || kernel == VPM_KERNEL_SINGULAR;
    DevPlan dp;
    if (int rc = make_dev_plan(f, tb, te, sb, se, npairs, true, 8, f->rng, dp)) return rc;
    return launch_zeta_plan(f, kernel, precision, dp);
}


}  // namespace vpm
