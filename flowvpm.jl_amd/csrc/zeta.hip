// zeta.hip -- the basis-function field the particles exert on each other
// (zeta_direct, src/FLOWVPM_viscous.jl:486-515): J[1:3] of every particle,
// static ones included, becomes
//     sum_j Gamma_j zeta(r_ij/sigma_j)/sigma_j^3
// over every particle j, the target itself included (there is no r == 0
// skip: the self term is Gamma_i zeta(0)/sigma_i^3).  The viscous scheme's
// RBF reset evaluates the vorticity field through it.
//
// F32C: the U/J tiling (lanes own targets, sources through LDS, Morton
// order, hi/lo positions relative to the block origin, per-tile TwoSum).
// zeta_sigma is (a, b) per source as in E_str: b exp2(a r^2) (Gaussian-erf),
// b exp2(a r^3) (Gaussian), b (1 + a r^2)^-3.5 (Winckelmans).  Sources whose
// distance to the wave's target box exceeds q_far sigma (Gaussian-erf 7,
// Gaussian 3: zeta below FP32 resolution of the self term) are skipped per
// wave with one ballot per 64 sources, and a block visits only the source
// tiles within reach of its targets (NearTiles, p2p_pair.h); Winckelmans
// evaluates every pair.
// The singular kernel's zeta is nonzero only at r == 0 and runs on the FP64
// kernel.
//
// F64: viscous.jl:500-509 statement by statement, storage order,
// contraction off.
#include <algorithm>
#include <cfloat>

#include "p2p_pair.h"

namespace vpm {


// Sorted source records: positions FP64 SoA, (Gamma, a) and (b, cut) FP32.
template <int K>
__global__ __launch_bounds__(kBlock) void pack_zeta_kernel(const double* __restrict__ P, int64_t np,
                                                           const uint32_t* __restrict__ perm,
                                                           double* __restrict__ pos, int64_t ld,
                                                           float4* __restrict__ gz, float2* __restrict__ bc)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= np) return;
    const int64_t p = perm ? (int64_t)perm[k] : k;
    const double* rec = P + p * NF;
    pos[k] = rec[RX];
    pos[ld + k] = rec[RX + 1];
    pos[2 * ld + k] = rec[RX + 2];
    const double sg = rec[RSIG];
    double a, b;
    zeta_params<K>(sg, a, b);
    gz[k] = make_float4((float)rec[RG], (float)rec[RG + 1], (float)rec[RG + 2], (float)a);
    bc[k] = make_float2((float)b, zeta_cut<K>(sg));
}

template <int K>
__global__ __launch_bounds__(kBlock) void zeta_f32c_kernel(const double* __restrict__ pos, int64_t ld,
                                                           const float4* __restrict__ gz,
                                                           const float2* __restrict__ bc,
                                                           const double* __restrict__ tbox, int64_t np,
                                                           int64_t chunk, double* __restrict__ part, int64_t pld)
{
    __shared__ float4 lp[kTile];   // s'_hi x, y, z, cut
    __shared__ float4 lg[kTile];   // Gamma x, y, z, b
    __shared__ float4 ll[kTile];   // s'_lo x, y, z, a
    __shared__ float wb[kBlock / 64][6];

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    // (target block, source chunk) from the linear block index renumbered
    // per XCD: the blocks one XCD runs at once are neighbouring target
    // blocks of one chunk, which share their near source tiles in its L2
    const int64_t L = xcd_block((int64_t)blockIdx.y * gridDim.x + blockIdx.x, (int64_t)gridDim.x * gridDim.y);
    const int64_t bx = L % gridDim.x, by = L / gridDim.x;
    const int64_t i = bx * kBlock + tid;
    const int64_t ib = bx * kBlock;
    const double ox = pos[ib], oy = pos[ld + ib], oz = pos[2 * ld + ib];
    const int64_t it = i < np ? i : ib;
    float tx, ty, tz, tlx, tly, tlz;
    split_rel(pos[it], ox, tx, tlx);
    split_rel(pos[ld + it], oy, ty, tly);
    split_rel(pos[2 * ld + it], oz, tz, tlz);
    const WaveBox box = wave_box(tx, ty, tz);
    float blo[3], bhi[3];
    block_box(box, wb, blo, bhi);
    const double org[3] = {ox, oy, oz};

    const int64_t sb = by * chunk;
    const int64_t se = min(sb + chunk, np);
    float sum[3] = {0.f, 0.f, 0.f}, err[3] = {0.f, 0.f, 0.f};
    __shared__ int nlist[kBlock];
    __shared__ int ncnt[kBlock / 64];
    NearTiles near(tbox, sb / kTile, (se + kTile - 1) / kTile, org, blo, bhi, 1.0001f, nlist, ncnt);
    for (int64_t tile; (tile = near.next()) >= 0;) {
        const int64_t base = tile * kTile;
        const int64_t j = base + tid;
        const int n = (int)min((int64_t)kTile, se - base);
        __syncthreads();
        if (tid < n) {
            const float4 g = gz[j];
            const float2 b = bc[j];
            float4 p, l;
            split_rel(pos[j], ox, p.x, l.x);
            split_rel(pos[ld + j], oy, p.y, l.y);
            split_rel(pos[2 * ld + j], oz, p.z, l.z);
            p.w = b.y;
            l.w = g.w;
            lp[tid] = p;
            lg[tid] = make_float4(g.x, g.y, g.z, b.x);
            ll[tid] = l;
        }
        __syncthreads();
        float acc[3] = {0.f, 0.f, 0.f};
#pragma unroll
        for (int grp = 0; grp < kTile / 64; ++grp) {
            const int s = grp * 64 + lane;
            bool nr = false;
            if (s < n) {
                const float4 a = lp[s];
                const float ex = fmaxf(fmaxf(box.lo[0] - a.x, a.x - box.hi[0]), 0.f);
                const float ey = fmaxf(fmaxf(box.lo[1] - a.y, a.y - box.hi[1]), 0.f);
                const float ez = fmaxf(fmaxf(box.lo[2] - a.z, a.z - box.hi[2]), 0.f);
                nr = fmaf(ez, ez, fmaf(ey, ey, ex * ex)) < a.w * 1.0001f;
            }
            uint64_t m = __ballot(nr);
            while (m) {
                const int q = grp * 64 + __builtin_ctzll(m);
                m &= m - 1;
                const float4 a = lp[q], g = lg[q], l = ll[q];
                const float dx = (tx - a.x) + (tlx - l.x);
                const float dy = (ty - a.y) + (tly - l.y);
                const float dz = (tz - a.z) + (tlz - l.z);
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                const float z = zeta_ab<K>(r2, l.w, g.w);
                acc[0] = fmaf(g.x, z, acc[0]);
                acc[1] = fmaf(g.y, z, acc[1]);
                acc[2] = fmaf(g.z, z, acc[2]);
            }
        }
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            const float t = sum[c] + acc[c];
            const float bp = t - sum[c];
            err[c] += (sum[c] - (t - bp)) + (acc[c] - bp);
            sum[c] = t;
        }
    }
    if (i < np) {
        double* out = part + by * 3 * pld + i;
#pragma unroll
        for (int c = 0; c < 3; ++c) out[c * pld] = (double)sum[c] + (double)err[c];
    }
}

#pragma clang fp contract(off)
// kernel.jl:45,51,60,69-74
template <int K>
__device__ __forceinline__ double zeta_ref(double r)
{
    if constexpr (K == VPM_KERNEL_SINGULAR) {
        return r == 0.0 ? 1.0 : 0.0;
    } else if constexpr (K == VPM_KERNEL_GAUSSIAN) {
        return kConst3 * exp(-r * r * r);
    } else if constexpr (K == VPM_KERNEL_GAUSSIANERF) {
        return kConst1 * exp(-r * r / 2);
    } else {
        double temp = r * r + 1;
        const double temp2 = temp * temp * temp;
        temp = temp2 * temp2 * temp;
        return kConst4 * 7.5 / sqrt(temp);
    }
}

// viscous.jl:500-509 in storage order
template <int K>
__global__ __launch_bounds__(kBlock) void zeta_f64_kernel(const double* __restrict__ P, int64_t np, int64_t chunk,
                                                          double* __restrict__ part, int64_t pld)
{
    __shared__ double4 ls[kTile];   // x, y, z, sigma
    __shared__ double4 lg[kTile];   // Gamma
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const int64_t it = i < np ? i : 0;
    const double* pi = P + it * NF;
    const double tx = pi[RX], ty = pi[RX + 1], tz = pi[RX + 2];
    const int64_t sb = (int64_t)blockIdx.y * chunk;
    const int64_t se = min(sb + chunk, np);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t base = sb; base < se; base += kTile) {
        const int n = (int)min((int64_t)kTile, se - base);
        __syncthreads();
        if (tid < n) {
            const double* pj = P + (base + tid) * NF;
            ls[tid] = make_double4(pj[RX], pj[RX + 1], pj[RX + 2], pj[RSIG]);
            lg[tid] = make_double4(pj[RG], pj[RG + 1], pj[RG + 2], 0.0);
        }
        __syncthreads();
        for (int s = 0; s < n; ++s) {
            const double4 a = ls[s];
            const double dX1 = tx - a.x, dX2 = ty - a.y, dX3 = tz - a.z;
            const double r = sqrt(dX1 * dX1 + dX2 * dX2 + dX3 * dX3);
            const double zeta_sgm = 1 / (a.w * a.w * a.w) * zeta_ref<K>(r / a.w);
            acc[0] += lg[s].x * zeta_sgm;
            acc[1] += lg[s].y * zeta_sgm;
            acc[2] += lg[s].z * zeta_sgm;
        }
    }
    if (i < np) {
        double* out = part + (int64_t)blockIdx.y * 3 * pld + i;
        for (int c = 0; c < 3; ++c) out[c * pld] = acc[c];
    }
}
#pragma clang fp contract(on)

// J[1:3] of record perm[k] (or k) <- sum of the K split partials
__global__ __launch_bounds__(kBlock) void zeta_finalize_kernel(const double* __restrict__ part, int64_t pld, int K,
                                                               int64_t np, const uint32_t* __restrict__ perm,
                                                               double* __restrict__ P)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= np) return;
    double a[3] = {0.0, 0.0, 0.0};
    for (int s = 0; s < K; ++s) {
        const double* p = part + (int64_t)s * 3 * pld + k;
        for (int c = 0; c < 3; ++c) a[c] += p[c * pld];
    }
    const int64_t idx = perm ? (int64_t)perm[k] : k;
    double* rec = P + idx * NF;
    for (int c = 0; c < 3; ++c) rec[RJ + c] = a[c];
}

int launch_zeta(vpm_field f, int kernel, int precision)
{
    const int64_t np = f->np;
    if (np == 0) return VPM_OK;
    const int64_t ld = f->cap;
    int rc;
    const unsigned nbt = (unsigned)((np + kBlock - 1) / kBlock);
    // source splits, chunks of whole tiles: enough blocks to fill the chip,
    // and up to 32 chunks of at least 16 tiles.  Near-field work per block is
    // uneven (a block whose targets straddle a jump of the sort order reaches
    // most tiles and finishes last); the chunks spread such a block's tiles
    // over that many blocks.  At N = 2^20 on the jet column: 21 ms with one
    // chunk, 12.4 with 8, 8.6 with 16, 6.6 with 32.
    const int64_t tiles = (np + kTile - 1) / kTile;
    const int64_t k_fill = std::min<int64_t>((2560 + nbt - 1) / nbt, tiles);
    const int64_t k_spread = std::min<int64_t>(32, tiles / 16);
    int64_t K = std::max<int64_t>(1, std::max(k_fill, k_spread));
    const int64_t chunk = ((tiles + K - 1) / K) * kTile;
    K = (np + chunk - 1) / chunk;
    if ((rc = ensure(f->part, sizeof(double) * 3 * (size_t)K * (size_t)np))) return rc;
    double* part = (double*)f->part.p;
    const dim3 grid(nbt, (unsigned)K);
    const uint32_t* perm = nullptr;
    hipError_t e;
    if (precision == VPM_PREC_F64 || kernel == VPM_KERNEL_SINGULAR) {
        timing_begin(f);
        switch (kernel) {
#define VPM_Z64(KID)                                                                                        \
    case KID:                                                                                               \
        hipLaunchKernelGGL(zeta_f64_kernel<KID>, grid, dim3(kBlock), 0, f->stream, f->P, np, chunk, part, np); \
        break;
            VPM_Z64(VPM_KERNEL_SINGULAR) VPM_Z64(VPM_KERNEL_GAUSSIAN) VPM_Z64(VPM_KERNEL_GAUSSIANERF)
            VPM_Z64(VPM_KERNEL_WINCKELMANS)
#undef VPM_Z64
        default:
            return fail(VPM_EINVAL, "unknown kernel id");
        }
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "zeta_f64");
        timing_end(f, 4);
    } else {
        if (np > 1 && (rc = sort_field(f, perm))) return rc;
        if ((rc = ensure(f->pos, sizeof(double) * 3 * (size_t)ld))) return rc;
        const int64_t ntile = (np + kTile - 1) / kTile;
        if ((rc = ensure(f->adj, (sizeof(float4) + sizeof(float2)) * (size_t)np + 64 * (size_t)ntile + 16)))
            return rc;
        float4* gz = (float4*)f->adj.p;
        float2* bc = (float2*)(gz + np);
        double* tbox = (double*)(bc + np);
        double* pos = (double*)f->pos.p;
        switch (kernel) {
#define VPM_Z32(KID)                                                                                           \
    case KID:                                                                                                  \
        hipLaunchKernelGGL(pack_zeta_kernel<KID>, dim3(nbt), dim3(kBlock), 0, f->stream, f->P, np, perm, pos, ld, \
                           gz, bc);                                                                            \
        hipLaunchKernelGGL(tile_box_kernel<2>, dim3((unsigned)ntile), dim3(kBlock), 0, f->stream,                \
                           (const double*)pos, ld, (const float*)bc + 1, np, tbox);                            \
        timing_begin(f);                                                                                       \
        hipLaunchKernelGGL(zeta_f32c_kernel<KID>, grid, dim3(kBlock), 0, f->stream, (const double*)pos, ld,     \
                           (const float4*)gz, (const float2*)bc, (const double*)tbox, np, chunk, part, np);  \
        break;
            VPM_Z32(VPM_KERNEL_GAUSSIAN) VPM_Z32(VPM_KERNEL_GAUSSIANERF) VPM_Z32(VPM_KERNEL_WINCKELMANS)
#undef VPM_Z32
        default:
            return fail(VPM_EINVAL, "unknown kernel id");
        }
        if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "zeta_f32c");
        timing_end(f, 4);
    }
    hipLaunchKernelGGL(zeta_finalize_kernel, dim3(nbt), dim3(kBlock), 0, f->stream, part, np, (int)K, np, perm,
                       f->P);
    if ((e = hipGetLastError()) != hipSuccess) return hip_fail(e, "zeta_finalize");
    return VPM_OK;
}

}  // namespace vpm
