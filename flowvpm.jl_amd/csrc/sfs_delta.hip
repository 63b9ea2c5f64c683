// sfs_delta.hip -- the test-minus-domain differences of the dynamic SFS
// procedure computed directly (dynamicprocedure_pseudo3level_beforeUJ /
// _afterUJ, src/FLOWVPM_subfilterscale.jl:447-673).
//
// The reference evaluates J and E_str twice, with the core sizes of the
// non-static particles scaled by alpha (test filter) and unscaled (domain
// filter), and the coefficient is built from the differences
//     dS = S(J_test) - S(J),   dE = SFS_test - SFS.
// For alpha close to 1 (SFS_Cd_twolevel_nobackscatter: 0.999) these are
// O(1 - alpha) of each term, so two FP32 evaluations would lose three
// digits.  Here the differences are evaluated as such:
//  * dJ: a pair changes only through g(q) with q = r/sigma; with
//    q' = q/alpha and k(q) = q g'(q),
//        dF'  = (g(q') - g(q))/r^3,   dH'' = ((g(q') - g(q)) - (k(q') - k(q))/3)/r^5,
//    where both differences are integrals of g' and of k' = g' + q g'' over
//    [q, q'], taken with 3-point Gauss-Legendre (exact to ~(q (1/alpha - 1))^6).
//    Pairs beyond the far radius have g == 1 at both widths (Gaussian-erf,
//    Gaussian) and contribute nothing; Winckelmans evaluates every pair.
//    Static sources keep their width and contribute nothing.
//  * dE: with J' = J - J_ref (the block's first target) and dJ likewise,
//        dE_i = sum_j dzeta_j (J'_i - J'_j)^T Gamma_j + zeta_t,j (dJ'_i - dJ'_j)^T Gamma_j,
//    dzeta = zeta_{alpha sigma}(r) - zeta_sigma(r) integrated over the width
//    (Gauss-Legendre in s), zeta_t = zeta_{alpha sigma}(r); static sources and
//    targets skipped as in Estr_direct!.
// The result is stored in M as the reference's test-filter values
// (M[0:3] = S(J + dJ), M[3:6] = SFS + dE, in FP64), so the coefficient
// kernel's subtraction of the domain values returns dS and dE to FP64
// rounding.  Used by the rk3 step for 1 - alpha <= 0.01.
#include <algorithm>
#include <cfloat>

#include "p2p_pair.h"

namespace vpm {

// 3-point Gauss-Legendre on [0, 1]
constexpr float kGlX0 = 0.1127016653792583f, kGlX1 = 0.5f, kGlX2 = 0.8872983346207417f;
constexpr float kGlW0 = 5.f / 18.f, kGlW1 = 8.f / 18.f;

// g'(s) and k'(s) = g'(s) + s g''(s) of the kernel at s
template <int K>
__device__ __forceinline__ void dg_dk(float s, float& gp, float& kp)
{
    const float s2 = s * s;
    if constexpr (K == VPM_KERNEL_GAUSSIANERF) {
        const float e = (float)kConst2 * ex2(-0.5f * kLog2e * s2);
        gp = e * s2;
        kp = e * s2 * (3.f - s2);
    } else if constexpr (K == VPM_KERNEL_GAUSSIAN) {
        const float s3 = s2 * s;
        const float e = ex2(-kLog2e * s3);
        gp = 3.f * s2 * e;
        kp = 9.f * s2 * (1.f - s3) * e;
    } else {   // Winckelmans
        const float w = 1.f / (s2 + 1.f);
        const float r = rsq(s2 + 1.f);
        const float w35 = w * w * w * r;
        gp = 7.5f * s2 * w35;
        kp = 7.5f * s2 * (3.f - 4.f * s2) * w35 * w;
    }
}

// (g(q') - g(q), k(q') - k(q)) for q' = q (1 + delta)
template <int K>
__device__ __forceinline__ void delta_gk(float q, float delta, float& dg, float& dk)
{
    const float w = q * delta;
    float g0, k0, g1, k1, g2, k2;
    dg_dk<K>(fmaf(w, kGlX0, q), g0, k0);
    dg_dk<K>(fmaf(w, kGlX1, q), g1, k1);
    dg_dk<K>(fmaf(w, kGlX2, q), g2, k2);
    dg = w * (kGlW0 * (g0 + g2) + kGlW1 * g1);
    dk = w * (kGlW0 * (k0 + k2) + kGlW1 * k1);
}

// zeta_s(r) = zeta(r/s)/s^3 and -d/ds of it, at width s (z0 = zeta(0) scale)
template <int K>
__device__ __forceinline__ void zeta_and_rate(float r2, float s, float& z, float& rate)
{
    const float si = 1.f / s;
    const float s3 = si * si * si;
    const float q2 = r2 * si * si;
    if constexpr (K == VPM_KERNEL_GAUSSIANERF) {
        z = (float)kConst1 * s3 * ex2(-0.5f * kLog2e * q2);
        rate = z * si * (3.f - q2);                  // s^-4 zeta (3 - q^2)
    } else if constexpr (K == VPM_KERNEL_GAUSSIAN) {
        const float q3 = q2 * sqrtf(q2);
        z = (float)kConst3 * s3 * ex2(-kLog2e * q3);
        rate = z * si * (3.f - 3.f * q3);
    } else {
        const float w = 1.f / (q2 + 1.f);
        const float r = rsq(q2 + 1.f);
        z = (float)(kConst4 * 7.5) * s3 * w * w * w * r;
        rate = z * si * w * (3.f - 4.f * q2);        // s^-4 c (q^2+1)^-4.5 (3 - 4 q^2)
    }
}

// -------------------------------------------------------------- dJ pass
// Sources: s'_hi (xyz, cut), (Gamma', 1/sigma), s'_lo; static sources carry
// Gamma' = 0.  Accumulators: dW' (3), dJhv (8, traceless).
template <int K>
__global__ __launch_bounds__(kBlock) void dj_f32c_kernel(const double* __restrict__ pos, int64_t ld,
                                                         const float4* __restrict__ sgs,
                                                         const float* __restrict__ scut,
                                                         const double* __restrict__ tbox, int64_t np, float delta,
                                                         int64_t chunk, double* __restrict__ part)
{
    __shared__ float4 lpos[kTile];
    __shared__ float4 lgs[kTile];
    __shared__ float4 llo[kTile];
    __shared__ float wb[kBlock / 64][6];
    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const int64_t ib = (int64_t)blockIdx.x * kBlock;
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
    constexpr int NA = 11;
    float sum[NA], err[NA];
#pragma unroll
    for (int c = 0; c < NA; ++c) { sum[c] = 0.f; err[c] = 0.f; }
    __shared__ int nlist[kBlock];
    __shared__ int ncnt[kBlock / 64];
    const int64_t s0 = (int64_t)blockIdx.y * chunk, s1 = min(s0 + chunk, np);
    NearTiles near(tbox, s0 / kTile, (s1 + kTile - 1) / kTile, org, blo, bhi, 1.0001f, nlist, ncnt);
    for (int64_t tile; (tile = near.next()) >= 0;) {
        const int64_t base = tile * kTile;
        const int n = (int)min((int64_t)kTile, s1 - base);
        __syncthreads();
        if (tid < n) {
            const int64_t j = base + tid;
            float4 p, l;
            split_rel(pos[j], ox, p.x, l.x);
            split_rel(pos[ld + j], oy, p.y, l.y);
            split_rel(pos[2 * ld + j], oz, p.z, l.z);
            p.w = scut[j];
            l.w = 0.f;
            lpos[tid] = p;
            lgs[tid] = sgs[j];
            llo[tid] = l;
        }
        __syncthreads();
        float acc[NA];
#pragma unroll
        for (int c = 0; c < NA; ++c) acc[c] = 0.f;
#pragma unroll
        for (int g = 0; g < kTile / 64; ++g) {
            const int s = g * 64 + lane;
            bool nr = false;
            if (s < n) {
                const float4 a = lpos[s];
                const float ex = fmaxf(fmaxf(box.lo[0] - a.x, a.x - box.hi[0]), 0.f);
                const float ey = fmaxf(fmaxf(box.lo[1] - a.y, a.y - box.hi[1]), 0.f);
                const float ez = fmaxf(fmaxf(box.lo[2] - a.z, a.z - box.hi[2]), 0.f);
                nr = fmaf(ez, ez, fmaf(ey, ey, ex * ex)) < a.w * 1.0001f;
            }
            uint64_t m = __ballot(nr);
            while (m) {
                const int q = g * 64 + __builtin_ctzll(m);
                m &= m - 1;
                const float4 a = lpos[q], b = lgs[q], lo = llo[q];
                const float dx = (tx - a.x) + (tlx - lo.x);
                const float dy = (ty - a.y) + (tly - lo.y);
                const float dz = (tz - a.z) + (tlz - lo.z);
                const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                if (!(r2 > 0.f)) continue;
                const float ri = rsq(r2);
                const float qq = r2 * ri * b.w;
                float dgv, dkv;
                delta_gk<K>(qq, delta, dgv, dkv);
                const float ri2 = ri * ri, ri3 = ri2 * ri;
                const float dF = dgv * ri3;
                const float dH = fmaf(-(1.f / 3.f), dkv, dgv) * ri3 * ri2;
                const float vx = fmaf(dy, b.z, -dz * b.y);
                const float vy = fmaf(dz, b.x, -dx * b.z);
                const float vz = fmaf(dx, b.y, -dy * b.x);
                acc[0] = fmaf(dF, b.x, acc[0]);
                acc[1] = fmaf(dF, b.y, acc[1]);
                acc[2] = fmaf(dF, b.z, acc[2]);
                const float hx = dH * vx, hy = dH * vy, hz = dH * vz;
                acc[3] = fmaf(hx, dx, acc[3]);
                acc[4] = fmaf(hy, dx, acc[4]);
                acc[5] = fmaf(hz, dx, acc[5]);
                acc[6] = fmaf(hx, dy, acc[6]);
                acc[7] = fmaf(hy, dy, acc[7]);
                acc[8] = fmaf(hz, dy, acc[8]);
                acc[9] = fmaf(hx, dz, acc[9]);
                acc[10] = fmaf(hy, dz, acc[10]);
            }
        }
#pragma unroll
        for (int c = 0; c < NA; ++c) {
            const float x = acc[c];
            const float t = sum[c] + x;
            const float bp = t - sum[c];
            err[c] += (sum[c] - (t - bp)) + (x - bp);
            sum[c] = t;
        }
    }
    if (i < np) {
        double* out = part + (int64_t)blockIdx.y * NA * np + i;
#pragma unroll
        for (int c = 0; c < NA; ++c) out[c * np] = (double)sum[c] + (double)err[c];
    }
}

// dJ from the chunk partials of dj_f32c_kernel, summed in chunk order.
__global__ __launch_bounds__(kBlock) void dj_finalize_kernel(const double* __restrict__ part, int nchunk, int64_t np,
                                                             double* __restrict__ dj)
{
    constexpr int NA = 11;
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i < np) {
        double a[NA];
#pragma unroll
        for (int c = 0; c < NA; ++c) a[c] = 0.0;
        for (int k = 0; k < nchunk; ++k) {
            const double* pk = part + (int64_t)k * NA * np + i;
#pragma unroll
            for (int c = 0; c < NA; ++c) a[c] += pk[c * np];
        }
        // dJ = -3 dJhv + eps . dW' (column-major), dJhv[2][2] = -dJhv00 - dJhv11
        const double Wx = a[0], Wy = a[1], Wz = a[2];
        double J[9];
        J[0] = -3.0 * a[3];
        J[1] = -3.0 * a[4] - Wz;
        J[2] = -3.0 * a[5] + Wy;
        J[3] = -3.0 * a[6] + Wz;
        J[4] = -3.0 * a[7];
        J[5] = -3.0 * a[8] - Wx;
        J[6] = -3.0 * a[9] - Wy;
        J[7] = -3.0 * a[10] + Wx;
        J[8] = 3.0 * (a[3] + a[7]);
#pragma unroll
        for (int c = 0; c < 9; ++c) dj[c * np + i] = J[c];
    }
}

// Source records of both passes, sorted order: Gamma' = -Gamma/(4 pi) and
// 1/sigma for dJ (zero strength for static particles), cut = 49 sigma^2 or
// 9 sigma^2 (FLT_MAX for Winckelmans); Gamma, sigma and J for dE.
template <int K>
__global__ __launch_bounds__(kBlock) void pack_sfsd_kernel(const double* __restrict__ P, int64_t np,
                                                           const uint32_t* __restrict__ perm,
                                                           double* __restrict__ pos, double* __restrict__ J,
                                                           int64_t ld, float4* __restrict__ gs, float* __restrict__ cut,
                                                           float4* __restrict__ gz)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= np) return;
    const int64_t p = perm ? (int64_t)perm[k] : k;
    const double* rec = P + p * NF;
    pos[k] = rec[RX];
    pos[ld + k] = rec[RX + 1];
    pos[2 * ld + k] = rec[RX + 2];
#pragma unroll
    for (int c = 0; c < 9; ++c) J[c * ld + k] = rec[RJ + c];
    const double sg = rec[RSIG];
    const bool st = rec[RSTATIC] != 0;
    const double m = st ? 0.0 : -kConst4;
    gs[k] = make_float4((float)(m * rec[RG]), (float)(m * rec[RG + 1]), (float)(m * rec[RG + 2]), (float)(1.0 / sg));
    cut[k] = K == VPM_KERNEL_WINCKELMANS ? FLT_MAX : (float)((K == VPM_KERNEL_GAUSSIAN ? 9.0 : 49.0) * sg * sg);
    // dE: Gamma and sigma; sigma = 0 marks a static source (skipped)
    gz[k] = make_float4((float)rec[RG], (float)rec[RG + 1], (float)rec[RG + 2], st ? 0.f : (float)sg);
}

// -------------------------------------------------------------- dE pass
// 12 accumulators: A1 = sum dzeta Gamma, B1 = sum dzeta S(J'_j, Gamma_j),
// A2 = sum zeta_t Gamma, B2 = sum zeta_t S(dJ'_j, Gamma_j).
template <int K>
__global__ __launch_bounds__(kBlock) void de_f32c_kernel(const double* __restrict__ pos, const double* __restrict__ J,
                                                         const double* __restrict__ dj, int64_t ld,
                                                         const float4* __restrict__ gz, const float* __restrict__ scut,
                                                         const double* __restrict__ tbox, int64_t np, float alpha,
                                                         int transposed, int64_t chunk, double* __restrict__ part)
{
    __shared__ float4 lp[kTile];    // s'_hi, sigma (0: static)
    __shared__ float4 lg[kTile];    // Gamma, cut
    __shared__ float4 l1[kTile];    // S(J'_j, Gamma_j)
    __shared__ float4 l2[kTile];    // S(dJ'_j, Gamma_j)
    __shared__ float4 ll[kTile];    // s'_lo
    __shared__ float wb[kBlock / 64][6];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const int64_t ib = (int64_t)blockIdx.x * kBlock;
    const double ox = pos[ib], oy = pos[ld + ib], oz = pos[2 * ld + ib];
    double Jr[9], dJr[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) { Jr[c] = J[c * ld + ib]; dJr[c] = dj[c * ld + ib]; }
    const int64_t it = i < np ? i : ib;
    float tx, ty, tz, tlx, tly, tlz;
    split_rel(pos[it], ox, tx, tlx);
    split_rel(pos[ld + it], oy, ty, tly);
    split_rel(pos[2 * ld + it], oz, tz, tlz);
    float blo[3], bhi[3];
    block_box(wave_box(tx, ty, tz), wb, blo, bhi);
    const double org[3] = {ox, oy, oz};
    float sum[12], err[12];
#pragma unroll
    for (int c = 0; c < 12; ++c) { sum[c] = 0.f; err[c] = 0.f; }
    const float oma = 1.f - alpha;
    __shared__ int nlist[kBlock];
    __shared__ int ncnt[kBlock / 64];
    const int64_t s0 = (int64_t)blockIdx.y * chunk, s1 = min(s0 + chunk, np);
    NearTiles near(tbox, s0 / kTile, (s1 + kTile - 1) / kTile, org, blo, bhi, 1.0f, nlist, ncnt);
    for (int64_t tile; (tile = near.next()) >= 0;) {
        const int64_t base = tile * kTile;
        const int n = (int)min((int64_t)kTile, s1 - base);
        __syncthreads();
        if (tid < n) {
            const int64_t j = base + tid;
            const float4 g = gz[j];
            float4 p, l;
            split_rel(pos[j], ox, p.x, l.x);
            split_rel(pos[ld + j], oy, p.y, l.y);
            split_rel(pos[2 * ld + j], oz, p.z, l.z);
            p.w = g.w;
            l.w = 0.f;
            double Jp[9], dJp[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) { Jp[c] = J[c * ld + j] - Jr[c]; dJp[c] = dj[c * ld + j] - dJr[c]; }
            const double G0 = g.x, G1 = g.y, G2 = g.z;
            float4 s1, s2;
            if (transposed) {
                s1 = make_float4((float)(Jp[0] * G0 + Jp[1] * G1 + Jp[2] * G2), (float)(Jp[3] * G0 + Jp[4] * G1 + Jp[5] * G2),
                                 (float)(Jp[6] * G0 + Jp[7] * G1 + Jp[8] * G2), 0.f);
                s2 = make_float4((float)(dJp[0] * G0 + dJp[1] * G1 + dJp[2] * G2),
                                 (float)(dJp[3] * G0 + dJp[4] * G1 + dJp[5] * G2),
                                 (float)(dJp[6] * G0 + dJp[7] * G1 + dJp[8] * G2), 0.f);
            } else {
                s1 = make_float4((float)(Jp[0] * G0 + Jp[3] * G1 + Jp[6] * G2), (float)(Jp[1] * G0 + Jp[4] * G1 + Jp[7] * G2),
                                 (float)(Jp[2] * G0 + Jp[5] * G1 + Jp[8] * G2), 0.f);
                s2 = make_float4((float)(dJp[0] * G0 + dJp[3] * G1 + dJp[6] * G2),
                                 (float)(dJp[1] * G0 + dJp[4] * G1 + dJp[7] * G2),
                                 (float)(dJp[2] * G0 + dJp[5] * G1 + dJp[8] * G2), 0.f);
            }
            lp[tid] = p;
            lg[tid] = make_float4(g.x, g.y, g.z, scut[j]);
            l1[tid] = s1;
            l2[tid] = s2;
            ll[tid] = l;
        }
        __syncthreads();
        for (int s0 = 0; s0 < n; s0 += 32) {
            const int s1e = min(s0 + 32, n);
            float acc[12];
#pragma unroll
            for (int c = 0; c < 12; ++c) acc[c] = 0.f;
            for (int s = s0; s < s1e; ++s) {
                const float4 a = lp[s];
                const float4 g = lg[s];
                float dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
                float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                if (__any(r2 < g.w && a.w > 0.f)) {
                    if (!(a.w > 0.f)) continue;   // static source: no E_str term
                    const float4 lo = ll[s];
                    dx += tlx - lo.x;
                    dy += tly - lo.y;
                    dz += tlz - lo.z;
                    r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    // dzeta = zeta_{alpha sigma} - zeta_sigma = int_{alpha sigma}^{sigma} rate(s) ds
                    const float sg = a.w, w = oma * sg;
                    float zt, r0, r1, r2v, z_;
                    zeta_and_rate<K>(r2, alpha * sg, zt, r0);
                    zeta_and_rate<K>(r2, fmaf(w, kGlX1, alpha * sg), z_, r1);
                    zeta_and_rate<K>(r2, fmaf(w, kGlX2, alpha * sg), z_, r2v);
                    float rq;
                    zeta_and_rate<K>(r2, fmaf(w, kGlX0, alpha * sg), z_, rq);
                    float dz_ = w * (kGlW0 * (rq + r2v) + kGlW1 * r1);
                    // the self pair is exactly zero in the reference ((J_i -
                    // J_i) and (dJ_i - dJ_i)); the factored form would add it
                    // as cancelling terms
                    if (base + s == i) { dz_ = 0.f; zt = 0.f; }
                    const float4 c1 = l1[s], c2 = l2[s];
                    acc[0] = fmaf(dz_, g.x, acc[0]);
                    acc[1] = fmaf(dz_, g.y, acc[1]);
                    acc[2] = fmaf(dz_, g.z, acc[2]);
                    acc[3] = fmaf(dz_, c1.x, acc[3]);
                    acc[4] = fmaf(dz_, c1.y, acc[4]);
                    acc[5] = fmaf(dz_, c1.z, acc[5]);
                    acc[6] = fmaf(zt, g.x, acc[6]);
                    acc[7] = fmaf(zt, g.y, acc[7]);
                    acc[8] = fmaf(zt, g.z, acc[8]);
                    acc[9] = fmaf(zt, c2.x, acc[9]);
                    acc[10] = fmaf(zt, c2.y, acc[10]);
                    acc[11] = fmaf(zt, c2.z, acc[11]);
                }
            }
#pragma unroll
            for (int c = 0; c < 12; ++c) {
                const float x = acc[c];
                const float t = sum[c] + x;
                const float bp = t - sum[c];
                err[c] += (sum[c] - (t - bp)) + (x - bp);
                sum[c] = t;
            }
        }
    }
    if (i < np) {
        double* o = part + (int64_t)blockIdx.y * 12 * np + i;
#pragma unroll
        for (int c = 0; c < 12; ++c) o[c * np] = (double)sum[c] + (double)err[c];
    }
}

// dE from the chunk partials of de_f32c_kernel, summed in chunk order; J'
// is relative to the J of the first target of the 256-target block, as in
// the kernel.
__global__ __launch_bounds__(kBlock) void de_finalize_kernel(const double* __restrict__ part, int nchunk,
                                                             const double* __restrict__ J,
                                                             const double* __restrict__ dj, int64_t ld, int64_t np,
                                                             int transposed, double* __restrict__ out)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    const int64_t ib = (int64_t)blockIdx.x * kBlock;
    if (i < np) {
        double A[12];
#pragma unroll
        for (int c = 0; c < 12; ++c) A[c] = 0.0;
        for (int k = 0; k < nchunk; ++k) {
            const double* pk = part + (int64_t)k * 12 * np + i;
#pragma unroll
            for (int c = 0; c < 12; ++c) A[c] += pk[c * np];
        }
        double Jr[9], dJr[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) { Jr[c] = J[c * ld + ib]; dJr[c] = dj[c * ld + ib]; }
        double Jp[9], dJp[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) { Jp[c] = J[c * ld + i] - Jr[c]; dJp[c] = dj[c * ld + i] - dJr[c]; }
        for (int q = 0; q < 3; ++q) {
            double t1, t2;
            if (transposed) {
                t1 = Jp[3 * q] * A[0] + Jp[3 * q + 1] * A[1] + Jp[3 * q + 2] * A[2];
                t2 = dJp[3 * q] * A[6] + dJp[3 * q + 1] * A[7] + dJp[3 * q + 2] * A[8];
            } else {
                t1 = Jp[q] * A[0] + Jp[3 + q] * A[1] + Jp[6 + q] * A[2];
                t2 = dJp[q] * A[6] + dJp[3 + q] * A[7] + dJp[6 + q] * A[8];
            }
            out[q * ld + i] = (t1 - A[3 + q]) + (t2 - A[9 + q]);
        }
    }
}

// M of non-static particles <- the reference's test-filter values:
// M[0:3] = S(J + dJ, Gamma), M[3:6] = SFS + dE, M[6:9] = 0.
__global__ __launch_bounds__(kBlock) void sfsd_store_kernel(const double* __restrict__ dj,
                                                            const double* __restrict__ de, int64_t ld, int64_t np,
                                                            const uint32_t* __restrict__ perm, int transposed,
                                                            double* __restrict__ P)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= np) return;
    const int64_t p = perm ? (int64_t)perm[k] : k;
    double* rec = P + p * NF;
    if (rec[RSTATIC] != 0) return;
    double Jt[9];
    for (int c = 0; c < 9; ++c) Jt[c] = rec[RJ + c] + dj[c * ld + k];
    const double* G = rec + RG;
    double S[3];
    for (int q = 0; q < 3; ++q)
        S[q] = transposed ? Jt[3 * q] * G[0] + Jt[3 * q + 1] * G[1] + Jt[3 * q + 2] * G[2]
                          : Jt[q] * G[0] + Jt[3 + q] * G[1] + Jt[6 + q] * G[2];
    double* M = rec + RM;
    for (int c = 0; c < 9; ++c) M[c] = 0.0;
    M[0] = S[0]; M[1] = S[1]; M[2] = S[2];
    M[3] = rec[RSFS] + de[k];
    M[4] = rec[RSFS + 1] + de[ld + k];
    M[5] = rec[RSFS + 2] + de[2 * ld + k];
}

int launch_sfs_delta(vpm_field f, int kernel, int transposed, double alpha)
{
    const int64_t np = f->np;
    if (np == 0) return VPM_OK;
    if (kernel == VPM_KERNEL_SINGULAR) return fail(VPM_EINVAL, "sfs difference kernels: singular kernel");
    const int64_t ld = np;
    int rc;
    const uint32_t* perm = nullptr;
    if (np > 1 && (rc = sort_field(f, perm))) return rc;
    // scratch: pos 3, J 9, dJ 9, dE 3 doubles; gs, gz float4; tile boxes; cut float
    const size_t nd = 24 * (size_t)np;
    const int64_t ntile = (np + kTile - 1) / kTile;
    const size_t bytes = sizeof(double) * nd + (2 * sizeof(float4) + sizeof(float)) * (size_t)np +
                         64 * (size_t)ntile + 64;
    if ((rc = ensure(f->adj, bytes))) return rc;
    double* pos = (double*)f->adj.p;
    double* J = pos + 3 * np;
    double* dj = J + 9 * np;
    double* de = dj + 9 * np;
    float4* gs = (float4*)(de + 3 * np);
    float4* gz = gs + np;
    double* tbox = (double*)(gz + np);
    float* cut = (float*)(tbox + 8 * ntile);
    const unsigned nb = (unsigned)((np + kBlock - 1) / kBlock);
    const float delta = (float)(1.0 / alpha - 1.0);
    // source chunks as for zeta and E_str (up to 32 of at least 16 tiles)
    const int64_t k_fill = std::min<int64_t>((2560 + nb - 1) / nb, ntile);
    const int64_t k_spread = std::min<int64_t>(32, ntile / 16);
    int64_t nchunk = std::max<int64_t>(1, std::max(k_fill, k_spread));
    const int64_t chunk = ((ntile + nchunk - 1) / nchunk) * kTile;
    nchunk = (np + chunk - 1) / chunk;
    if ((rc = ensure(f->part, sizeof(double) * 12 * (size_t)nchunk * (size_t)np))) return rc;
    double* part = (double*)f->part.p;
    const dim3 grid(nb, (unsigned)nchunk);
    switch (kernel) {
#define VPM_SD(KID)                                                                                              \
    case KID:                                                                                                    \
        hipLaunchKernelGGL(pack_sfsd_kernel<KID>, dim3(nb), dim3(kBlock), 0, f->stream, f->P, np, perm, pos, J, ld, \
                           gs, cut, gz);                                                                         \
        hipLaunchKernelGGL(tile_box_kernel<1>, dim3((unsigned)ntile), dim3(kBlock), 0, f->stream,                \
                           (const double*)pos, ld, (const float*)cut, np, tbox);                               \
        hipLaunchKernelGGL(dj_f32c_kernel<KID>, grid, dim3(kBlock), 0, f->stream, (const double*)pos, ld,         \
                           (const float4*)gs, (const float*)cut, (const double*)tbox, np, delta, chunk, part);   \
        hipLaunchKernelGGL(dj_finalize_kernel, dim3(nb), dim3(kBlock), 0, f->stream, (const double*)part,         \
                           (int)nchunk, np, dj);                                                                 \
        hipLaunchKernelGGL(de_f32c_kernel<KID>, grid, dim3(kBlock), 0, f->stream, (const double*)pos,             \
                           (const double*)J, (const double*)dj, ld, (const float4*)gz, (const float*)cut,         \
                           (const double*)tbox, np, (float)alpha, transposed, chunk, part);                      \
        hipLaunchKernelGGL(de_finalize_kernel, dim3(nb), dim3(kBlock), 0, f->stream, (const double*)part,         \
                           (int)nchunk, (const double*)J, (const double*)dj, ld, np, transposed, de);           \
        break;
        VPM_SD(VPM_KERNEL_GAUSSIAN) VPM_SD(VPM_KERNEL_GAUSSIANERF) VPM_SD(VPM_KERNEL_WINCKELMANS)
#undef VPM_SD
    default:
        return fail(VPM_EINVAL, "unknown kernel id");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "sfs difference kernels");
    hipLaunchKernelGGL(sfsd_store_kernel, dim3(nb), dim3(kBlock), 0, f->stream, (const double*)dj, (const double*)de,
                       ld, np, perm, transposed, f->P);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "sfsd_store");
    return VPM_OK;
}

}  // namespace vpm
