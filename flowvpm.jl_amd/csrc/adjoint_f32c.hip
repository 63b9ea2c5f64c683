// adjoint_f32c.hip -- fused all-pairs adjoint of U and J in FP32 pair
// arithmetic with FP64 partial sums (the VPM_PREC_F32C path of vpm_uj_adjoint;
// replaces the ChainRules pullback UJS_pullback, src/FLOWVPM_rrules.jl:179-330).
//
// For L = sum_i Ubar_i . U_i + sum_i Jbar_i : J_i, with Gamma' = -Gamma/(4 pi),
// d = x_i - x_j (target i, source j), v = d x Gamma'_j, w = Jbar_i d,
// tau_i = (Jbar_23 - Jbar_32, Jbar_31 - Jbar_13, Jbar_12 - Jbar_21):
//     A + C = Ubar_i . v + tau_i . Gamma'_j,     B = v . w,
//     z     = F Ubar_i + H w,                   s = H (A + C) + K B,
//     dl/dd      = s d + Gamma'_j x z + H Jbar_i^T v
//     dl/dGamma' = z x d + F tau_i               (dl/dGamma = -dl/dGamma' / (4 pi))
//     dl/dsigma  = Fs (A + C) + Hs B
// with, in terms of q = r/sigma and the kernel's g(q),
//     F  = g/r^3,                      H  = (q g' - 3 g)/r^5,
//     K  = (q^2 g'' - 7 q g' + 15 g)/r^7,
//     Fs = -q g'/(sigma r^3),          Hs = -q (q g'' - 2 g')/(sigma r^5).
// dL/dx_i gets +dl/dd, dL/dx_j gets -dl/dd.
//
// One pass computes every pair once and feeds both sides:
//  * a block owns a (target super-tile, source super-tile) pair; sources are
//    staged through LDS in chunks of 256 (4 groups of 64, each row stored
//    twice so that index lane + k needs no wrap), lanes own targets;
//  * in a group, lane l visits source (l + k) mod 64 at step k; the seven
//    source-side sums travel with their source around the wave by a DPP
//    wave rotation fused into the accumulating add (v_add_f32 wave_rol:1),
//    so after 64 steps each lane holds the complete group sum of one source
//    and no cross-lane reduction is ever needed;
//  * the four waves take the four groups in staggered order (wave w, phase p
//    -> group (w + p) mod 4) with a barrier per phase, so each group's sums
//    are added into an FP64 LDS accumulator by one wave at a time, in a fixed
//    order: the result is deterministic;
//  * near/far split per (wave, group): if the box of the wave's targets and
//    the box of the group's sources are farther apart than the group's largest
//    far-field radius, every pair uses F = r^-3, H = -3 r^-5, K = 15 r^-7 and
//    no sigma term (Gaussian-erf q >= 7, Gaussian q >= 3; Winckelmans keeps its
//    full formula); otherwise the full evaluation with hi/lo positions;
//  * near-field reduced forms come from kernel_fits.h (phi0, phi1, phi2, chi)
//    wherever the closed forms cancel;
//  * per-target sums over a 256-source chunk and per-source sums over a
//    64-target group are FP32; everything above that is FP64.
#include <cfloat>

#include "p2p_pair.h"

namespace vpm {

constexpr int kAdjGroups = kTile / 64;   // groups of 64 sources per LDS chunk
constexpr int kAdjAcc = 7;               // source side: -dl/dd (3), dl/dGamma' (3), dl/dsigma

// DPP control "wave_rol:1": lane l receives lane (l + 1) mod 64.
constexpr int kDppWaveRol1 = 0x134;

__device__ __forceinline__ float rot1(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, x), kDppWaveRol1,
                                                                 0xf, 0xf, false));
}

// Near-field coefficients (F, H, K, Fs, Hs) at any q; all zero for r == 0
// (coincident pairs contribute nothing, fmm.jl:118).
template <int K>
__device__ __forceinline__ void adj_coef(float r2, float sinv, float& F, float& H, float& Kc, float& Fs, float& Hs);

template <>
__device__ __forceinline__ void adj_coef<VPM_KERNEL_SINGULAR>(float r2, float, float& F, float& H, float& Kc,
                                                              float& Fs, float& Hs)
{
    const float ri = r2 > 0.f ? rsq(r2) : 0.f;
    const float ri2 = ri * ri;
    F = ri2 * ri;
    H = -3.f * F * ri2;
    Kc = -5.f * H * ri2;
    Fs = 0.f;
    Hs = 0.f;
}

template <>
__device__ __forceinline__ void adj_coef<VPM_KERNEL_GAUSSIANERF>(float r2, float sinv, float& F, float& H,
                                                                 float& Kc, float& Fs, float& Hs)
{
    // g = erf(q/sqrt2) - c2 q E, q g' = c2 q^3 E, E = e^{-q^2/2}, c2 = sqrt(2/pi)
    const float ri = r2 > 0.f ? rsq(r2) : 0.f;
    const float q = r2 * ri * sinv;
    const float t = q * q;
    const float s2 = sinv * sinv, s3 = s2 * sinv, s4 = s2 * s2, s5 = s4 * sinv, s6 = s4 * s2, s7 = s6 * sinv;
    const float ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2, ri7 = ri5 * ri2;
    const float E = ex2(-0.5f * kLog2e * t);
    const float c2E = (float)kConst2 * E;
    const float c2qE = c2E * q;
    // erf(q/sqrt2) = 1 - E S(1/q)/q above the series (kErfcS)
    const float w = rsq(t);
    const float g = fmaf(-E, fmaf(horner(kErfcS, w), w, (float)kConst2 * q), 1.f);
    const float hm = fmaf(-(1.f / 3.f) * t, c2qE, g);          // g - q g'/3
    const float kn = fmaf(-c2qE * t, 5.f + t, 15.f * g);        // q^2 g'' - 7 q g' + 15 g
    const bool sm1 = t < kErfFitTMax, sm2 = t < kErfPhi2TMax;
    F = sm1 ? horner(kErfPhi0, t) * s3 : g * ri3;
    H = -3.f * (sm1 ? horner(kErfPhi1, t) * s5 : hm * ri5);
    Kc = sm2 ? horner(kErfPhi2, t) * s7 : kn * ri7;
    Fs = -c2E * s4;
    Hs = c2E * s6;
    if (!(r2 > 0.f)) { F = 0.f; H = 0.f; Kc = 0.f; Fs = 0.f; Hs = 0.f; }
}

template <>
__device__ __forceinline__ void adj_coef<VPM_KERNEL_GAUSSIAN>(float r2, float sinv, float& F, float& H, float& Kc,
                                                              float& Fs, float& Hs)
{
    // g = 1 - E, E = e^{-u}, u = q^3
    const float ri = r2 > 0.f ? rsq(r2) : 0.f;
    const float q = r2 * ri * sinv;
    const float u = q * q * q;
    const float s2 = sinv * sinv, s3 = s2 * sinv, s4 = s2 * s2, s5 = s4 * sinv, s6 = s4 * s2;
    const float ri2 = ri * ri, ri3 = ri2 * ri, ri5 = ri3 * ri2, ri7 = ri5 * ri2;
    const float E = ex2(-kLog2e * u);
    const bool sm = u < kGauFitUMax;
    F = sm ? horner(kGauPhi0, u) * s3 : (1.f - E) * ri3;
    H = -3.f * (sm ? q * horner(kGauPsi, u) * s5 : fmaf(-E, 1.f + u, 1.f) * ri5);
    Kc = sm ? horner(kGauChi, u) * s6 * ri : fmaf(-E, fmaf(9.f * u, u, fmaf(15.f, u, 15.f)), 15.f) * ri7;
    Fs = -3.f * E * s4;
    Hs = 9.f * q * E * s6;
    if (!(r2 > 0.f)) { F = 0.f; H = 0.f; Kc = 0.f; Fs = 0.f; Hs = 0.f; }
}

template <>
__device__ __forceinline__ void adj_coef<VPM_KERNEL_WINCKELMANS>(float r2, float sinv, float& F, float& H,
                                                                 float& Kc, float& Fs, float& Hs)
{
    // w = 1/(t + 1), t = q^2:  F = (t + 5/2) w^{5/2} sigma^-3,
    // H = -3 (t + 7/2) w^{7/2} sigma^-5, K = 15/2 (2 t + 9) w^{9/2} sigma^-7,
    // Fs = -15/2 w^{7/2} sigma^-4, Hs = 105/2 w^{9/2} sigma^-6
    const float s2 = sinv * sinv;
    const float t = r2 * s2;
    const float sr = rsq(t + 1.f);
    const float w = sr * sr;
    const float w25 = w * w * sr, w35 = w25 * w, w45 = w35 * w;
    const float s3 = s2 * sinv, s4 = s2 * s2, s5 = s4 * sinv, s6 = s4 * s2, s7 = s6 * sinv;
    F = (t + 2.5f) * w25 * s3;
    H = -3.f * (t + 3.5f) * w35 * s5;
    Kc = 7.5f * fmaf(2.f, t, 9.f) * w45 * s7;
    Fs = -7.5f * w35 * s4;
    Hs = 52.5f * w45 * s6;
    if (!(r2 > 0.f)) { F = 0.f; H = 0.f; Kc = 0.f; Fs = 0.f; Hs = 0.f; }
}

// A lane's target: position (hi/lo, relative to the chunk origin), Ubar,
// the columns of Jbar (jb[a] = Jbar[a][b]) and tau.
struct AdjTarget {
    float x, y, z, lx, ly, lz;
    float u0, u1, u2;
    float a0, a1, a2;   // column 0
    float b0, b1, b2;   // column 1
    float c0, c1, c2;   // column 2
    float t0, t1, t2;   // tau
};

// One pair: target T (the lane's), source (a = s'_hi, b = (Gamma', 1/sigma),
// l = s'_lo).  X gets +dl/dd; the rotating source sums R get -dl/dd,
// dl/dGamma', dl/dsigma.  JB = false is the loss on U alone (Jbar = 0, so
// tau = 0, w = 0, B = 0): the terms dropped are exact zeros, and every kept
// term is rounded as in the general form.
template <int K, bool NEAR, bool JB>
__device__ __forceinline__ void adj_pair(const AdjTarget& T, const float4 a, const float4 b, const float4 l,
                                         float (&X)[3], float (&R)[kAdjAcc])
{
    constexpr bool kCheap = !NEAR && Far<K>::cheap;
    constexpr bool kSigma = K != VPM_KERNEL_SINGULAR && !kCheap;
    float dx, dy, dz;
    if constexpr (NEAR) {
        dx = (T.x - a.x) + (T.lx - l.x);
        dy = (T.y - a.y) + (T.ly - l.y);
        dz = (T.z - a.z) + (T.lz - l.z);
    } else {
        dx = T.x - a.x;
        dy = T.y - a.y;
        dz = T.z - a.z;
    }
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    float F, H, Kc, Fs = 0.f, Hs = 0.f;
    if constexpr (kCheap) {
        const float ri = rsq(r2);
        const float ri2 = ri * ri;
        F = ri2 * ri;
        H = -3.f * F * ri2;
        Kc = -5.f * H * ri2;
    } else {
        adj_coef<K>(r2, b.w, F, H, Kc, Fs, Hs);
        if constexpr (Far<K>::cheap) {
            // pairs beyond the source's far radius take the far form in near
            // blocks too, so a pair's value does not depend on its block
            // (padding rows carry a zero radius: r2 == 0 keeps the zeros)
            if (r2 >= a.w && r2 > 0.f) {
                const float ri = rsq(r2);
                const float ri2 = ri * ri;
                F = ri2 * ri;
                H = -3.f * F * ri2;
                Kc = -5.f * H * ri2;
                Fs = 0.f;
                Hs = 0.f;
            }
        }
    }
    // v = d x Gamma'
    const float vx = fmaf(dy, b.z, -dz * b.y);
    const float vy = fmaf(dz, b.x, -dx * b.z);
    const float vz = fmaf(dx, b.y, -dy * b.x);
    if constexpr (!JB) {
        const float AC = fmaf(T.u2, vz, fmaf(T.u1, vy, T.u0 * vx));
        const float s = H * AC;
        const float zx = F * T.u0, zy = F * T.u1, zz = F * T.u2;
        const float gx = fmaf(s, dx, fmaf(b.y, zz, -b.z * zy));
        const float gy = fmaf(s, dy, fmaf(b.z, zx, -b.x * zz));
        const float gz = fmaf(s, dz, fmaf(b.x, zy, -b.y * zx));
        X[0] += gx;
        X[1] += gy;
        X[2] += gz;
        R[0] = rot1(R[0]) - gx;
        R[1] = rot1(R[1]) - gy;
        R[2] = rot1(R[2]) - gz;
        R[3] = rot1(R[3]) + fmaf(zy, dz, -zz * dy);
        R[4] = rot1(R[4]) + fmaf(zz, dx, -zx * dz);
        R[5] = rot1(R[5]) + fmaf(zx, dy, -zy * dx);
        // the product rounded on its own, as inside fmaf(Hs, B, Fs * AC) below
        if constexpr (kSigma) R[6] = rot1(R[6]) + __fmul_rn(Fs, AC);
        return;
    }
    // w = Jbar d
    const float wx = fmaf(T.c0, dz, fmaf(T.b0, dy, T.a0 * dx));
    const float wy = fmaf(T.c1, dz, fmaf(T.b1, dy, T.a1 * dx));
    const float wz = fmaf(T.c2, dz, fmaf(T.b2, dy, T.a2 * dx));
    const float AC = fmaf(T.u2, vz, fmaf(T.u1, vy, fmaf(T.u0, vx, fmaf(T.t2, b.z, fmaf(T.t1, b.y, T.t0 * b.x)))));
    const float B = fmaf(vz, wz, fmaf(vy, wy, vx * wx));
    const float s = fmaf(Kc, B, H * AC);
    // z = F Ubar + H w
    const float zx = fmaf(H, wx, F * T.u0);
    const float zy = fmaf(H, wy, F * T.u1);
    const float zz = fmaf(H, wz, F * T.u2);
    // Jbar^T v
    const float jx = fmaf(T.a2, vz, fmaf(T.a1, vy, T.a0 * vx));
    const float jy = fmaf(T.b2, vz, fmaf(T.b1, vy, T.b0 * vx));
    const float jz = fmaf(T.c2, vz, fmaf(T.c1, vy, T.c0 * vx));
    // dl/dd = s d + Gamma' x z + H Jbar^T v
    const float gx = fmaf(s, dx, fmaf(H, jx, fmaf(b.y, zz, -b.z * zy)));
    const float gy = fmaf(s, dy, fmaf(H, jy, fmaf(b.z, zx, -b.x * zz)));
    const float gz = fmaf(s, dz, fmaf(H, jz, fmaf(b.x, zy, -b.y * zx)));
    X[0] += gx;
    X[1] += gy;
    X[2] += gz;
    R[0] = rot1(R[0]) - gx;
    R[1] = rot1(R[1]) - gy;
    R[2] = rot1(R[2]) - gz;
    // dl/dGamma' = z x d + F tau
    R[3] = rot1(R[3]) + fmaf(zy, dz, fmaf(-zz, dy, F * T.t0));
    R[4] = rot1(R[4]) + fmaf(zz, dx, fmaf(-zx, dz, F * T.t1));
    R[5] = rot1(R[5]) + fmaf(zx, dy, fmaf(-zy, dx, F * T.t2));
    if constexpr (kSigma) R[6] = rot1(R[6]) + fmaf(Hs, B, Fs * AC);
}

// 64 steps over one group.  pa/pg/pl point at the lane's own row of the
// doubled group arrays; step k reads row lane + k.
template <int K, bool NEAR, bool JB>
__device__ __forceinline__ void adj_group(const AdjTarget& T, const float4* pa, const float4* pg, const float4* pl,
                                          float (&X)[3], float (&R)[kAdjAcc])
{
#pragma unroll 8
    for (int k = 0; k < 64; ++k) {
        const float4 l = NEAR ? pl[k] : make_float4(0.f, 0.f, 0.f, 0.f);
        adj_pair<K, NEAR, JB>(T, pa[k], pg[k], l, X, R);
    }
}

// partT [nss][3][np]: target-side sums per source super-tile.
// partS [ntt][7][np]: source-side sums per target super-tile.
// tcot [4][np] float4: (Ubar, tau_x), (Jbar col 0, tau_y), (col 1, tau_z), (col 2, 0).
// JB = false: every Jbar is zero, and only tcot[0] is read.
template <int K, bool JB>
__global__ __launch_bounds__(kBlock) void adj_f32c_kernel(
    const double* __restrict__ pos, int64_t ld, const float4* __restrict__ sgs, const float* __restrict__ scut,
    const float4* __restrict__ tcot, int64_t np, int64_t tsz, int64_t ssz, int ntt, int nss, int remap,
    double* __restrict__ partT, double* __restrict__ partS)
{
    __shared__ float4 lpos[kAdjGroups][128];   // s'_hi, far-field r^2 threshold
    __shared__ float4 lgs[kAdjGroups][128];    // Gamma', 1/sigma
    __shared__ float4 llo[kAdjGroups][128];    // s'_lo
    __shared__ float lbox[kAdjGroups][8];      // group box lo, hi, largest threshold
    __shared__ double lacc[kAdjAcc][kTile];    // source-side sums of the chunk

    const int tid = threadIdx.x;
    const int lane = tid & 63;
    const int wave = tid >> 6;
    // Blocks are dealt round-robin to the 8 XCDs; with `remap` the blocks of
    // one XCD get consecutive target super-tiles, which keeps the target data
    // they re-read in that XCD's L2.
    int64_t L = blockIdx.x;
    if (remap) L = (int64_t)(blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
    const int tt = (int)(L / nss);
    const int ss = (int)(L % nss);
    const int64_t tb = (int64_t)tt * tsz, te = min(tb + tsz, np);
    const int64_t sb = (int64_t)ss * ssz, se = min(sb + ssz, np);

    for (int64_t sc = sb; sc < se; sc += kTile) {
        const int nv = (int)min((int64_t)kTile, se - sc);
        const double ox = pos[sc], oy = pos[ld + sc], oz = pos[2 * ld + sc];
        __syncthreads();
        {
            // padding rows repeat the chunk's last source with zero strength
            const int64_t j = sc + min(tid, nv - 1);
            float4 p, lo;
            float4 g = sgs[j];
            split_rel(pos[j], ox, p.x, lo.x);
            split_rel(pos[ld + j], oy, p.y, lo.y);
            split_rel(pos[2 * ld + j], oz, p.z, lo.z);
            p.w = tid < nv ? scut[j] : 0.f;
            lo.w = 0.f;
            if (tid >= nv) { g.x = 0.f; g.y = 0.f; g.z = 0.f; }
            lpos[wave][lane] = p;
            lpos[wave][lane + 64] = p;
            lgs[wave][lane] = g;
            lgs[wave][lane + 64] = g;
            llo[wave][lane] = lo;
            llo[wave][lane + 64] = lo;
            const WaveBox bx = wave_box(p.x, p.y, p.z);
            float mc = p.w;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) mc = fmaxf(mc, __shfl_xor(mc, off));
            if (lane == 0) {
                lbox[wave][0] = bx.lo[0]; lbox[wave][1] = bx.lo[1]; lbox[wave][2] = bx.lo[2];
                lbox[wave][3] = bx.hi[0]; lbox[wave][4] = bx.hi[1]; lbox[wave][5] = bx.hi[2];
                lbox[wave][6] = mc;
            }
#pragma unroll
            for (int c = 0; c < kAdjAcc; ++c) lacc[c][tid] = 0.0;
        }
        __syncthreads();
        const bool first = sc == sb;
        for (int64_t tc = tb; tc < te; tc += kTile) {
            const int64_t i = tc + tid;
            const bool tvalid = i < te;
            const int64_t it = tvalid ? i : te - 1;
            AdjTarget T;
            split_rel(pos[it], ox, T.x, T.lx);
            split_rel(pos[ld + it], oy, T.y, T.ly);
            split_rel(pos[2 * ld + it], oz, T.z, T.lz);
            {
                const float m = tvalid ? 1.f : 0.f;   // padding targets carry zero cotangents
                const float4 c0 = tcot[it];
                T.u0 = m * c0.x; T.u1 = m * c0.y; T.u2 = m * c0.z; T.t0 = m * c0.w;
                if constexpr (JB) {
                    const float4 c1 = tcot[np + it], c2 = tcot[2 * np + it], c3 = tcot[3 * np + it];
                    T.a0 = m * c1.x; T.a1 = m * c1.y; T.a2 = m * c1.z; T.t1 = m * c1.w;
                    T.b0 = m * c2.x; T.b1 = m * c2.y; T.b2 = m * c2.z; T.t2 = m * c2.w;
                    T.c0 = m * c3.x; T.c1 = m * c3.y; T.c2 = m * c3.z;
                } else {
                    T.a0 = T.a1 = T.a2 = T.b0 = T.b1 = T.b2 = T.c0 = T.c1 = T.c2 = 0.f;
                    T.t1 = T.t2 = 0.f;
                }
            }
            double prev[3] = {0.0, 0.0, 0.0};
            if (tvalid && !first) {
#pragma unroll
                for (int c = 0; c < 3; ++c) prev[c] = partT[((int64_t)ss * 3 + c) * np + i];
            }
            const WaveBox tbx = wave_box(T.x, T.y, T.z);
            const bool wactive = tc + wave * 64 < te;
            float X[3] = {0.f, 0.f, 0.f};
            for (int ph = 0; ph < kAdjGroups; ++ph) {
                const int g = (wave + ph) & (kAdjGroups - 1);
                if (wactive && g * 64 < nv) {
                    const float* gb = lbox[g];
                    const float ex = fmaxf(fmaxf(tbx.lo[0] - gb[3], gb[0] - tbx.hi[0]), 0.f);
                    const float ey = fmaxf(fmaxf(tbx.lo[1] - gb[4], gb[1] - tbx.hi[1]), 0.f);
                    const float ez = fmaxf(fmaxf(tbx.lo[2] - gb[5], gb[2] - tbx.hi[2]), 0.f);
                    // 1e-4 margin covers the FP32 rounding of the box test
                    const bool near = fmaf(ez, ez, fmaf(ey, ey, ex * ex)) < gb[6] * 1.0001f;
                    float R[kAdjAcc];
#pragma unroll
                    for (int c = 0; c < kAdjAcc; ++c) R[c] = 0.f;
                    if (near)
                        adj_group<K, true, JB>(T, &lpos[g][lane], &lgs[g][lane], &llo[g][lane], X, R);
                    else
                        adj_group<K, false, JB>(T, &lpos[g][lane], &lgs[g][lane], &llo[g][lane], X, R);
                    // after 64 rotations lane l holds the sums of row (l - 1) mod 64
                    const int slot = g * 64 + ((lane + 63) & 63);
#pragma unroll
                    for (int c = 0; c < kAdjAcc; ++c) lacc[c][slot] += (double)R[c];
                }
                __syncthreads();
            }
            if (tvalid) {
#pragma unroll
                for (int c = 0; c < 3; ++c) partT[((int64_t)ss * 3 + c) * np + i] = prev[c] + (double)X[c];
            }
        }
        if (tid < nv) {
#pragma unroll
            for (int c = 0; c < kAdjAcc; ++c) partS[((int64_t)tt * kAdjAcc + c) * np + sc + tid] = lacc[c][tid];
        }
    }
}

// Sorted cotangents for the kernel above (perm: sorted index -> particle).
__global__ __launch_bounds__(kBlock) void pack_adj_cot_kernel(const double* __restrict__ Ubar,
                                                              const double* __restrict__ Jbar,
                                                              const uint32_t* __restrict__ perm, int64_t np,
                                                              float4* __restrict__ tcot)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= np) return;
    const int64_t p = perm ? (int64_t)perm[k] : k;
    const double* u = Ubar + 3 * p;
    const double* J = Jbar + 9 * p;   // column-major: J[a][b] = J[3 b + a]
    const double t0 = J[7] - J[5], t1 = J[2] - J[6], t2 = J[3] - J[1];
    tcot[k] = make_float4((float)u[0], (float)u[1], (float)u[2], (float)t0);
    tcot[np + k] = make_float4((float)J[0], (float)J[1], (float)J[2], (float)t1);
    tcot[2 * np + k] = make_float4((float)J[3], (float)J[4], (float)J[5], (float)t2);
    tcot[3 * np + k] = make_float4((float)J[6], (float)J[7], (float)J[8], 0.f);
}

// Sums the super-tile partials in fixed order and scatters to particle order.
__global__ __launch_bounds__(kBlock) void adj_finalize_kernel(const double* __restrict__ partT, int nss,
                                                              const double* __restrict__ partS, int ntt, int64_t np,
                                                              const uint32_t* __restrict__ perm,
                                                              double* __restrict__ Xbar, double* __restrict__ Gbar,
                                                              double* __restrict__ Sbar)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= np) return;
    double x[3] = {0.0, 0.0, 0.0}, g[3] = {0.0, 0.0, 0.0}, s = 0.0;
    for (int b = 0; b < nss; ++b) {
#pragma unroll
        for (int c = 0; c < 3; ++c) x[c] += partT[((int64_t)b * 3 + c) * np + k];
    }
    for (int b = 0; b < ntt; ++b) {
        const double* ps = partS + (int64_t)b * kAdjAcc * np + k;
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            x[c] += ps[c * np];
            g[c] += ps[(3 + c) * np];
        }
        s += ps[6 * np];
    }
    const int64_t p = perm ? (int64_t)perm[k] : k;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        Xbar[3 * p + c] = x[c];
        Gbar[3 * p + c] = -kConst4 * g[c];
    }
    Sbar[p] = s;
}

// Super-tiles, each a whole number of 256-particle chunks: up to 64 on the
// target side and up to 256 on the source side.  Blocks differ in near-field
// work, and more, shorter blocks balance the CUs better: at N = 2^20, 64 x 256
// runs 10.6 % faster than 64 x 64 (1,898 against 2,122 ms).  A source
// super-tile costs 3 FP64 partial sums per target and a target super-tile 7
// per source, so the source side grows only while the partials fit in
// 12 GiB (10.2 GB at 2^20), and never below 64.
static void adj_tiles(int64_t np, int64_t& tsz, int& ntt, int64_t& ssz, int& nss)
{
    const int64_t chunks = (np + kTile - 1) / kTile;
    const int64_t side_t = std::min<int64_t>(64, chunks);
    const int64_t per_part = 8 * np;                      // bytes of one partial row
    const int64_t budget = (int64_t)12 << 30;
    int64_t side_s = (budget / per_part - kAdjAcc * side_t) / 3;
    side_s = std::min<int64_t>(chunks, std::max<int64_t>(std::min<int64_t>(64, chunks), std::min<int64_t>(256, side_s)));
    tsz = ((chunks + side_t - 1) / side_t) * kTile;
    ssz = ((chunks + side_s - 1) / side_s) * kTile;
    ntt = (int)((np + tsz - 1) / tsz);
    nss = (int)((np + ssz - 1) / ssz);
}

int launch_uj_adjoint_f32c(vpm_field f, int kernel, const double* dU, const double* dJ, double* dX, double* dG,
                           double* dS, float4* tcot, bool jbar)
{
    const int64_t np = f->np;
    int rc;
    Staged S;
    if ((rc = stage_field(f, kernel, VPM_PREC_F32C, true, S))) return rc;
    const unsigned nb = (unsigned)((np + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(pack_adj_cot_kernel, dim3(nb), dim3(kBlock), 0, f->stream, dU, dJ, S.perm, np, tcot);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "pack_adj_cot");
    int64_t tsz, ssz;
    int ntt, nss;
    adj_tiles(np, tsz, ntt, ssz, nss);
    const size_t nT = (size_t)nss * 3 * (size_t)np, nS = (size_t)ntt * kAdjAcc * (size_t)np;
    if ((rc = ensure(f->part, sizeof(double) * (nT + nS)))) return rc;
    double* partT = (double*)f->part.p;
    double* partS = partT + nT;
    const unsigned grid = (unsigned)ntt * (unsigned)nss;
    const int remap = (grid % 8 == 0) ? 1 : 0;
    timing_begin(f);
    switch (kernel) {
#define VPM_ADJ32(KID)                                                                                        \
    case KID:                                                                                                 \
        if (jbar)                                                                                             \
            hipLaunchKernelGGL((adj_f32c_kernel<KID, true>), dim3(grid), dim3(kBlock), 0, f->stream, S.pos,     \
                               S.ld, S.gs, S.cut, (const float4*)tcot, np, tsz, ssz, ntt, nss, remap, partT,  \
                               partS);                                                                        \
        else                                                                                                  \
            hipLaunchKernelGGL((adj_f32c_kernel<KID, false>), dim3(grid), dim3(kBlock), 0, f->stream, S.pos,    \
                               S.ld, S.gs, S.cut, (const float4*)tcot, np, tsz, ssz, ntt, nss, remap, partT,  \
                               partS);                                                                        \
        break;
        VPM_ADJ32(VPM_KERNEL_SINGULAR)
        VPM_ADJ32(VPM_KERNEL_GAUSSIAN)
        VPM_ADJ32(VPM_KERNEL_GAUSSIANERF)
        VPM_ADJ32(VPM_KERNEL_WINCKELMANS)
#undef VPM_ADJ32
    default:
        return fail(VPM_EINVAL, "unknown kernel id");
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "adj_f32c");
    timing_end(f, 3);
    hipLaunchKernelGGL(adj_finalize_kernel, dim3(nb), dim3(kBlock), 0, f->stream, partT, nss, partS, ntt, np, S.perm,
                       dX, dG, dS);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "adj_finalize");
    return VPM_OK;
}

}  // namespace vpm
