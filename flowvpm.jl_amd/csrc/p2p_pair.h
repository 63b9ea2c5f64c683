// p2p_pair.h -- per-pair device arithmetic shared by the all-pairs and the
// leaf-pair-list U/J kernels (src/FLOWVPM_fmm.jl:113-161).
#pragma once
#include <cfloat>

#include "kernel_fits.h"
#include "vpm_internal.h"

namespace vpm {

// ---------------------------------------------------------------- FP32 math
__device__ __forceinline__ float rsq(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float ex2(float x) { return __builtin_amdgcn_exp2f(x); }
constexpr float kLog2e = 1.4426950408889634f;

template <int N>
__device__ __forceinline__ float horner(const float (&c)[N], float x)
{
    float a = c[N - 1];
#pragma unroll
    for (int k = N - 2; k >= 0; --k) a = fmaf(a, x, c[k]);
    return a;
}

// Near/far split per source: a source is "near" a wave when some lane has
// r^2 < q_far^2 sigma^2.  Near sources take the full kernel evaluation and
// the hi/lo position correction; far sources use the cheap form.
//  * Gaussian-erf: for q >= 7, 1 - g < 2e-10 and |q g'| < 3e-9, so g = 1 and
//    q g' - 3 g = -3 to FP32 rounding: F' = r^-3, H'' = r^-5.
//  * Gaussian: for q >= 3, e^{-q^3} (1 + q^3) < 6e-11: same far form.
//  * Winckelmans: g -> 1 only as 1 - 1.9/q^4, so the far form is the full
//    formula; q_far = 7 only decides where the hi/lo correction is applied.
//  * singular: every source is near (its pair error grows like 1/r).
template <int K> struct Far;
template <> struct Far<VPM_KERNEL_GAUSSIANERF> { static constexpr float q2 = 49.f; static constexpr bool cheap = true; };
template <> struct Far<VPM_KERNEL_GAUSSIAN> { static constexpr float q2 = 9.f; static constexpr bool cheap = true; };
template <> struct Far<VPM_KERNEL_WINCKELMANS> { static constexpr float q2 = 49.f; static constexpr bool cheap = false; };
template <> struct Far<VPM_KERNEL_SINGULAR> { static constexpr float q2 = 0.f; static constexpr bool cheap = false; };

// far-field r^2 threshold of a source of core size sigma
template <int K>
__host__ __device__ __forceinline__ float far_cut(double sigma)
{
    if (K == VPM_KERNEL_SINGULAR) return FLT_MAX;
    const float c = (float)((double)Far<K>::q2 * sigma * sigma);
    return c > FLT_MIN ? c : FLT_MIN;
}

// Full (any-q) evaluation of F' = g/r^3 and H'' = -(q g' - 3g)/(3 r^5).
// Coincident pairs (r2 == 0) contribute nothing (fmm.jl:118).
template <int K>
__device__ __forceinline__ void coef_full(float r2, float sinv, float& Fp, float& Hpp);

template <>
__device__ __forceinline__ void coef_full<VPM_KERNEL_SINGULAR>(float r2, float, float& Fp, float& Hpp)
{
    const float ri = r2 > 0.f ? rsq(r2) : 0.f;
    const float ri2 = ri * ri;
    Fp = ri2 * ri;
    Hpp = Fp * ri2;
}

template <>
__device__ __forceinline__ void coef_full<VPM_KERNEL_GAUSSIANERF>(float r2, float sinv, float& Fp, float& Hpp)
{
    // g = erf(q/sqrt2) - sqrt(2/pi) q e^{-q^2/2},  q g' = sqrt(2/pi) q^3 e^{-q^2/2};
    // above the series, erfc(q/sqrt2) = E S(1/q)/q with E = e^{-q^2/2}, so
    // g = 1 - E (S(1/q)/q + sqrt(2/pi) q) (kErfcS; the unselected value at
    // small q may be non-finite)
    const float ri = r2 > 0.f ? rsq(r2) : 0.f;
    const float q = r2 * ri * sinv;
    const float t = q * q;
    const float s3 = sinv * sinv * sinv;
    const float s5 = s3 * sinv * sinv;
    const float Fs = horner(kErfPhi0, t) * s3;
    const float Hs = horner(kErfPhi1, t) * s5;
    const float E = ex2(-0.5f * kLog2e * t);
    const float c2q = (float)kConst2 * q;
    const float w = rsq(t);
    const float g = fmaf(-E, fmaf(horner(kErfcS, w), w, c2q), 1.f);
    const float hm = fmaf(-(1.f / 3.f) * t, c2q * E, g);    // g - q g'/3
    const float ri2 = ri * ri;
    const float ri3 = ri2 * ri;
    const bool small = t < kErfFitTMax;
    Fp = small ? Fs : g * ri3;
    Hpp = small ? Hs : hm * ri3 * ri2;
    if (!(r2 > 0.f)) { Fp = 0.f; Hpp = 0.f; }
}

template <>
__device__ __forceinline__ void coef_full<VPM_KERNEL_GAUSSIAN>(float r2, float sinv, float& Fp, float& Hpp)
{
    // g = 1 - e^{-u}, u = q^3;  -(q g' - 3 g)/3 = 1 - e^{-u}(1 + u)
    const float ri = r2 > 0.f ? rsq(r2) : 0.f;
    const float q = r2 * ri * sinv;
    const float u = q * q * q;
    const float s3 = sinv * sinv * sinv;
    const float s5 = s3 * sinv * sinv;
    const float Fs = horner(kGauPhi0, u) * s3;
    const float Hs = q * horner(kGauPsi, u) * s5;
    const float E = ex2(-kLog2e * u);
    const float ri2 = ri * ri;
    const float ri3 = ri2 * ri;
    const bool small = u < kGauFitUMax;
    Fp = small ? Fs : (1.f - E) * ri3;
    Hpp = small ? Hs : fmaf(-E, 1.f + u, 1.f) * ri3 * ri2;
    if (!(r2 > 0.f)) { Fp = 0.f; Hpp = 0.f; }
}

template <>
__device__ __forceinline__ void coef_full<VPM_KERNEL_WINCKELMANS>(float r2, float sinv, float& Fp, float& Hpp)
{
    // phi0 = (t + 5/2)/(t + 1)^{5/2},  -phi1/3 = (t + 7/2)/(t + 1)^{7/2},  t = q^2
    const float s2 = sinv * sinv;
    const float t = r2 * s2;
    const float sr = rsq(t + 1.f);
    const float w = sr * sr;
    const float w2s = w * w * sr;
    const float s3 = s2 * sinv;
    Fp = (t + 2.5f) * w2s * s3;
    Hpp = (t + 3.5f) * w2s * w * s3 * s2;
    if (!(r2 > 0.f)) { Fp = 0.f; Hpp = 0.f; }
}

// Accumulates one pair given its separation d and coefficients F', H''.
__device__ __forceinline__ void accum_pair(float dx, float dy, float dz, float Fp, float Hpp, const float4 b,
                                           float (&acc)[kAccF32])
{
    const float vx = fmaf(dy, b.z, -dz * b.y);
    const float vy = fmaf(dz, b.x, -dx * b.z);
    const float vz = fmaf(dx, b.y, -dy * b.x);
    acc[0] = fmaf(Fp, vx, acc[0]);
    acc[1] = fmaf(Fp, vy, acc[1]);
    acc[2] = fmaf(Fp, vz, acc[2]);
    acc[3] = fmaf(Fp, b.x, acc[3]);
    acc[4] = fmaf(Fp, b.y, acc[4]);
    acc[5] = fmaf(Fp, b.z, acc[5]);
    const float hx = Hpp * vx, hy = Hpp * vy, hz = Hpp * vz;
    acc[6] = fmaf(hx, dx, acc[6]);     // Jhv[0][0]
    acc[7] = fmaf(hy, dx, acc[7]);     // Jhv[1][0]
    acc[8] = fmaf(hz, dx, acc[8]);     // Jhv[2][0]
    acc[9] = fmaf(hx, dy, acc[9]);     // Jhv[0][1]
    acc[10] = fmaf(hy, dy, acc[10]);   // Jhv[1][1]
    acc[11] = fmaf(hz, dy, acc[11]);   // Jhv[2][1]
    acc[12] = fmaf(hx, dz, acc[12]);   // Jhv[0][2]
    acc[13] = fmaf(hy, dz, acc[13]);   // Jhv[1][2]
}

// Far form of one pair: F' = r^-3, H'' = r^-5 (kernels with Far<K>::cheap),
// or the full formula without the hi/lo correction (Winckelmans).
template <int K>
__device__ __forceinline__ void far_pair(float tx, float ty, float tz, const float4 a, const float4 b,
                                         float (&acc)[kAccF32])
{
    const float dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    float Fp, Hpp;
    if constexpr (Far<K>::cheap) {
        const float ri = rsq(r2);
        const float ri2 = ri * ri;
        Fp = ri2 * ri;
        Hpp = Fp * ri2;
    } else {
        coef_full<K>(r2, b.w, Fp, Hpp);
    }
    accum_pair(dx, dy, dz, Fp, Hpp, b, acc);
}

// Near form: d from the hi/lo split (t - O = t'_hi + t'_lo to ~2^-48, so d
// keeps FP32 relative accuracy however close the pair), full kernel formula.
template <int K>
__device__ __forceinline__ void near_pair(float tx, float ty, float tz, float tlx, float tly, float tlz,
                                          const float4 a, const float4 b, const float4 l, float (&acc)[kAccF32])
{
    const float dx = (tx - a.x) + (tlx - l.x);
    const float dy = (ty - a.y) + (tly - l.y);
    const float dz = (tz - a.z) + (tlz - l.z);
    const float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    float Fp, Hpp;
    coef_full<K>(r2, b.w, Fp, Hpp);
    accum_pair(dx, dy, dz, Fp, Hpp, b, acc);
}

// Bounding box of the wave's targets (t'_hi), computed once per block; the
// result is wave-uniform and is kept in scalar registers.
struct WaveBox { float lo[3], hi[3]; };

__device__ __forceinline__ float uniform_f(float x)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(int, x)));
}

__device__ __forceinline__ WaveBox wave_box(float tx, float ty, float tz)
{
    WaveBox b;
    float v[3] = {tx, ty, tz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        float mn = v[c], mx = v[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn = fminf(mn, __shfl_xor(mn, off));
            mx = fmaxf(mx, __shfl_xor(mx, off));
        }
        b.lo[c] = uniform_f(mn);
        b.hi[c] = uniform_f(mx);
    }
    return b;
}

// TwoSum fold of FP32 partials into (sum, err)
__device__ __forceinline__ void fold_twosum(const float (&acc)[kAccF32], float (&sum)[kAccF32],
                                            float (&err)[kAccF32])
{
#pragma unroll
    for (int c = 0; c < kAccF32; ++c) {
        const float x = acc[c];
        const float t = sum[c] + x;
        const float bp = t - sum[c];
        err[c] += (sum[c] - (t - bp)) + (x - bp);
        sum[c] = t;
    }
}

// Near pairs per fold of the U partials (tile_f32c over 256-source tiles).
// On the N = 65,536 random cloud, whose U sums cancel by up to ~2,400x at
// some particles, the worst particle's U error against FP64 falls from
// 1.5e-5 (one fold per tile) to 8.1e-6, measured on the GPU; an FP32
// emulation of the kernel gave 8.7, 8.6, 7.9, 9.0 and 10.3e-6 for folds
// every 1, 2, 4, 8 and 16 near pairs, what remains being the rounding of
// the terms themselves.  Folding U alone costs 0.7 % on the jet and 5 % on
// that cloud; folding all 14 accumulators also takes J from 3.3e-6 to
// 1.3e-6 but costs 1 % and 12 %, and J is well inside 1e-5 without it.
constexpr int kNearFold = 4;

// TwoSum of the U partials (acc[0..2]) into (sum, err), partials reset
__device__ __forceinline__ void fold_u(float (&acc)[kAccF32], float (&sum)[kAccF32], float (&err)[kAccF32])
{
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float x = acc[c];
        const float t = sum[c] + x;
        const float bp = t - sum[c];
        err[c] += (sum[c] - (t - bp)) + (x - bp);
        sum[c] = t;
        acc[c] = 0.f;
    }
}

// One LDS tile of n <= 64 NG sources against the lane's target (F32C).
// a = (s'_hi, far threshold), b = (Gamma', 1/sigma), l = s'_lo.
// Each wave first splits the tile with one ballot per 64 sources: a source
// is near when its distance to the wave's target box is below q_far sigma
// (then some lane may be near; the test is conservative).  The far and the
// near sources are then visited by iterating over the set bits of the two
// masks (scalar loop control), so the far loop body is straight-line.
template <int K, int NG = kTile / 64>
__device__ __forceinline__ void tile_f32c(int n, const float4* lpos, const float4* lgs, const float4* llo,
                                          float tx, float ty, float tz, float tlx, float tly, float tlz,
                                          const WaveBox& box, float (&acc)[kAccF32], float (&sum)[kAccF32],
                                          float (&err)[kAccF32])
{
    const int lane = threadIdx.x & 63;
    uint64_t near[NG], far[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int s = g * 64 + lane;
        bool nr = false, ok = s < n;
        if (ok) {
            const float4 a = lpos[s];
            const float ex = fmaxf(fmaxf(box.lo[0] - a.x, a.x - box.hi[0]), 0.f);
            const float ey = fmaxf(fmaxf(box.lo[1] - a.y, a.y - box.hi[1]), 0.f);
            const float ez = fmaxf(fmaxf(box.lo[2] - a.z, a.z - box.hi[2]), 0.f);
            // 1e-4 margin covers the FP32 rounding of the box test
            nr = fmaf(ez, ez, fmaf(ey, ey, ex * ex)) < a.w * 1.0001f;
        }
        near[g] = __ballot(nr);
        far[g] = __ballot(ok && !nr);
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        uint64_t fb = far[g];
        if (fb == ~0ull) {
            // the whole group is far (the common case in Morton order): a
            // straight-line loop with immediate LDS offsets, no scalar
            // bookkeeping per source
#pragma unroll 16
            for (int s = 0; s < 64; ++s) far_pair<K>(tx, ty, tz, lpos[g * 64 + s], lgs[g * 64 + s], acc);
        } else {
            while (fb) {
                const int s = g * 64 + __builtin_ctzll(fb);
                fb &= fb - 1;
                far_pair<K>(tx, ty, tz, lpos[s], lgs[s], acc);
            }
        }
    }
    // Near pairs carry the largest terms of a sum that may cancel to far
    // below them, so over 256-source tiles the U partials are folded into
    // the TwoSum pairs every kNearFold near pairs (the count is
    // wave-uniform).  One-group tiles (the wave-per-leaf list kernels) fold
    // every 64 sources anyway.
    int nn = 0;
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        uint64_t nb = near[g];
        while (nb) {
            const int s = g * 64 + __builtin_ctzll(nb);
            nb &= nb - 1;
            near_pair<K>(tx, ty, tz, tlx, tly, tlz, lpos[s], lgs[s], llo[s], acc);
            if constexpr (NG > 1) {
                if (++nn == kNearFold) {
                    fold_u(acc, sum, err);
                    nn = 0;
                }
            }
        }
    }
    (void)nn;
}

// hi/lo split of a position relative to the block origin
// Blocks are dealt round-robin to the 8 XCDs (block b runs on XCD b mod 8).
// This renumbering (a bijection of [0, G)) gives each XCD a contiguous run
// of blocks, so the blocks one XCD runs at once work on neighbouring targets
// and find their shared sources in that XCD's L2: on the 2^20 jet's
// 16-particle leaf lists E_str and zeta take a third less time, U/J
// (compute-bound) the same; the all-pairs zeta kernel 3 % less.
__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t G)
{
    const int64_t x = b & 7, q = G >> 3, r = G & 7;
    return x * q + min(x, r) + (b >> 3);
}

__device__ __forceinline__ void split_rel(double x, double o, float& hi, float& lo)
{
    const double d = x - o;
    hi = (float)d;
    lo = (float)(d - (double)hi);
}

// Adds the F32C accumulators (U', W', Jhv) of one target onto its record:
// U += U', J = -3 Jhv + eps . W' (column-major), Jhv[2][2] = -Jhv00 - Jhv11.
__device__ __forceinline__ void add_f32c_to_record(double* rec, const double (&a)[kAccF32])
{
    const double Wx = a[3], Wy = a[4], Wz = a[5];
    const double j8 = -(a[6] + a[10]);
    rec[RU + 0] += a[0];
    rec[RU + 1] += a[1];
    rec[RU + 2] += a[2];
    rec[RJ + 0] += -3.0 * a[6];
    rec[RJ + 1] += -3.0 * a[7] - Wz;
    rec[RJ + 2] += -3.0 * a[8] + Wy;
    rec[RJ + 3] += -3.0 * a[9] + Wz;
    rec[RJ + 4] += -3.0 * a[10];
    rec[RJ + 5] += -3.0 * a[11] - Wx;
    rec[RJ + 6] += -3.0 * a[12] - Wy;
    rec[RJ + 7] += -3.0 * a[13] + Wx;
    rec[RJ + 8] += -3.0 * j8;
}

#pragma clang fp contract(off)
// custom_erf64 (src/FLOWVPM_gpu_erf.jl:159-191): fdlibm s_erf.c branches and
// coefficients (the coefficient table is the published fdlibm one).
__device__ double erf64_ref(double x)
{
    const double xabs = fabs(x);
    const double sgn = x > 0 ? 1.0 : (x < 0 ? -1.0 : x);
    double val = sgn * 1.0;
    if (xabs < 0.84375) {
        const double z = x * x;
        const double r = 1.28379167095512558561e-01 + z * (-3.25042107247001499370e-01 + z * (-2.84817495755985104766e-02 + z * (-5.77027029648944159157e-03 + z * -2.37630166566501626084e-05)));
        const double s = 1.0 + z * (3.97917223959155352819e-01 + z * (6.50222499887672944485e-02 + z * (5.08130628187576562776e-03 + z * (1.32494738004321644526e-04 + z * -3.96022827877536812320e-06))));
        const double y = r / s;
        val = sgn * (xabs + xabs * y);
    } else if (xabs < 1.25) {
        const double s = xabs - 1.0;
        const double P = -2.36211856075265944077e-03 + s * (4.14856118683748331666e-01 + s * (-3.72207876035701323847e-01 + s * (3.18346619901161753674e-01 + s * (-1.10894694282396677476e-01 + s * (3.54783043256182359371e-02 + s * -2.16637559486879084300e-03)))));
        const double Q = 1.0 + s * (1.06420880400844228286e-01 + s * (5.40397917702171048937e-01 + s * (7.18286544141962662868e-02 + s * (1.26171219808761642112e-01 + s * (1.36370839120290507362e-02 + s * 1.19844998467991074170e-02)))));
        val = sgn * (8.45062911510467529297e-01 + P / Q);
    } else if (xabs < 2.857142857142857) {
        const double s = 1.0 / (x * x);
        const double R = -9.86494403484714822705e-03 + s * (-6.93858572707181764372e-01 + s * (-1.05586262253232909814e+01 + s * (-6.23753324503260060396e+01 + s * (-1.62396669462573470355e+02 + s * (-1.84605092906711035994e+02 + s * (-8.12874355063065934246e+01 + s * -9.81432934416914548592e+00))))));
        const double S = 1.0 + s * (1.96512716674392571292e+01 + s * (1.37657754143519042600e+02 + s * (4.34565877475229228821e+02 + s * (6.45387271733267880336e+02 + s * (4.29008140027567833386e+02 + s * (1.08635005541779435134e+02 + s * (6.57024977031928170135e+00 + s * -6.04244152148580987438e-02)))))));
        const double r = exp(-x * x - 0.5625 + R / S);
        val = sgn * (1.0 - r / xabs);
    } else if (xabs < 6.0) {
        const double s = 1.0 / (x * x);
        const double R = -9.86494292470009928597e-03 + s * (-7.99283237680523006574e-01 + s * (-1.77579549177547519889e+01 + s * (-1.60636384855821916062e+02 + s * (-6.37566443368389627722e+02 + s * (-1.02509513161107724954e+03 + s * -4.83519191608651397019e+02)))));
        const double S = 1.0 + s * (3.03380607434824582924e+01 + s * (3.25792512996573918826e+02 + s * (1.53672958608443695994e+03 + s * (3.19985821950859553908e+03 + s * (2.55305040643316442583e+03 + s * (4.74528541206955367215e+02 + s * -2.24409524465858183362e+01))))));
        const double r = exp(-x * x - 0.5625 + R / S);
        val = sgn * (1.0 - r / xabs);
    }
    return val;
}

// kernel.jl:45-84, statement by statement
template <int K>
__device__ __forceinline__ void g_dgdr_ref(double r, double& g, double& dg)
{
    if constexpr (K == VPM_KERNEL_SINGULAR) {
        g = 1.0; dg = 0.0;
    } else if constexpr (K == VPM_KERNEL_GAUSSIAN) {
        const double aux = exp(-r * r * r);
        g = 1 - aux; dg = 3 * r * r * aux;
    } else if constexpr (K == VPM_KERNEL_GAUSSIANERF) {
        const double aux = kConst2 * r * exp(-r * r / 2);
        g = erf64_ref(r / kSqr2) - aux; dg = r * aux;
    } else {
        double aux0 = r * r + 1;
        const double aux02 = aux0 * aux0;
        aux0 = aux02 * aux02 * aux0;
        aux0 = sqrt(aux0);
        g = r * r * r * (r * r + 2.5) / aux0;
        dg = 7.5 * r * r / (aux0 * (r * r + 1));
    }
}

// One source (a = (x, y, z, sigma), b = Gamma) acting on target t in FP64,
// statement by statement as fmm.jl:113-161.
template <int K>
__device__ __forceinline__ void pair_f64(double tx, double ty, double tz, const double4 a, const double4 b,
                                         double (&acc)[kAccF64])
{
    const double dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 != 0.0) {
        const double sigma = a.w;
        const double r = sqrt(r2);
        double g_sgm, dg_sgmdr;
        g_dgdr_ref<K>(r / sigma, g_sgm, dg_sgmdr);
        const double r3inv = 1.0 / (r2 * r);
        const double crss1 = -kConst4 * r3inv * (dy * b.z - dz * b.y);
        const double crss2 = -kConst4 * r3inv * (dz * b.x - dx * b.z);
        const double crss3 = -kConst4 * r3inv * (dx * b.y - dy * b.x);
        acc[0] += g_sgm * crss1;
        acc[1] += g_sgm * crss2;
        acc[2] += g_sgm * crss3;
        const double aux = dg_sgmdr / (sigma * r) - 3 * g_sgm / r2;
        const double aux2 = -kConst4 * g_sgm * r3inv;
        acc[3] += aux * crss1 * dx;
        acc[4] += aux * crss2 * dx - aux2 * b.z;
        acc[5] += aux * crss3 * dx + aux2 * b.y;
        acc[6] += aux * crss1 * dy + aux2 * b.z;
        acc[7] += aux * crss2 * dy;
        acc[8] += aux * crss3 * dy - aux2 * b.x;
        acc[9] += aux * crss1 * dz - aux2 * b.y;
        acc[10] += aux * crss2 * dz + aux2 * b.x;
        acc[11] += aux * crss3 * dz;
    }
}
#pragma clang fp contract(on)

// ------------------------------------------- FP64 pair with a far form
// The FP64 pair of the far-form kernel (VPM_UJ_F64_FAR): one rsqrt per pair
// instead of a sqrt and three divisions, and past a q where the kernel's g
// is 1 and q g' is 0 to FP64 rounding, the far form F' = r^-3, H'' = r^-5:
//  * Gaussian-erf, q >= 10: 1 - g = erfc(q/sqrt2) + sqrt(2/pi) q e^{-q^2/2}
//    < 2e-21 and q g' = sqrt(2/pi) q^3 e^{-q^2/2} < 2e-19;
//  * Gaussian, q^2 >= 13 (q >= 3.605): e^{-q^3} < 5e-21, 3 q^3 e^{-q^3} < 7e-19;
//  * Winckelmans (g -> 1 only as 1 - 1.9/q^4) and singular (r = 0 guard):
//    every pair takes the near form.
// Near pairs use the reference's g and dg/dq (g_dgdr_ref) at q = r / sigma,
// except Winckelmans, whose reduced form has no cancellation and no division.
template <int K> struct Far64 { static constexpr double q2 = 0.0; static constexpr bool cheap = false; };
template <> struct Far64<VPM_KERNEL_GAUSSIANERF> { static constexpr double q2 = 100.0; static constexpr bool cheap = true; };
template <> struct Far64<VPM_KERNEL_GAUSSIAN> { static constexpr double q2 = 13.0; static constexpr bool cheap = true; };

// far-field r^2 threshold of a source of core size sigma (FP64 kernel)
template <int K>
__device__ __forceinline__ double far_cut64(double sigma)
{
    return Far64<K>::cheap ? Far64<K>::q2 * sigma * sigma : DBL_MAX;
}

// 1/sqrt(x) for finite x > 0: the hardware estimate y and one third-order
// step y (1 + e/2 + 3 e^2/8), e = 1 - x y^2 (Householder's iteration for
// x^-1/2; ocml's rsqrt takes the same step and adds a guard for x = 0 and
// x = inf, which the pairs here never pass)
__device__ __forceinline__ double rsqrt_pos(double x)
{
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

// accum_pair in FP64 (the 14 sums of add_f32c_to_record)
__device__ __forceinline__ void accum_pair64(double dx, double dy, double dz, double Fp, double Hpp,
                                             const double4 b, double (&acc)[kAccF32])
{
    const double vx = fma(dy, b.z, -dz * b.y);
    const double vy = fma(dz, b.x, -dx * b.z);
    const double vz = fma(dx, b.y, -dy * b.x);
    acc[0] = fma(Fp, vx, acc[0]);
    acc[1] = fma(Fp, vy, acc[1]);
    acc[2] = fma(Fp, vz, acc[2]);
    acc[3] = fma(Fp, b.x, acc[3]);
    acc[4] = fma(Fp, b.y, acc[4]);
    acc[5] = fma(Fp, b.z, acc[5]);
    const double hx = Hpp * vx, hy = Hpp * vy, hz = Hpp * vz;
    acc[6] = fma(hx, dx, acc[6]);
    acc[7] = fma(hy, dx, acc[7]);
    acc[8] = fma(hz, dx, acc[8]);
    acc[9] = fma(hx, dy, acc[9]);
    acc[10] = fma(hy, dy, acc[10]);
    acc[11] = fma(hz, dy, acc[11]);
    acc[12] = fma(hx, dz, acc[12]);
    acc[13] = fma(hy, dz, acc[13]);
}

// a = (x, y, z, far threshold), b = (Gamma' = -Gamma / 4 pi, 1/sigma)
template <int K>
__device__ __forceinline__ void near_pair64(double tx, double ty, double tz, const double4 a, const double4 b,
                                            double (&acc)[kAccF32])
{
    const double dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
    const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
    if constexpr (K == VPM_KERNEL_WINCKELMANS) {
        // g/q^3 = (t + 5/2)/(t + 1)^{5/2} and (g - q g'/3)/q^5 =
        // (t + 7/2)/(t + 1)^{7/2}, t = q^2 (kernel.jl:69-74 reduced; no
        // cancellation at any q)
        if (r2 > 0.0) {   // coincident pairs contribute nothing (fmm.jl:118)
            const double s2 = b.w * b.w;
            const double t = r2 * s2;
            const double sr = rsqrt_pos(t + 1.0);
            const double w = sr * sr;
            const double w2s = w * w * sr;
            const double s3 = s2 * b.w;
            accum_pair64(dx, dy, dz, (t + 2.5) * w2s * s3, (t + 3.5) * w2s * w * s3 * s2, b, acc);
        }
        return;
    }
    if (r2 > 0.0) {   // coincident pairs contribute nothing (fmm.jl:118)
        const double ri = rsqrt_pos(r2);
        const double q = r2 * ri * b.w;
        double g, dg;
        g_dgdr_ref<K>(q, g, dg);
        const double ri2 = ri * ri;
        const double ri3 = ri2 * ri;
        accum_pair64(dx, dy, dz, g * ri3, fma((-1.0 / 3.0) * q, dg, g) * ri3 * ri2, b, acc);
    }
}

template <int K>
__device__ __forceinline__ void far_pair64(double tx, double ty, double tz, const double4 a, const double4 b,
                                           double (&acc)[kAccF32])
{
    if constexpr (Far64<K>::cheap) {
        const double dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
        const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
        const double ri = rsqrt_pos(r2);
        const double ri2 = ri * ri;
        const double Fp = ri2 * ri;
        accum_pair64(dx, dy, dz, Fp, Fp * ri2, b, acc);
    } else {
        near_pair64<K>(tx, ty, tz, a, b, acc);
    }
}

// Bounding box of the wave's targets in FP64 (wave-uniform)
struct WaveBox64 { double lo[3], hi[3]; };

__device__ __forceinline__ double uniform_d(double x)
{
    const uint64_t u = __builtin_bit_cast(uint64_t, x);
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)u);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(u >> 32));
    return __builtin_bit_cast(double, ((uint64_t)hi << 32) | lo);
}

__device__ __forceinline__ WaveBox64 wave_box64(double tx, double ty, double tz)
{
    WaveBox64 b;
    const double v[3] = {tx, ty, tz};
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        double mn = v[c], mx = v[c];
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            mn = fmin(mn, __shfl_xor(mn, off));
            mx = fmax(mx, __shfl_xor(mx, off));
        }
        b.lo[c] = uniform_d(mn);
        b.hi[c] = uniform_d(mx);
    }
    return b;
}

// One LDS tile of n <= kTile sources against the lane's target (FP64 far
// form), split per 64-source group into near and far sources against the
// wave's box as tile_f32c does; a far source is beyond q_far of every
// target of the wave.
template <int K>
__device__ __forceinline__ void tile_f64far(int n, const double4* lpos, const double4* lgs, double tx, double ty,
                                            double tz, const WaveBox64& box, double (&acc)[kAccF32])
{
    constexpr int NG = kTile / 64;
    const int lane = threadIdx.x & 63;
    uint64_t near[NG], far[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int s = g * 64 + lane;
        bool nr = false, ok = s < n;
        if (ok) {
            const double4 a = lpos[s];
            const double ex = fmax(fmax(box.lo[0] - a.x, a.x - box.hi[0]), 0.0);
            const double ey = fmax(fmax(box.lo[1] - a.y, a.y - box.hi[1]), 0.0);
            const double ez = fmax(fmax(box.lo[2] - a.z, a.z - box.hi[2]), 0.0);
            nr = fma(ez, ez, fma(ey, ey, ex * ex)) < a.w;
        }
        near[g] = __ballot(nr);
        far[g] = __ballot(ok && !nr);
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        uint64_t fb = far[g];
        if (fb == ~0ull) {
#pragma unroll 2
            for (int s = 0; s < 64; ++s) far_pair64<K>(tx, ty, tz, lpos[g * 64 + s], lgs[g * 64 + s], acc);
        } else {
            while (fb) {
                const int s = g * 64 + __builtin_ctzll(fb);
                fb &= fb - 1;
                far_pair64<K>(tx, ty, tz, lpos[s], lgs[s], acc);
            }
        }
    }
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        uint64_t nb = near[g];
        while (nb) {
            const int s = g * 64 + __builtin_ctzll(nb);
            nb &= nb - 1;
            near_pair64<K>(tx, ty, tz, lpos[s], lgs[s], acc);
        }
    }
}

// zeta(r/sigma)/sigma^3 of a source in FP32 as b exp2(a r^2) (Gaussian-erf),
// b exp2(a r^3) (Gaussian), b (1 + a r^2)^-3.5 (Winckelmans), with (a, b)
// from the source's sigma in FP64 (kernel.jl:51,60,69-74).
template <int K>
__host__ __device__ __forceinline__ void zeta_params(double sigma, double& a, double& b)
{
    const double si = 1.0 / sigma;
    const double s2 = si * si, s3 = s2 * si;
    const double l2e = 1.4426950408889634;
    if (K == VPM_KERNEL_GAUSSIANERF) { a = -0.5 * l2e * s2; b = kConst1 * s3; }
    else if (K == VPM_KERNEL_GAUSSIAN) { a = -l2e * s3; b = kConst3 * s3; }
    else { a = s2; b = kConst4 * 7.5 * s3; }
}

template <int K>
__device__ __forceinline__ float zeta_ab(float r2, float a, float b)
{
    if constexpr (K == VPM_KERNEL_GAUSSIANERF) {
        return b * ex2(a * r2);
    } else if constexpr (K == VPM_KERNEL_GAUSSIAN) {
        const float r = r2 > 0.f ? r2 * rsq(r2) : 0.f;
        return b * ex2(a * r2 * r);
    } else {
        const float s = rsq(fmaf(a, r2, 1.f));
        const float s2 = s * s;
        return b * (s2 * s2) * (s2 * s);
    }
}

// r^2 beyond which zeta is below FP32 resolution of the self term
// (q = 7 Gaussian-erf, q = 3 Gaussian); Winckelmans has no cut.
template <int K>
__host__ __device__ __forceinline__ float zeta_cut(double sigma)
{
    if (K == VPM_KERNEL_GAUSSIANERF) return (float)(49.0 * sigma * sigma);
    if (K == VPM_KERNEL_GAUSSIAN) return (float)(9.0 * sigma * sigma);
    return FLT_MAX;
}

// ---- source tiles beyond every cut ----
// The near-field-only kernels (zeta, E_str, the SFS differences) drop a
// source when its FP32 hi distance to the target (or to the wave's target
// box) is at least its cut.  A block skips a whole tile of kTile sorted
// sources, without loading it, when the FP32 gap between the block's target
// box and the tile's source box is already at least the tile's largest cut.
// Both boxes bound the FP32 hi coordinates the per-source tests use (x - o
// is rounded monotonically, and the double box extremes are attained), and
// every operation of the gap is monotone in those coordinates, so the gap
// never exceeds a tested distance: a skipped tile holds only sources the
// per-source test would drop, and results are unchanged bit for bit.

// FP64 bounding box and largest cut of each tile of kTile consecutive
// sources: tb[8 t + 0..2] = lo, [3..5] = hi, [6] = max cut[j * CS].
template <int CS>
__global__ __launch_bounds__(kBlock) void tile_box_kernel(const double* __restrict__ pos, int64_t ld,
                                                          const float* __restrict__ cut, int64_t np,
                                                          double* __restrict__ tb)
{
    static_assert(kTile == kBlock, "one thread per source of a tile");
    __shared__ double red[7][kBlock];
    const int tid = threadIdx.x;
    const int64_t j = (int64_t)blockIdx.x * kTile + tid;
    if (j < np) {
        for (int c = 0; c < 3; ++c) red[c][tid] = red[3 + c][tid] = pos[c * ld + j];
        red[6][tid] = (double)cut[j * CS];
    } else {
        for (int c = 0; c < 3; ++c) { red[c][tid] = DBL_MAX; red[3 + c][tid] = -DBL_MAX; }
        red[6][tid] = 0.0;
    }
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (tid < w) {
            for (int c = 0; c < 3; ++c) {
                red[c][tid] = fmin(red[c][tid], red[c][tid + w]);
                red[3 + c][tid] = fmax(red[3 + c][tid], red[3 + c][tid + w]);
            }
            red[6][tid] = fmax(red[6][tid], red[6][tid + w]);
        }
        __syncthreads();
    }
    if (tid < 7) tb[8 * (int64_t)blockIdx.x + tid] = red[tid][0];
}

// Union of the wave boxes of a block, in every thread.  wb is __shared__
// float[kBlock / 64][6]; contains a barrier (call from every thread).
__device__ __forceinline__ void block_box(const WaveBox& box, float (*wb)[6], float (&blo)[3], float (&bhi)[3])
{
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int c = 0; c < 3; ++c) { wb[wave][c] = box.lo[c]; wb[wave][3 + c] = box.hi[c]; }
    __syncthreads();
    for (int c = 0; c < 3; ++c) {
        blo[c] = wb[0][c];
        bhi[c] = wb[0][3 + c];
        for (int w = 1; w < kBlock / 64; ++w) {
            blo[c] = fminf(blo[c], wb[w][c]);
            bhi[c] = fmaxf(bhi[c], wb[w][3 + c]);
        }
    }
}

// Squared FP32 gap between the block's target box and tile box T, both
// relative to the block origin org (same evaluation order as the per-source
// distances: x, then y, then z by fma).
__device__ __forceinline__ float tile_gap2(const double* __restrict__ T, const double (&org)[3],
                                           const float (&blo)[3], const float (&bhi)[3])
{
    float g2 = 0.f;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float lo = (float)(T[c] - org[c]);
        const float hi = (float)(T[3 + c] - org[c]);
        const float g = fmaxf(fmaxf(blo[c] - hi, lo - bhi[c]), 0.f);
        g2 = c == 0 ? g * g : fmaf(g, g, g2);
    }
    return g2;
}

// The near tiles of [t0, t1) for a block, in ascending order: kBlock tiles
// are tested at once (one per thread, independent loads), the near ones are
// compacted into the __shared__ list, and next() walks the list.  A tile is
// near when tile_gap2 < its largest cut times margin (the margin of the
// kernel's per-source test).  All threads must call next() together: it
// holds barriers and returns the same tile (or -1 at the end) to each.
struct NearTiles {
    const double* tbox;
    int64_t c0, t1;
    const double* org;
    const float* blo;
    const float* bhi;
    float margin;
    int* list;      // __shared__ int[kBlock]
    int* wcnt;      // __shared__ int[kBlock / 64]
    int cnt = 0, k = 0;
    int64_t cbase = 0;

    __device__ NearTiles(const double* tb, int64_t first, int64_t end, const double (&o)[3], const float (&lo)[3],
                         const float (&hi)[3], float m, int* l, int* w)
        : tbox(tb), c0(first), t1(end), org(o), blo(lo), bhi(hi), margin(m), list(l), wcnt(w) {}

    __device__ int64_t next()
    {
        while (k >= cnt) {
            if (c0 >= t1) return -1;
            const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
            const int64_t t = c0 + tid;
            bool nr = false;
            if (t < t1) {
                const double* T = tbox + 8 * t;
                const double o3[3] = {org[0], org[1], org[2]};
                const float l3[3] = {blo[0], blo[1], blo[2]}, h3[3] = {bhi[0], bhi[1], bhi[2]};
                nr = tile_gap2(T, o3, l3, h3) < (float)T[6] * margin;
            }
            const uint64_t m = __ballot(nr);
            __syncthreads();   // readers of the previous list are done
            if (lane == 0) wcnt[wave] = __popcll(m);
            __syncthreads();
            int off = 0, tot = 0;
            for (int w = 0; w < kBlock / 64; ++w) {
                off += w < wave ? wcnt[w] : 0;
                tot += wcnt[w];
            }
            if (nr) list[off + __popcll(m & ((1ull << lane) - 1ull))] = tid;
            __syncthreads();
            cbase = c0;
            c0 += kBlock;
            cnt = tot;
            k = 0;
        }
        return cbase + list[k++];
    }
};

}  // namespace vpm
