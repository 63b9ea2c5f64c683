// estr.hip -- all-pairs subfilter-scale stretching E_str
// (Estr_direct!, src/FLOWVPM_subfilterscale_models.jl:16-92).
//
// SFS_i += sum_j zeta_sigma_j(r_ij) S_ij over non-static targets i and
// non-static sources j, with S_ij = (J_i - J_j)^T Gamma_j (transposed scheme)
// or (J_i - J_j) Gamma_j (classic) and zeta_sigma(r) = zeta(r/sigma)/sigma^3.
//
// F32C: write J'_k = J_k - J_ref with J_ref the J of the block's first
// (spatially sorted) target, subtracted in FP64.  Then
//     sum_j zeta_j (J'_i - J'_j)^T Gamma_j = J'_i^T A_i - B_i,
//     A_i = sum_j zeta_j Gamma_j,  B_i = sum_j zeta_j J'_j^T Gamma_j,
// so a pair costs one zeta and six FMAs, and every FP32 rounding is relative
// to the local variation of J (not to |J|).  The self pair, exactly zero in
// the reference, is left out: in the factored form it would enter A and B as
// two terms zeta(0) J'_i^T Gamma_i that cancel, which for the Gaussian kernel
// (zeta(0) = e^8 zeta(2)) can dwarf every other term.  Distances use the hi/lo split of
// positions relative to the block origin (zeta is steep, so r needs full
// FP32 relative accuracy).  Sources beyond q_far, where zeta is below FP32
// resolution of the near sum, are skipped per wave.  A and B use the same
// per-tile TwoSum compensation as the U/J kernel; J'_i^T A - B is formed in
// FP64.  The singular kernel (zeta nonzero only at r = 0) uses the F64 path.
//
// F64: the reference's pair expression (subfilterscale_models.jl:22-40) in
// storage order, contraction off -- the GPU oracle for E_str.
#include <cfloat>

#include "p2p_pair.h"

namespace vpm {

__device__ __forceinline__ float rsq_e(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ float ex2_e(float x) { return __builtin_amdgcn_exp2f(x); }
constexpr double kLog2eD = 1.4426950408889634;

// zeta_sigma(r) from per-source (a, b) (see pack_estr_kernel)
template <int K>
__device__ __forceinline__ float zeta_sig(float r2, float a, float b)
{
    if constexpr (K == VPM_KERNEL_GAUSSIANERF) {
        return b * ex2_e(a * r2);                    // const1 s^-3 exp(-q^2/2)
    } else if constexpr (K == VPM_KERNEL_GAUSSIAN) {
        const float r = r2 > 0.f ? r2 * rsq_e(r2) : 0.f;
        return b * ex2_e(a * r2 * r);                // const3 s^-3 exp(-q^3)
    } else {
        const float s = rsq_e(fmaf(a, r2, 1.f));     // (q^2 + 1)^-1/2
        const float s2 = s * s;
        return b * (s2 * s2) * (s2 * s);             // 15/(8 pi) s^-3 (q^2+1)^-3.5
    }
}

// Sources beyond q_far contribute below FP32 resolution of the near sum:
// Gaussian-erf zeta(7)/zeta(0) = 2.3e-11, Gaussian zeta(3)/zeta(0) = 1.9e-12.
// Winckelmans decays only like q^-7, so every source is evaluated.
template <int K>
__device__ __forceinline__ float estr_cut(double sigma)
{
    if (K == VPM_KERNEL_GAUSSIANERF) return (float)(49.0 * sigma * sigma);
    if (K == VPM_KERNEL_GAUSSIAN) return (float)(9.0 * sigma * sigma);
    return FLT_MAX;
}

// Source records for E_str in sorted order: positions and J in FP64 (the
// kernel subtracts the block origin and J_ref in FP64), (Gamma, a) and
// (b, cut) in FP32.  b = 0 for static sources, which the reference skips
// (subfilterscale_models.jl:66).
template <int K>
__global__ __launch_bounds__(kBlock) void pack_estr_kernel(const double* __restrict__ P, int64_t np,
                                                           const uint32_t* __restrict__ perm,
                                                           double* __restrict__ pos, double* __restrict__ J,
                                                           int64_t ld, float4* __restrict__ gz,
                                                           float2* __restrict__ bc)
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
    const double si = 1.0 / sg;
    const double s2 = si * si, s3 = s2 * si;
    double a, b;
    if (K == VPM_KERNEL_GAUSSIANERF) { a = -0.5 * kLog2eD * s2; b = kConst1 * s3; }
    else if (K == VPM_KERNEL_GAUSSIAN) { a = -kLog2eD * s3; b = kConst3 * s3; }
    else { a = s2; b = kConst4 * 7.5 * s3; }
    if (rec[RSTATIC] != 0) b = 0.0;
    gz[k] = make_float4((float)rec[RG], (float)rec[RG + 1], (float)rec[RG + 2], (float)a);
    bc[k] = make_float2((float)b, estr_cut<K>(sg));
}

__device__ __forceinline__ void split_rel_e(double x, double o, float& hi, float& lo)
{
    const double d = x - o;
    hi = (float)d;
    lo = (float)(d - (double)hi);
}

template <int K>
__global__ __launch_bounds__(kBlock) void estr_f32c_kernel(
    const double* __restrict__ pos, const double* __restrict__ J, int64_t ld,
    const float4* __restrict__ gz, const float2* __restrict__ bc, const double* __restrict__ tbox, int64_t np,
    int transposed, int64_t chunk, double* __restrict__ part, int64_t pld)
{
    __shared__ float4 lp[kTile];   // s'_hi x, y, z, a
    __shared__ float4 lg[kTile];   // Gamma x, y, z, b
    __shared__ float4 lq[kTile];   // J'^T Gamma (or J' Gamma), cut
    __shared__ float4 ll[kTile];   // s'_lo x, y, z
    __shared__ float wb[kBlock / 64][6];

    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const int64_t ib = (int64_t)blockIdx.x * kBlock;
    const double ox = pos[ib], oy = pos[ld + ib], oz = pos[2 * ld + ib];
    double Jr[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) Jr[c] = J[c * ld + ib];
    const int64_t it = i < np ? i : ib;
    float tx, ty, tz, tlx, tly, tlz;
    split_rel_e(pos[it], ox, tx, tlx);
    split_rel_e(pos[ld + it], oy, ty, tly);
    split_rel_e(pos[2 * ld + it], oz, tz, tlz);
    float blo[3], bhi[3];
    block_box(wave_box(tx, ty, tz), wb, blo, bhi);
    const double org[3] = {ox, oy, oz};

    const int64_t sb = (int64_t)blockIdx.y * chunk;
    const int64_t se = min(sb + chunk, np);
    float sum[6], err[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) { sum[c] = 0.f; err[c] = 0.f; }

    __shared__ int nlist[kBlock];
    __shared__ int ncnt[kBlock / 64];
    NearTiles near(tbox, sb / kTile, (se + kTile - 1) / kTile, org, blo, bhi, 1.0f, nlist, ncnt);
    for (int64_t tile; (tile = near.next()) >= 0;) {
        const int64_t base = tile * kTile;
        const int64_t j = base + tid;
        float4 p, g, q, l;
        if (j < se) {
            const float4 gg = gz[j];
            const float2 b = bc[j];
            split_rel_e(pos[j], ox, p.x, l.x);
            split_rel_e(pos[ld + j], oy, p.y, l.y);
            split_rel_e(pos[2 * ld + j], oz, p.z, l.z);
            p.w = gg.w;
            l.w = 0.f;
            g = make_float4(gg.x, gg.y, gg.z, b.x);
            double Jp[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) Jp[c] = J[c * ld + j] - Jr[c];
            const double G0 = gg.x, G1 = gg.y, G2 = gg.z;
            if (transposed) {   // S_a = sum_c J'[3a + c] Gamma_c
                q = make_float4((float)(Jp[0] * G0 + Jp[1] * G1 + Jp[2] * G2),
                                (float)(Jp[3] * G0 + Jp[4] * G1 + Jp[5] * G2),
                                (float)(Jp[6] * G0 + Jp[7] * G1 + Jp[8] * G2), b.y);
            } else {            // S_a = sum_c J'[3c + a] Gamma_c
                q = make_float4((float)(Jp[0] * G0 + Jp[3] * G1 + Jp[6] * G2),
                                (float)(Jp[1] * G0 + Jp[4] * G1 + Jp[7] * G2),
                                (float)(Jp[2] * G0 + Jp[5] * G1 + Jp[8] * G2), b.y);
            }
        } else {
            p = make_float4(1e15f, 1e15f, 1e15f, 0.f);
            g = make_float4(0.f, 0.f, 0.f, 0.f);
            q = make_float4(0.f, 0.f, 0.f, 0.f);
            l = make_float4(0.f, 0.f, 0.f, 0.f);
        }
        __syncthreads();
        lp[tid] = p;
        lg[tid] = g;
        lq[tid] = q;
        ll[tid] = l;
        __syncthreads();
        // fold into the compensated sums every 32 sources: E_str sums cancel
        // strongly, so the FP32 tile sums are kept short
        for (int s0 = 0; s0 < kTile; s0 += 32) {
            float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll 2
            for (int s = s0; s < s0 + 32; ++s) {
                const float4 a = lp[s];
                const float4 c = lq[s];
                float dx = tx - a.x, dy = ty - a.y, dz = tz - a.z;
                float r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                if (__any(r2 < c.w)) {
                    const float4 lo = ll[s];
                    const float4 b = lg[s];
                    dx += tlx - lo.x;
                    dy += tly - lo.y;
                    dz += tlz - lo.z;
                    r2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
                    // the self pair adds (J_i - J_i)^T Gamma_i = 0 exactly in
                    // the reference, but the factored form would add it as
                    // zeta(0) J'_i^T Gamma_i twice with opposite signs
                    const float z = base + s == i ? 0.f : zeta_sig<K>(r2, a.w, b.w);
                    acc[0] = fmaf(z, b.x, acc[0]);
                    acc[1] = fmaf(z, b.y, acc[1]);
                    acc[2] = fmaf(z, b.z, acc[2]);
                    acc[3] = fmaf(z, c.x, acc[3]);
                    acc[4] = fmaf(z, c.y, acc[4]);
                    acc[5] = fmaf(z, c.z, acc[5]);
                }
            }
#pragma unroll
            for (int c = 0; c < 6; ++c) {
                const float x = acc[c];
                const float t = sum[c] + x;
                const float bp = t - sum[c];
                err[c] += (sum[c] - (t - bp)) + (x - bp);
                sum[c] = t;
            }
        }
    }
    if (i < np) {
        double* out = part + (int64_t)blockIdx.y * 6 * pld + i;
#pragma unroll
        for (int c = 0; c < 6; ++c) out[c * pld] = (double)sum[c] + (double)err[c];
    }
}

__global__ __launch_bounds__(kBlock) void estr_finalize_kernel(const double* __restrict__ part, int64_t pld, int K,
                                                              const double* __restrict__ J, int64_t ld, int64_t np,
                                                              const uint32_t* __restrict__ perm, int transposed,
                                                              double* __restrict__ P)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    const int64_t p = perm ? (int64_t)perm[i] : i;
    double* rec = P + p * NF;
    if (rec[RSTATIC] != 0) return;                      // subfilterscale_models.jl:63
    double a[6] = {0, 0, 0, 0, 0, 0};
    for (int k = 0; k < K; ++k)
#pragma unroll
        for (int c = 0; c < 6; ++c) a[c] += part[((int64_t)k * 6 + c) * pld + i];
    const int64_t ib = (i / kBlock) * kBlock;
    double Jp[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) Jp[c] = J[c * ld + i] - J[c * ld + ib];
    double S[3];
    if (transposed) {
        for (int q = 0; q < 3; ++q) S[q] = Jp[3 * q] * a[0] + Jp[3 * q + 1] * a[1] + Jp[3 * q + 2] * a[2] - a[3 + q];
    } else {
        for (int q = 0; q < 3; ++q) S[q] = Jp[q] * a[0] + Jp[3 + q] * a[1] + Jp[6 + q] * a[2] - a[3 + q];
    }
    rec[RSFS + 0] += S[0];
    rec[RSFS + 1] += S[1];
    rec[RSFS + 2] += S[2];
}

// ---------------------------------------------------------------- F64
#pragma clang fp contract(off)
template <int K>
__device__ __forceinline__ double zeta_ref(double r)
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

// One block per 256 targets (storage order); sources in storage order, whole
// range per block (E_str is not the headline path; this is the oracle).
template <int K>
__global__ __launch_bounds__(kBlock) void estr_f64_kernel(double* __restrict__ P, int64_t np, int transposed)
{
    __shared__ double lx[kTile][4];
    __shared__ double lg[kTile][4];
    __shared__ double lj[kTile][9];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const int64_t it = i < np ? i : 0;
    const double* tr = P + it * NF;
    const double tx = tr[RX], ty = tr[RX + 1], tz = tr[RX + 2];
    double JT[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) JT[c] = tr[RJ + c];
    const bool tstatic = tr[RSTATIC] != 0;
    double s1 = 0, s2 = 0, s3 = 0;
    for (int64_t base = 0; base < np; base += kTile) {
        const int n = (int)min((int64_t)kTile, np - base);
        __syncthreads();
        if (tid < n) {
            const double* sr = P + (base + tid) * NF;
            lx[tid][0] = sr[RX]; lx[tid][1] = sr[RX + 1]; lx[tid][2] = sr[RX + 2]; lx[tid][3] = sr[RSIG];
            lg[tid][0] = sr[RG]; lg[tid][1] = sr[RG + 1]; lg[tid][2] = sr[RG + 2]; lg[tid][3] = sr[RSTATIC];
#pragma unroll
            for (int c = 0; c < 9; ++c) lj[tid][c] = sr[RJ + c];
        }
        __syncthreads();
        for (int s = 0; s < n; ++s) {
            if (lg[s][3] != 0) continue;                    // static source
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
            const double zeta_sgm = zeta_ref<K>(r * sigma_inv) * sigma_inv * sigma_inv * sigma_inv;
            s1 += zeta_sgm * S1;
            s2 += zeta_sgm * S2;
            s3 += zeta_sgm * S3;
        }
    }
    if (i < np && !tstatic) {
        double* rec = P + i * NF;
        rec[RSFS] += s1;
        rec[RSFS + 1] += s2;
        rec[RSFS + 2] += s3;
    }
}
#pragma clang fp contract(on)

static inline unsigned nblk(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// ------------------------------------------------------ F64 with a far cut
// E_str in FP64 for VPM_UJ_F64_FAR: the reference's pair expression
// (subfilterscale_models.jl:22-40) with zeta_sigma from r^2 (no sqrt or
// division for Gaussian-erf and Winckelmans), sources in the U/J kernels'
// spatial order, and only the source tiles and the pairs within q_cut:
// Gaussian-erf q_cut = 10 (zeta(10)/zeta(0) = e^-50 = 2e-22), Gaussian
// q_cut^2 = 14 (e^-52.4 = 2e-23), below FP64 resolution of the near terms.
// Winckelmans (zeta ~ q^-7) has no cut; the singular kernel's zeta is
// nonzero at r = 0 only, so only coincident pairs are evaluated.
template <int K>
__device__ __forceinline__ double estr_q2cut64()
{
    return K == VPM_KERNEL_GAUSSIANERF ? 100.0 : K == VPM_KERNEL_GAUSSIAN ? 14.0 : 0.0;
}

// r^2 cut of a source from its 1/sigma (-1: a static source, never paired)
template <int K>
__device__ __forceinline__ double estr_cut64(double si)
{
    if (!(si > 0.0)) return -1.0;
    if constexpr (K == VPM_KERNEL_WINCKELMANS) return DBL_MAX;
    else if constexpr (K == VPM_KERNEL_SINGULAR) return DBL_MIN;
    else return estr_q2cut64<K>() / (si * si);
}

// zeta(r/sigma)/sigma^3 (kernel.jl:51,60,69-74,78)
template <int K>
__device__ __forceinline__ double zeta64(double r2, double si)
{
    const double s3 = si * si * si;
    if constexpr (K == VPM_KERNEL_SINGULAR) {
        return r2 == 0.0 ? s3 : 0.0;
    } else if constexpr (K == VPM_KERNEL_GAUSSIAN) {
        const double q = (r2 > 0.0 ? r2 * rsqrt_pos(r2) : 0.0) * si;
        return kConst3 * exp(-q * q * q) * s3;
    } else if constexpr (K == VPM_KERNEL_GAUSSIANERF) {
        return kConst1 * exp(-0.5 * r2 * si * si) * s3;
    } else {
        const double w = rsqrt_pos(fma(r2, si * si, 1.0));   // (q^2 + 1)^-1/2
        const double w2 = w * w;
        return kConst4 * 7.5 * (w2 * w2 * w2 * w) * s3;
    }
}

// Sources of the far-cut kernel in the order of perm: positions and J in
// FP64 (SoA, stride ld), (Gamma, 1/sigma) with 1/sigma = 0 for static
// sources, and the r^2 cut in FP32 (rounded up) for the tile boxes.
template <int K>
__global__ __launch_bounds__(kBlock) void pack_estr64_kernel(const double* __restrict__ P, int64_t np,
                                                             const uint32_t* __restrict__ perm,
                                                             double* __restrict__ pos, double* __restrict__ J,
                                                             int64_t ld, double4* __restrict__ g4,
                                                             float* __restrict__ cutf)
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
    const double si = rec[RSTATIC] != 0 ? 0.0 : 1.0 / rec[RSIG];
    g4[k] = make_double4(rec[RG], rec[RG + 1], rec[RG + 2], si);
    const double c = estr_cut64<K>(si);
    cutf[k] = c < 0.0 ? 0.f : c >= (double)FLT_MAX ? FLT_MAX : (float)c * 1.0001f;
}

// One block of 256 targets (sorted) over the near source tiles (NearTiles on
// the tile boxes); in a tile, a source is evaluated when some lane of the
// wave is within its cut, and each lane adds the pair only within the cut.
template <int K>
__global__ __launch_bounds__(kBlock) void estr_f64far_kernel(const double* __restrict__ pos,
                                                             const double* __restrict__ J, int64_t ld,
                                                             const double4* __restrict__ g4,
                                                             const double* __restrict__ tbox,
                                                             const uint32_t* __restrict__ perm, int64_t np,
                                                             int transposed, double* __restrict__ P)
{
    __shared__ double4 lp[kTile];     // x, y, z, r^2 cut (-1: static source)
    __shared__ double4 lg[kTile];     // Gamma, 1/sigma
    __shared__ double lj[9][kTile];   // J
    __shared__ float wb[kBlock / 64][6];
    __shared__ int nlist[kBlock];
    __shared__ int ncnt[kBlock / 64];

    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const int64_t ib = (int64_t)blockIdx.x * kBlock;
    const int64_t it = i < np ? i : ib;
    const double tx = pos[it], ty = pos[ld + it], tz = pos[2 * ld + it];
    double JT[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) JT[c] = J[c * ld + it];
    const double org[3] = {pos[ib], pos[ld + ib], pos[2 * ld + ib]};
    float blo[3], bhi[3];
    block_box(wave_box((float)(tx - org[0]), (float)(ty - org[1]), (float)(tz - org[2])), wb, blo, bhi);

    double s1 = 0.0, s2 = 0.0, s3 = 0.0;
    NearTiles near(tbox, 0, (np + kTile - 1) / kTile, org, blo, bhi, 1.0001f, nlist, ncnt);
    for (int64_t tile; (tile = near.next()) >= 0;) {
        const int64_t base = tile * kTile;
        const int64_t j = base + tid;
        __syncthreads();   // the previous tile's readers are done
        if (j < np) {
            const double4 g = g4[j];
            lp[tid] = make_double4(pos[j], pos[ld + j], pos[2 * ld + j], estr_cut64<K>(g.w));
            lg[tid] = g;
#pragma unroll
            for (int c = 0; c < 9; ++c) lj[c][tid] = J[c * ld + j];
        }
        __syncthreads();
        const int n = (int)min((int64_t)kTile, np - base);
        for (int s = 0; s < n; ++s) {
            const double4 a = lp[s];
            const double dx = a.x - tx, dy = a.y - ty, dz = a.z - tz;
            const double r2 = fma(dz, dz, fma(dy, dy, dx * dx));
            const bool in = r2 < a.w;
            if (__any(in)) {
                const double4 b = lg[s];
                double S1, S2, S3;
                if (transposed) {
                    S1 = (JT[0] - lj[0][s]) * b.x + (JT[1] - lj[1][s]) * b.y + (JT[2] - lj[2][s]) * b.z;
                    S2 = (JT[3] - lj[3][s]) * b.x + (JT[4] - lj[4][s]) * b.y + (JT[5] - lj[5][s]) * b.z;
                    S3 = (JT[6] - lj[6][s]) * b.x + (JT[7] - lj[7][s]) * b.y + (JT[8] - lj[8][s]) * b.z;
                } else {
                    S1 = (JT[0] - lj[0][s]) * b.x + (JT[3] - lj[3][s]) * b.y + (JT[6] - lj[6][s]) * b.z;
                    S2 = (JT[1] - lj[1][s]) * b.x + (JT[4] - lj[4][s]) * b.y + (JT[7] - lj[7][s]) * b.z;
                    S3 = (JT[2] - lj[2][s]) * b.x + (JT[5] - lj[5][s]) * b.y + (JT[8] - lj[8][s]) * b.z;
                }
                const double z = in ? zeta64<K>(r2, b.w) : 0.0;
                s1 = fma(z, S1, s1);
                s2 = fma(z, S2, s2);
                s3 = fma(z, S3, s3);
            }
        }
    }
    if (i < np) {
        double* rec = P + (perm ? (int64_t)perm[i] : i) * NF;
        if (rec[RSTATIC] == 0) {
            rec[RSFS] += s1;
            rec[RSFS + 1] += s2;
            rec[RSFS + 2] += s3;
        }
    }
}

int launch_estr_impl(vpm_field f, int kernel, int transposed, int precision, const uint32_t* perm)
{
    const int64_t np = f->np;
    if (np == 0) return VPM_OK;
    int rc;
    // The singular kernel's zeta is nonzero only at r = 0 (coincident pairs),
    // where the reference's (J_t - J_s) difference is exact: use FP64.
    if (precision == VPM_PREC_F64 || kernel == VPM_KERNEL_SINGULAR) {
        timing_begin(f);
        switch (kernel) {
        case VPM_KERNEL_SINGULAR: hipLaunchKernelGGL(estr_f64_kernel<VPM_KERNEL_SINGULAR>, dim3(nblk(np)), dim3(kBlock), 0, f->stream, f->P, np, transposed); break;
        case VPM_KERNEL_GAUSSIAN: hipLaunchKernelGGL(estr_f64_kernel<VPM_KERNEL_GAUSSIAN>, dim3(nblk(np)), dim3(kBlock), 0, f->stream, f->P, np, transposed); break;
        case VPM_KERNEL_GAUSSIANERF: hipLaunchKernelGGL(estr_f64_kernel<VPM_KERNEL_GAUSSIANERF>, dim3(nblk(np)), dim3(kBlock), 0, f->stream, f->P, np, transposed); break;
        case VPM_KERNEL_WINCKELMANS: hipLaunchKernelGGL(estr_f64_kernel<VPM_KERNEL_WINCKELMANS>, dim3(nblk(np)), dim3(kBlock), 0, f->stream, f->P, np, transposed); break;
        default: return fail(VPM_EINVAL, "unknown kernel id");
        }
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return hip_fail(e, "estr_f64");
        timing_end(f, 2);
        return VPM_OK;
    }
    const int64_t ld = f->cap;
    // [3][ld] positions, [9][ld] J (double), [ld] float4, [ld] float2, tile boxes
    const int64_t ntile = (np + kTile - 1) / kTile;
    if ((rc = ensure(f->adj, sizeof(double) * 12 * (size_t)ld + sizeof(float4) * (size_t)ld +
                                 sizeof(float2) * (size_t)ld + 64 * (size_t)ntile)))
        return rc;
    double* pos = (double*)f->adj.p;
    double* J = pos + 3 * ld;
    float4* gz = (float4*)(J + 9 * ld);
    float2* bc = (float2*)(gz + ld);
    double* tbox = (double*)(bc + ld);
    switch (kernel) {
#define VPM_PE(KID) case KID: hipLaunchKernelGGL(pack_estr_kernel<KID>, dim3(nblk(np)), dim3(kBlock), 0, f->stream, f->P, np, perm, pos, J, ld, gz, bc); break;
        VPM_PE(VPM_KERNEL_GAUSSIAN) VPM_PE(VPM_KERNEL_GAUSSIANERF) VPM_PE(VPM_KERNEL_WINCKELMANS)
#undef VPM_PE
    default: return fail(VPM_EINVAL, "unknown kernel id");
    }
    hipLaunchKernelGGL(tile_box_kernel<2>, dim3((unsigned)ntile), dim3(kBlock), 0, f->stream, (const double*)pos, ld,
                       (const float*)bc + 1, np, tbox);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "pack_estr");
    // source splits, chunks of whole tiles: enough blocks to fill the chip,
    // and up to 32 chunks of at least 16 tiles to spread the near-field work
    // of blocks whose targets straddle a jump of the sort order (as for zeta)
    const int64_t nb = (np + kBlock - 1) / kBlock;
    const int64_t tiles = (np + kTile - 1) / kTile;
    const int64_t k_fill = std::min<int64_t>((2560 + nb - 1) / nb, tiles);
    const int64_t k_spread = std::min<int64_t>(32, tiles / 16);
    int64_t K = std::max<int64_t>(1, std::max(k_fill, k_spread));
    const int64_t chunk = ((tiles + K - 1) / K) * kTile;
    K = (np + chunk - 1) / chunk;
    if ((rc = ensure(f->part, sizeof(double) * 6 * (size_t)K * (size_t)np))) return rc;
    double* part = (double*)f->part.p;
    timing_begin(f);
    switch (kernel) {
#define VPM_ES(KID) case KID: hipLaunchKernelGGL(estr_f32c_kernel<KID>, dim3(nblk(np), (unsigned)K), dim3(kBlock), 0, f->stream, pos, J, ld, gz, bc, (const double*)tbox, np, transposed, chunk, part, np); break;
        VPM_ES(VPM_KERNEL_GAUSSIAN) VPM_ES(VPM_KERNEL_GAUSSIANERF) VPM_ES(VPM_KERNEL_WINCKELMANS)
#undef VPM_ES
    default: return fail(VPM_EINVAL, "unknown kernel id");
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "estr_f32c");
    timing_end(f, 2);
    hipLaunchKernelGGL(estr_finalize_kernel, dim3(nblk(np)), dim3(kBlock), 0, f->stream, part, np, (int)K, J, ld, np,
                       perm, transposed, f->P);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "estr_finalize");
    return VPM_OK;
}

int launch_estr_f64far(vpm_field f, int kernel, int transposed)
{
    const int64_t np = f->np;
    if (np == 0) return VPM_OK;
    int rc;
    const uint32_t* perm = nullptr;
    if ((rc = sort_field(f, perm))) return rc;
    const int64_t ld = f->cap;
    const int64_t ntile = (np + kTile - 1) / kTile;
    auto up = [](size_t b) { return (b + 255) & ~(size_t)255; };
    const size_t o_j = up(sizeof(double) * 3 * (size_t)ld), o_g = o_j + up(sizeof(double) * 9 * (size_t)ld);
    const size_t o_c = o_g + up(sizeof(double4) * (size_t)ld), o_t = o_c + up(sizeof(float) * (size_t)ld);
    if ((rc = ensure(f->adj, o_t + sizeof(double) * 8 * (size_t)ntile))) return rc;
    char* base = (char*)f->adj.p;
    double* pos = (double*)base;
    double* J = (double*)(base + o_j);
    double4* g4 = (double4*)(base + o_g);
    float* cutf = (float*)(base + o_c);
    double* tbox = (double*)(base + o_t);
    switch (kernel) {
#define VPM_PE64(KID) case KID: hipLaunchKernelGGL(pack_estr64_kernel<KID>, dim3(nblk(np)), dim3(kBlock), 0, f->stream, f->P, np, perm, pos, J, ld, g4, cutf); break;
        VPM_PE64(VPM_KERNEL_SINGULAR) VPM_PE64(VPM_KERNEL_GAUSSIAN) VPM_PE64(VPM_KERNEL_GAUSSIANERF)
        VPM_PE64(VPM_KERNEL_WINCKELMANS)
#undef VPM_PE64
    default: return fail(VPM_EINVAL, "unknown kernel id");
    }
    hipLaunchKernelGGL(tile_box_kernel<1>, dim3((unsigned)ntile), dim3(kBlock), 0, f->stream, (const double*)pos, ld,
                       (const float*)cutf, np, tbox);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "pack_estr64");
    timing_begin(f);
    switch (kernel) {
#define VPM_EF64(KID) case KID: hipLaunchKernelGGL(estr_f64far_kernel<KID>, dim3(nblk(np)), dim3(kBlock), 0, f->stream, (const double*)pos, (const double*)J, ld, (const double4*)g4, (const double*)tbox, perm, np, transposed, f->P); break;
        VPM_EF64(VPM_KERNEL_SINGULAR) VPM_EF64(VPM_KERNEL_GAUSSIAN) VPM_EF64(VPM_KERNEL_GAUSSIANERF)
        VPM_EF64(VPM_KERNEL_WINCKELMANS)
#undef VPM_EF64
    default: return fail(VPM_EINVAL, "unknown kernel id");
    }
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "estr_f64far");
    timing_end(f, 2);
    return VPM_OK;
}

}  // namespace vpm
