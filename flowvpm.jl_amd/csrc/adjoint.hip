// adjoint.hip -- hand-written adjoint of the all-pairs U/J evaluation
// (replaces the ChainRules pullback UJS_pullback, src/FLOWVPM_rrules.jl:179-330).
//
// For L = sum_i Ubar_i . U_i + sum_i Jbar_i : J_i the contribution of one
// pair (target i, source j; d = x_i - x_j, v = d x Gamma_j) is
//     l = F (A + C) + H B,
//     A = Ubar_i . v,  B = v . (Jbar_i d),  C = tau_i . Gamma_j,
//     tau = (Jbar_23 - Jbar_32, Jbar_31 - Jbar_13, Jbar_12 - Jbar_21),
// with F = -c g/r^3 and H = F'(r)/r the two coefficients of fmm.jl:125-158
// (U += F v, J[a][b] += H v_a d_b + F eps_abc Gamma_c).  With K = H'(r)/r,
//     dl/dd      = (H (A + C) + K B) d + F Gamma x Ubar + H (Gamma x (Jbar d) + Jbar^T v)
//     dl/dGamma  = F (Ubar x d + tau) + H (Jbar d) x d
//     dl/dsigma  = F_s (A + C) + H_s B,
// dl/dx_i = +dl/dd, dl/dx_j = -dl/dd.  No N^2 intermediate is stored: two
// gather passes recompute the pair terms, one with lanes on targets (dL/dx as
// target) and one with lanes on sources (dL/dx as source, dL/dGamma,
// dL/dsigma), each streaming the other side through LDS.
//
// The kernels below run the pair arithmetic in FP64 with the closed-form
// derivatives of g (VPM_PREC_F64, the FP64 reference of the adjoint); the
// fused FP32 kernel of VPM_PREC_F32C is in adjoint_f32c.hip.
#include "vpm_internal.h"

namespace vpm {

struct AdjCoef { double F, H, K, Fs, Hs; };

template <int KID>
__device__ __forceinline__ void g_derivs(double q, double& g, double& dg, double& d2g)
{
    if constexpr (KID == VPM_KERNEL_SINGULAR) {
        g = 1.0; dg = 0.0; d2g = 0.0;
    } else if constexpr (KID == VPM_KERNEL_GAUSSIAN) {
        const double E = exp(-q * q * q);
        g = 1.0 - E; dg = 3.0 * q * q * E; d2g = (6.0 * q - 9.0 * q * q * q * q) * E;
    } else if constexpr (KID == VPM_KERNEL_GAUSSIANERF) {
        const double E = exp(-0.5 * q * q);
        g = erf(q / kSqr2) - kConst2 * q * E;
        dg = kConst2 * q * q * E;
        d2g = kConst2 * (2.0 * q - q * q * q) * E;
    } else {
        const double t = q * q + 1.0;
        const double s = rsqrt(t);           // t^-1/2
        const double t35 = s * s * s * s * s * s * s;
        g = q * q * q * (q * q + 2.5) * t * t35;
        dg = 7.5 * q * q * t35;
        d2g = 7.5 * q * (2.0 - 5.0 * q * q) * t35 / t;
    }
}

template <int KID>
__device__ __forceinline__ AdjCoef adj_coef(double r2, double sigma)
{
    const double r = sqrt(r2);
    const double q = r / sigma;
    double g, dg, d2g;
    g_derivs<KID>(q, g, dg, d2g);
    const double ri2 = 1.0 / r2, ri3 = ri2 / r, ri5 = ri3 * ri2, ri7 = ri5 * ri2;
    AdjCoef k;
    k.F = -kConst4 * g * ri3;
    k.H = -kConst4 * (q * dg - 3.0 * g) * ri5;
    k.K = -kConst4 * (q * q * d2g - 7.0 * q * dg + 15.0 * g) * ri7;
    k.Fs = kConst4 * dg * q / sigma * ri3;
    k.Hs = kConst4 * q * (q * d2g - 2.0 * dg) / sigma * ri5;
    return k;
}

__device__ __forceinline__ void cross(const double a[3], const double b[3], double o[3])
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

// pair adjoint: target data (ub, Jb: Jbar flat, f = 3b + a is J[a][b]), source
// (G, sigma), d = x_i - x_j.  Outputs gd (dl/dd), gG, gs.
template <int KID>
__device__ __forceinline__ void pair_adj(const double d[3], const double G[3], double sigma, const double ub[3],
                                         const double* Jb, double gd[3], double gG[3], double& gs)
{
    const double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    if (r2 == 0.0) {
        gd[0] = gd[1] = gd[2] = gG[0] = gG[1] = gG[2] = 0.0;
        gs = 0.0;
        return;
    }
    const AdjCoef k = adj_coef<KID>(r2, sigma);
    double v[3];
    cross(d, G, v);
    double w[3], JTv[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        w[a] = Jb[a] * d[0] + Jb[3 + a] * d[1] + Jb[6 + a] * d[2];
        JTv[a] = Jb[3 * a] * v[0] + Jb[3 * a + 1] * v[1] + Jb[3 * a + 2] * v[2];
    }
    const double tau[3] = {Jb[7] - Jb[5], Jb[2] - Jb[6], Jb[3] - Jb[1]};
    const double A = ub[0] * v[0] + ub[1] * v[1] + ub[2] * v[2];
    const double B = v[0] * w[0] + v[1] * w[1] + v[2] * w[2];
    const double C = tau[0] * G[0] + tau[1] * G[1] + tau[2] * G[2];
    double Gxu[3], Gxw[3], uxd[3], wxd[3];
    cross(G, ub, Gxu);
    cross(G, w, Gxw);
    cross(ub, d, uxd);
    cross(w, d, wxd);
    const double s = k.H * (A + C) + k.K * B;
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        gd[a] = s * d[a] + k.F * Gxu[a] + k.H * (Gxw[a] + JTv[a]);
        gG[a] = k.F * (uxd[a] + tau[a]) + k.H * wxd[a];
    }
    gs = k.Fs * (A + C) + k.Hs * B;
}

// Target pass: lanes own targets i; sources stream through LDS.
template <int KID>
__global__ __launch_bounds__(kBlock) void adj_target_kernel(const double* __restrict__ P, int64_t np,
                                                            const double* __restrict__ Ubar,
                                                            const double* __restrict__ Jbar,
                                                            double* __restrict__ Xbar)
{
    __shared__ double lx[kTile][4];
    __shared__ double lg[kTile][4];
    const int tid = threadIdx.x;
    const int64_t i = (int64_t)blockIdx.x * kBlock + tid;
    const int64_t it = i < np ? i : 0;
    const double* tr = P + it * NF;
    const double tx = tr[RX], ty = tr[RX + 1], tz = tr[RX + 2];
    double ub[3], Jb[9];
    for (int c = 0; c < 3; ++c) ub[c] = Ubar[3 * it + c];
    for (int c = 0; c < 9; ++c) Jb[c] = Jbar[9 * it + c];
    double ax[3] = {0.0, 0.0, 0.0};
    for (int64_t base = 0; base < np; base += kTile) {
        const int n = (int)min((int64_t)kTile, np - base);
        __syncthreads();
        if (tid < n) {
            const double* sr = P + (base + tid) * NF;
            lx[tid][0] = sr[RX]; lx[tid][1] = sr[RX + 1]; lx[tid][2] = sr[RX + 2]; lx[tid][3] = sr[RSIG];
            lg[tid][0] = sr[RG]; lg[tid][1] = sr[RG + 1]; lg[tid][2] = sr[RG + 2];
        }
        __syncthreads();
        for (int s = 0; s < n; ++s) {
            const double d[3] = {tx - lx[s][0], ty - lx[s][1], tz - lx[s][2]};
            const double G[3] = {lg[s][0], lg[s][1], lg[s][2]};
            double gd[3], gG[3], gs;
            pair_adj<KID>(d, G, lx[s][3], ub, Jb, gd, gG, gs);
            ax[0] += gd[0]; ax[1] += gd[1]; ax[2] += gd[2];
        }
    }
    if (i < np) {
        Xbar[3 * i] = ax[0]; Xbar[3 * i + 1] = ax[1]; Xbar[3 * i + 2] = ax[2];
    }
}

// Source pass: lanes own sources j; targets (position, Ubar, Jbar) stream through LDS.
template <int KID>
__global__ __launch_bounds__(kBlock) void adj_source_kernel(const double* __restrict__ P, int64_t np,
                                                            const double* __restrict__ Ubar,
                                                            const double* __restrict__ Jbar,
                                                            double* __restrict__ Xbar, double* __restrict__ Gbar,
                                                            double* __restrict__ Sbar)
{
    __shared__ double lt[kTile][3];
    __shared__ double lu[kTile][3];
    __shared__ double lj[kTile][9];
    const int tid = threadIdx.x;
    const int64_t j = (int64_t)blockIdx.x * kBlock + tid;
    const int64_t jt = j < np ? j : 0;
    const double* sr = P + jt * NF;
    const double sx = sr[RX], sy = sr[RX + 1], sz = sr[RX + 2], sg = sr[RSIG];
    const double G[3] = {sr[RG], sr[RG + 1], sr[RG + 2]};
    double ax[3] = {0.0, 0.0, 0.0}, aG[3] = {0.0, 0.0, 0.0}, as = 0.0;
    for (int64_t base = 0; base < np; base += kTile) {
        const int n = (int)min((int64_t)kTile, np - base);
        __syncthreads();
        if (tid < n) {
            const int64_t t = base + tid;
            const double* tr = P + t * NF;
            lt[tid][0] = tr[RX]; lt[tid][1] = tr[RX + 1]; lt[tid][2] = tr[RX + 2];
            for (int c = 0; c < 3; ++c) lu[tid][c] = Ubar[3 * t + c];
            for (int c = 0; c < 9; ++c) lj[tid][c] = Jbar[9 * t + c];
        }
        __syncthreads();
        for (int s = 0; s < n; ++s) {
            const double d[3] = {lt[s][0] - sx, lt[s][1] - sy, lt[s][2] - sz};
            double gd[3], gG[3], gs;
            pair_adj<KID>(d, G, sg, lu[s], lj[s], gd, gG, gs);
            for (int a = 0; a < 3; ++a) { ax[a] -= gd[a]; aG[a] += gG[a]; }
            as += gs;
        }
    }
    if (j < np) {
        for (int a = 0; a < 3; ++a) { Xbar[3 * j + a] += ax[a]; Gbar[3 * j + a] = aG[a]; }
        Sbar[j] = as;
    }
}

int launch_uj_adjoint_f32c(vpm_field f, int kernel, const double* dU, const double* dJ, double* dX, double* dG,
                           double* dS, float4* tcot, bool jbar);

// Device staging of the U/J adjoint: Ubar 3n, Jbar 9n, Xbar 3n, Gbar 3n,
// Sbar n (doubles), then the sorted FP32 cotangents of the F32C path (4n
// float4, 16-byte aligned).
int adj_buffers(vpm_field f, AdjBufs& b)
{
    const size_t n = (size_t)f->np;
    const size_t cot_off = (19 * n + 1) & ~(size_t)1;
    if (int rc = ensure(f->adj, sizeof(double) * (cot_off + 8 * n + 2))) return rc;
    b.U = (double*)f->adj.p;
    b.J = b.U + 3 * n;
    b.X = b.J + 9 * n;
    b.G = b.X + 3 * n;
    b.S = b.G + 3 * n;
    b.cot = (float4*)(b.U + cot_off);
    return VPM_OK;
}

// U/J adjoint of the field's current particles with device-resident
// cotangents b.U, b.J; gradients into b.X, b.G, b.S (particle order).
int launch_uj_adjoint_dev(vpm_field f, int kernel, int precision, const AdjBufs& b)
{
    const int64_t np = f->np;
    if (np == 0) return VPM_OK;
    if (precision != VPM_PREC_F64) return launch_uj_adjoint_f32c(f, kernel, b.U, b.J, b.X, b.G, b.S, b.cot, b.jbar);
    const dim3 grid((unsigned)((np + kBlock - 1) / kBlock));
    timing_begin(f);
    switch (kernel) {
#define VPM_ADJ(KID)                                                                                            \
    case KID:                                                                                                   \
        hipLaunchKernelGGL(adj_target_kernel<KID>, grid, dim3(kBlock), 0, f->stream, f->P, np, b.U, b.J, b.X);  \
        hipLaunchKernelGGL(adj_source_kernel<KID>, grid, dim3(kBlock), 0, f->stream, f->P, np, b.U, b.J, b.X,   \
                           b.G, b.S);                                                                           \
        break;
        VPM_ADJ(VPM_KERNEL_SINGULAR) VPM_ADJ(VPM_KERNEL_GAUSSIAN) VPM_ADJ(VPM_KERNEL_GAUSSIANERF)
        VPM_ADJ(VPM_KERNEL_WINCKELMANS)
#undef VPM_ADJ
    default:
        return fail(VPM_EINVAL, "unknown kernel id");
    }
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "adjoint kernels");
    timing_end(f, 3);
    return VPM_OK;
}

int launch_uj_adjoint(vpm_field f, int kernel, int precision, const double* Ubar, const double* Jbar, double* Xbar,
                      double* Gbar, double* Sbar)
{
    const int64_t np = f->np;
    if (np == 0) return VPM_OK;
    const size_t n = (size_t)np;
    AdjBufs b;
    if (int rc = adj_buffers(f, b)) return rc;
    // a loss on U alone (every Jbar zero) takes the kernel without Jbar terms
    b.jbar = false;
    for (size_t k = 0; k < 9 * n && !b.jbar; ++k) b.jbar = !(Jbar[k] == 0.0);
    hipError_t e;
    if ((e = hipMemcpyAsync(b.U, Ubar, sizeof(double) * 3 * n, hipMemcpyHostToDevice, f->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(b.J, Jbar, sizeof(double) * 9 * n, hipMemcpyHostToDevice, f->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(cotangents)");
    if (int rc = launch_uj_adjoint_dev(f, kernel, precision, b)) return rc;
    if ((e = hipMemcpyAsync(Xbar, b.X, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, f->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(Gbar, b.G, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, f->stream)) != hipSuccess ||
        (e = hipMemcpyAsync(Sbar, b.S, sizeof(double) * n, hipMemcpyDeviceToHost, f->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(gradients)");
    return VPM_OK;
}

}  // namespace vpm
