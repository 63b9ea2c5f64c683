// particle_ops.hip -- per-particle FP64 kernels around the all-pairs passes.
//
// One lane per particle record, in place on the device-resident 46-double
// records.  These are O(N) and a negligible part of a step next to the
// O(N^2) passes; they run in FP64 with contraction off so that every
// expression rounds as the cited reference line does.
#include "vpm_internal.h"

#pragma clang fp contract(off)

namespace vpm {

__device__ __forceinline__ double sgn_jl(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }
__device__ __forceinline__ double cube(double x) { return x * x * x; }

// which: 0 -> U, vorticity, J, PSE (particlefield.jl:485-490);
//        1 -> SFS (:505-507);  2 -> M (timeintegration.jl:401-414)
__global__ __launch_bounds__(kBlock) void reset_kernel(double* __restrict__ P, int64_t np, int which)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    if (which == 0) {
        for (int k = RU; k < RPSE + 3; ++k) p[k] = 0.0;
    } else if (which == 1) {
        p[RSFS] = 0.0; p[RSFS + 1] = 0.0; p[RSFS + 2] = 0.0;
    } else {
        for (int k = 0; k < 9; ++k) p[RM + k] = 0.0;
    }
}

__device__ __forceinline__ void stretch(const double* J, const double* G, int transposed, double S[3])
{
    if (transposed) {     // (Gamma . grad') U (timeintegration.jl:489-493)
        S[0] = J[0] * G[0] + J[1] * G[1] + J[2] * G[2];
        S[1] = J[3] * G[0] + J[4] * G[1] + J[5] * G[2];
        S[2] = J[6] * G[0] + J[7] * G[1] + J[8] * G[2];
    } else {              // (Gamma . grad) U (:494-499)
        S[0] = J[0] * G[0] + J[3] * G[1] + J[6] * G[2];
        S[1] = J[1] * G[0] + J[4] * G[1] + J[7] * G[2];
        S[2] = J[2] * G[0] + J[5] * G[1] + J[8] * G[2];
    }
}

// update_particle_states, rVPM (timeintegration.jl:463-532)
__global__ __launch_bounds__(kBlock) void update_states_kernel(double* __restrict__ P, int64_t np, double a, double b,
                                                               double dt, double U0, double U1, double U2, double f,
                                                               double g, double zeta0, int transposed)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    const double C = p[RC];
    double* M = p + RM;
    double* G = p + RG;
    const double* J = p + RJ;
    M[0] = a * M[0] + dt * (p[RU + 0] + U0);
    M[1] = a * M[1] + dt * (p[RU + 1] + U1);
    M[2] = a * M[2] + dt * (p[RU + 2] + U2);
    p[RX + 0] += b * M[0];
    p[RX + 1] += b * M[1];
    p[RX + 2] += b * M[2];
    double MM[3];
    stretch(J, G, transposed, MM);
    const double Gnorm2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
    const double sgm = p[RSIG];
    double MM4;
    if (Gnorm2 > 0) {
        MM4 = (f + g) / (1 + 3 * f) * (MM[0] * G[0] + MM[1] * G[1] + MM[2] * G[2]);
        MM4 -= f / (1 + 3 * f) * (C * p[RSFS] * G[0] + C * p[RSFS + 1] * G[1] + C * p[RSFS + 2] * G[2]) * cube(sgm) / zeta0;
        MM4 /= Gnorm2;
    } else {
        MM4 = 0.0;
    }
    M[3] = a * M[3] + dt * (MM[0] - 3 * MM4 * G[0] - C * p[RSFS + 0] * cube(sgm) / zeta0);
    M[4] = a * M[4] + dt * (MM[1] - 3 * MM4 * G[1] - C * p[RSFS + 1] * cube(sgm) / zeta0);
    M[5] = a * M[5] + dt * (MM[2] - 3 * MM4 * G[2] - C * p[RSFS + 2] * cube(sgm) / zeta0);
    M[7] = a * M[7] - dt * (sgm * MM4);
    G[0] += b * M[3];
    G[1] += b * M[4];
    G[2] += b * M[5];
    p[RSIG] += b * M[7];
}

// relax_pedrizzetti / relax_correctedpedrizzetti (relaxation.jl:62-81,117-142)
__global__ __launch_bounds__(kBlock) void relax_kernel(double* __restrict__ P, int64_t np, double rlxf, int corrected)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    const double* J = p + RJ;
    double* G = p + RG;
    const double nrmw = sqrt((J[5] - J[7]) * (J[5] - J[7]) + (J[6] - J[2]) * (J[6] - J[2]) +
                             (J[1] - J[3]) * (J[1] - J[3]));
    if (nrmw != 0) {
        const double nrmGamma = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
        double b2 = 0;
        if (corrected)
            b2 = 1 - 2 * (1 - rlxf) * rlxf *
                         (1 - (G[0] * (J[5] - J[7]) + G[1] * (J[6] - J[2]) + G[2] * (J[1] - J[3])) / (nrmGamma * nrmw));
        G[0] = (1 - rlxf) * G[0] + rlxf * nrmGamma * (J[5] - J[7]) / nrmw;
        G[1] = (1 - rlxf) * G[1] + rlxf * nrmGamma * (J[6] - J[2]) / nrmw;
        G[2] = (1 - rlxf) * G[2] + rlxf * nrmGamma * (J[1] - J[3]) / nrmw;
        if (corrected) {
            const double sb2 = sqrt(b2);
            G[0] /= sb2; G[1] /= sb2; G[2] /= sb2;
        }
    }
}

// CoreSpreading, rungekutta3 branch (viscous.jl:170-175)
__global__ __launch_bounds__(kBlock) void core_spreading_kernel(double* __restrict__ P, int64_t np, double nu,
                                                                double a, double b, double dt)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    p[RM + 6] = a * p[RM + 6] + dt * 2 * nu;
    p[RSIG] = sqrt(p[RSIG] * p[RSIG] + b * p[RM + 6]);
}

// sigma *= alpha / sigma /= alpha on non-static particles (subfilterscale.jl:465-475, 526-536)
__global__ __launch_bounds__(kBlock) void scale_sigma_kernel(double* __restrict__ P, int64_t np, double alpha,
                                                             int divide)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    if (divide) p[RSIG] /= alpha;
    else p[RSIG] *= alpha;
}

// Test-filter bookkeeping of the pseudo-3-level procedure
// (subfilterscale.jl:481-521): M = 0, M[1:3] = stretching, M[4:6] = SFS.
__global__ __launch_bounds__(kBlock) void sfs_store_test_kernel(double* __restrict__ P, int64_t np, int transposed)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    for (int k = 0; k < 9; ++k) p[RM + k] = 0.0;
    double S[3];
    stretch(p + RJ, p + RG, transposed, S);
    p[RM + 0] = S[0]; p[RM + 1] = S[1]; p[RM + 2] = S[2];
    p[RM + 3] = p[RSFS]; p[RM + 4] = p[RSFS + 1]; p[RM + 5] = p[RSFS + 2];
}

// dynamicprocedure_pseudo3level_afterUJ (subfilterscale.jl:559-670) and
// clipping_backscatter (:225-247, 287-296).
__global__ __launch_bounds__(kBlock) void sfs_coefficient_kernel(double* __restrict__ P, int64_t np, int transposed,
                                                                 double alpha, double rlxf, double minC, double maxC,
                                                                 double zeta0, int force_positive, int clip,
                                                                 int* __restrict__ nan_flag)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    double* M = p + RM;
    const double* J = p + RJ;
    const double* G = p + RG;
    double* C = p + RC;
    if (transposed) {
        M[0] -= J[0] * G[0] + J[1] * G[1] + J[2] * G[2];
        M[1] -= J[3] * G[0] + J[4] * G[1] + J[5] * G[2];
        M[2] -= J[6] * G[0] + J[7] * G[1] + J[8] * G[2];
    } else {
        M[0] -= J[0] * G[0] + J[3] * G[1] + J[6] * G[2];
        M[1] -= J[1] * G[0] + J[4] * G[1] + J[7] * G[2];
        M[2] -= J[2] * G[0] + J[5] * G[1] + J[8] * G[2];
    }
    M[3] -= p[RSFS]; M[4] -= p[RSFS + 1]; M[5] -= p[RSFS + 2];

    double nume = M[0] * G[0] + M[1] * G[1] + M[2] * G[2];
    nume *= 3 * alpha - 2;
    double deno = M[3] * G[0] + M[4] * G[1] + M[5] * G[2];
    deno /= zeta0 / cube(p[RSIG]);
    if (C[2] == 0) {
        C[2] = deno;
        if (C[2] == 0) C[2] = 2.220446049250313e-16;   // eps()
    }
    nume = rlxf * nume + (1 - rlxf) * C[1];
    deno = rlxf * deno + (1 - rlxf) * C[2];
    if (fabs(nume / deno) > maxC) {
        if (fabs(deno) < fabs(C[2])) deno = sgn_jl(deno) * fabs(C[2]);
        if (fabs(nume / deno) >= maxC) nume = sgn_jl(nume) * fabs(deno) * maxC;
    } else if (fabs(nume / deno) < minC) {
        nume = sgn_jl(nume) * fabs(deno) * minC;
    }
    C[1] = nume;
    C[2] = deno;
    C[0] = C[1] / C[2];
    if (isnan(C[0])) *nan_flag = 1;     // every writer stores the same value
    if (force_positive) C[0] *= sgn_jl(C[0]);
    for (int k = 0; k < 9; ++k) M[k] = 0.0;
    if (clip) {
        if (C[0] * (G[0] * p[RSFS] + G[1] * p[RSFS + 1] + G[2] * p[RSFS + 2]) < 0) C[0] *= 0;
    }
}

// nextstep: U_prev = |U| for every particle (particlefield.jl:443-455)
__global__ __launch_bounds__(kBlock) void uprev_kernel(double* __restrict__ P, int64_t np)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    p[RUPREV] = sqrt(p[RU] * p[RU] + p[RU + 1] * p[RU + 1] + p[RU + 2] * p[RU + 2]);
}

// Pullback of update_states_kernel for ReformulatedVPM without SFS (C S = 0;
// the state_pullback of rrules.jl:436-555, derived from the update
// timeintegration.jl:467-528 itself).  With m1 = Mbar1' + b Xbar',
// m23 = Mbar23' + b sigmabar', m2 = Mbar2' + b Gammabar', MM = stretching,
// G2 = |Gamma|^2, MM4 = A (MM . Gamma)/G2, A = (f + g)/(1 + 3 f):
//   Xbar = Xbar', Mbar1 = a m1, Ubar = dt m1, Mbar23 = a m23,
//   sigmabar = sigmabar' - dt MM4 m23, Mbar2 = a m2,
//   MM4bar = -dt sigma m23 - 3 dt Gamma . m2,
//   MMbar = dt m2 + MM4bar A Gamma/G2,
//   Gammabar = Gammabar' - 3 dt MM4 m2 + MM4bar (A MM - 2 MM4 Gamma)/G2 + (dMM/dGamma)^T MMbar,
//   Jbar = (dMM/dJ)^T MMbar
// (the full sum over components in d(MM4 Gamma)/dJ; rrules.jl:521 keeps only
// the diagonal term).  B: cotangent records (NF rows per particle) of the
// updated state in, of the state before the update out; P: the state before
// the update with its U and J.  Ub (3 np) and Jb (9 np) receive Ubar, Jbar.
__global__ __launch_bounds__(kBlock) void update_states_adjoint_kernel(
    const double* __restrict__ P, double* __restrict__ B, int64_t np, double a, double b, double dt, double f,
    double g, int transposed, double* __restrict__ Ub, double* __restrict__ Jb)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    const double* p = P + i * NF;
    double* q = B + i * NF;
    if (p[RSTATIC] != 0) {
        for (int k = 0; k < 3; ++k) Ub[3 * i + k] = 0.0;
        for (int k = 0; k < 9; ++k) Jb[9 * i + k] = 0.0;
        return;
    }
    const double A = (f + g) / (1 + 3 * f);
    const double* G = p + RG;
    const double* J = p + RJ;
    const double sgm = p[RSIG];
    double MM[3];
    stretch(J, G, transposed, MM);
    const double G2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
    const double MM4 = G2 > 0 ? A * (MM[0] * G[0] + MM[1] * G[1] + MM[2] * G[2]) / G2 : 0.0;
    double m1[3], m2[3];
    for (int k = 0; k < 3; ++k) {
        m1[k] = q[RM + k] + b * q[RX + k];
        m2[k] = q[RM + 3 + k] + b * q[RG + k];
    }
    const double m23 = q[RM + 7] + b * q[RSIG];
    const double MM4b = -dt * sgm * m23 - 3 * dt * (G[0] * m2[0] + G[1] * m2[1] + G[2] * m2[2]);
    double MMb[3], Gb[3];
    for (int k = 0; k < 3; ++k) {
        q[RM + k] = a * m1[k];
        Ub[3 * i + k] = dt * m1[k];
        q[RM + 3 + k] = a * m2[k];
        MMb[k] = dt * m2[k];
        Gb[k] = q[RG + k] - 3 * dt * MM4 * m2[k];
    }
    q[RM + 7] = a * m23;
    q[RSIG] = q[RSIG] - dt * MM4 * m23;
    if (G2 > 0) {
        for (int k = 0; k < 3; ++k) {
            MMb[k] += MM4b * A * G[k] / G2;
            Gb[k] += MM4b * (A * MM[k] - 2 * MM4 * G[k]) / G2;
        }
    }
    double jb[9];
    for (int k = 0; k < 9; ++k) jb[k] = 0.0;
    // MM[e] = sum_x J(x, e) G[x] (transposed) or J(e, x) G[x]; J(r, c) = J[3 c + r]
    for (int e = 0; e < 3; ++e)
        for (int x = 0; x < 3; ++x) {
            const int idx = transposed ? 3 * e + x : 3 * x + e;
            jb[idx] += MMb[e] * G[x];
            Gb[x] += J[idx] * MMb[e];
        }
    for (int k = 0; k < 9; ++k) Jb[9 * i + k] = jb[k];
    for (int k = 0; k < 3; ++k) q[RG + k] = Gb[k];
}

// out[i][r] = P[i][row0 + r]: rows of the records, packed
__global__ __launch_bounds__(kBlock) void gather_rows_kernel(const double* __restrict__ P, int64_t np, int row0,
                                                             int nr, double* __restrict__ out)
{
    const int64_t k = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (k >= np * nr) return;
    const int64_t i = k / nr;
    const int r = (int)(k - i * nr);
    out[k] = P[i * NF + row0 + r];
}

// B[i] += (dX, dGamma, dsigma) of the U/J adjoint
__global__ __launch_bounds__(kBlock) void adj_add_bar_kernel(double* __restrict__ B, int64_t np,
                                                             const double* __restrict__ dX,
                                                             const double* __restrict__ dG,
                                                             const double* __restrict__ dS)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* q = B + i * NF;
    for (int k = 0; k < 3; ++k) {
        q[RX + k] += dX[3 * i + k];
        q[RG + k] += dG[3 * i + k];
    }
    q[RSIG] += dS[i];
}

// Cotangent of the state before an rk3 step: the step starts from M = 0 for
// non-static particles (timeintegration.jl:401-414), and U, J are outputs.
__global__ __launch_bounds__(kBlock) void adj_step_start_kernel(const double* __restrict__ P, double* __restrict__ B,
                                                                int64_t np)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* q = B + i * NF;
    for (int k = 0; k < 3; ++k) q[RU + k] = 0.0;
    for (int k = 0; k < 9; ++k) q[RJ + k] = 0.0;
    if (P[i * NF + RSTATIC] == 0)
        for (int k = 0; k < 9; ++k) q[RM + k] = 0.0;
}

static inline unsigned nb(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

#define VPM_CHECK_LAUNCH(what)                                                  \
    do {                                                                        \
        hipError_t e_ = hipGetLastError();                                      \
        if (e_ != hipSuccess) return hip_fail(e_, what);                        \
    } while (0)

int launch_reset(vpm_field f, int which)
{
    if (f->np == 0) return VPM_OK;
    hipLaunchKernelGGL(reset_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, f->P, f->np, which);
    VPM_CHECK_LAUNCH("reset");
    return VPM_OK;
}

int launch_update_states(vpm_field f, double a, double b, double dt, const double Uinf[3], double fr, double gr,
                         double zeta0, int transposed)
{
    if (f->np == 0) return VPM_OK;
    hipLaunchKernelGGL(update_states_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, f->P, f->np, a, b, dt,
                       Uinf[0], Uinf[1], Uinf[2], fr, gr, zeta0, transposed);
    VPM_CHECK_LAUNCH("update_particle_states");
    f->perm_valid = false;
    return VPM_OK;
}

int launch_relax(vpm_field f, double rlxf, int corrected)
{
    if (f->np == 0) return VPM_OK;
    hipLaunchKernelGGL(relax_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, f->P, f->np, rlxf, corrected);
    VPM_CHECK_LAUNCH("relax");
    return VPM_OK;
}

int launch_core_spreading(vpm_field f, double nu, double a, double b, double dt)
{
    if (f->np == 0) return VPM_OK;
    hipLaunchKernelGGL(core_spreading_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, f->P, f->np, nu, a, b, dt);
    VPM_CHECK_LAUNCH("core_spreading");
    return VPM_OK;
}

int launch_scale_sigma(vpm_field f, double alpha, int divide)
{
    if (f->np == 0) return VPM_OK;
    hipLaunchKernelGGL(scale_sigma_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, f->P, f->np, alpha, divide);
    VPM_CHECK_LAUNCH("scale_sigma");
    return VPM_OK;
}

int launch_sfs_store_test(vpm_field f, int transposed)
{
    if (f->np == 0) return VPM_OK;
    hipLaunchKernelGGL(sfs_store_test_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, f->P, f->np, transposed);
    VPM_CHECK_LAUNCH("sfs_store_test");
    return VPM_OK;
}

int launch_sfs_coefficient(vpm_field f, int transposed, double alpha, double rlxf, double minC, double maxC,
                           double zeta0, int force_positive, int clip, int* nan_flag)
{
    *nan_flag = 0;
    if (f->np == 0) return VPM_OK;
    int rc;
    if ((rc = ensure(f->flag, sizeof(int)))) return rc;
    hipError_t e = hipMemsetAsync(f->flag.p, 0, sizeof(int), f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemsetAsync(flag)");
    hipLaunchKernelGGL(sfs_coefficient_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, f->P, f->np, transposed,
                       alpha, rlxf, minC, maxC, zeta0, force_positive, clip, (int*)f->flag.p);
    VPM_CHECK_LAUNCH("sfs_coefficient");
    e = hipMemcpyAsync(nan_flag, f->flag.p, sizeof(int), hipMemcpyDeviceToHost, f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpyAsync(flag)");
    e = hipStreamSynchronize(f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return VPM_OK;
}

int launch_update_uprev(vpm_field f)
{
    if (f->np == 0) return VPM_OK;
    hipLaunchKernelGGL(uprev_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, f->P, f->np);
    VPM_CHECK_LAUNCH("uprev");
    return VPM_OK;
}

int launch_update_states_adjoint(vpm_field f, double* B, double a, double b, double dt, double fr, double gr,
                                 int transposed, double* Ub, double* Jb)
{
    if (f->np == 0) return VPM_OK;
    hipLaunchKernelGGL(update_states_adjoint_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, f->P, B, f->np, a,
                       b, dt, fr, gr, transposed, Ub, Jb);
    VPM_CHECK_LAUNCH("update_states_adjoint");
    return VPM_OK;
}

int launch_gather_rows(vpm_field f, int row0, int nrows, double* out)
{
    const int64_t n = f->np * nrows;
    if (n == 0) return VPM_OK;
    hipLaunchKernelGGL(gather_rows_kernel, dim3(nb(n)), dim3(kBlock), 0, f->stream, f->P, f->np, row0, nrows, out);
    VPM_CHECK_LAUNCH("gather_rows");
    return VPM_OK;
}

int launch_adj_add_bar(vpm_field f, double* B, const double* dX, const double* dG, const double* dS)
{
    if (f->np == 0) return VPM_OK;
    hipLaunchKernelGGL(adj_add_bar_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, B, f->np, dX, dG, dS);
    VPM_CHECK_LAUNCH("adj_add_bar");
    return VPM_OK;
}

int launch_adj_step_start(vpm_field f, double* B)
{
    if (f->np == 0) return VPM_OK;
    hipLaunchKernelGGL(adj_step_start_kernel, dim3(nb(f->np)), dim3(kBlock), 0, f->stream, f->P, B, f->np);
    VPM_CHECK_LAUNCH("adj_step_start");
    return VPM_OK;
}

}  // namespace vpm
