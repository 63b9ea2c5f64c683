// rbf.hip -- core-spreading reset: target vorticity from the zeta field,
// core sizes back to sgm0, and new strengths by RBF interpolation with the
// conjugate gradient method (src/FLOWVPM_viscous.jl:202-220 and
// rbf_conjugategradient, :314-478).
//
// Storage as in the reference: target vorticity in M[6:9], solution x in
// M[0:3], residual r in M[3:6], search direction p in Gamma (the zeta
// field A p lands in J[0:3]).  Vectors and dot products run over the
// non-static particles (iterator(pfield)); the zeta field over all of them
// (zeta_direct, viscous.jl:486-493).  The three components are three
// independent CG solves sharing each zeta evaluation.  Dot products are
// FP64 block partials summed in a fixed order, so the iteration is
// deterministic.  The scalars alpha, beta and the convergence flags live on
// the host.
#include <cmath>

#include "vpm_internal.h"

namespace vpm {

#pragma clang fp contract(off)

// per-block FP64 sums of three products over non-static particles
__device__ __forceinline__ void block_sum3(double a0, double a1, double a2, double* __restrict__ part)
{
    __shared__ double red[3][kBlock];
    const int t = threadIdx.x;
    red[0][t] = a0; red[1][t] = a1; red[2][t] = a2;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (t < w) {
            red[0][t] += red[0][t + w];
            red[1][t] += red[1][t + w];
            red[2][t] += red[2][t + w];
        }
        __syncthreads();
    }
    if (t < 3) part[3 * blockIdx.x + t] = red[t][0];
}

// M[6:9] <- J[0:3] (zeta field = target vorticity), sigma <- sgm0
__global__ __launch_bounds__(kBlock) void rbf_target_kernel(double* __restrict__ P, int64_t np, double sgm0)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    for (int c = 0; c < 3; ++c) p[RM + 6 + c] = p[RJ + c];
    p[RSIG] = sgm0;
}

// x0 = omega vol; Gamma = x0
__global__ __launch_bounds__(kBlock) void rbf_x0_kernel(double* __restrict__ P, int64_t np)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    for (int c = 0; c < 3; ++c) {
        p[RM + c] = p[RM + 6 + c] * p[RVOL];
        p[RG + c] = p[RM + c];
    }
}

// r0 = omega - A x0; p0 = r0; partial sums of r0.r0
__global__ __launch_bounds__(kBlock) void rbf_r0_kernel(double* __restrict__ P, int64_t np, double* __restrict__ part)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double s[3] = {0.0, 0.0, 0.0};
    if (i < np) {
        double* p = P + i * NF;
        if (p[RSTATIC] == 0) {
            for (int c = 0; c < 3; ++c) {
                p[RM + 3 + c] = p[RM + 6 + c] - p[RJ + c];
                p[RG + c] = p[RM + 3 + c];
                s[c] = p[RM + 3 + c] * p[RM + 3 + c];
            }
        }
    }
    block_sum3(s[0], s[1], s[2], part);
}

// partial sums of p.Ap (Gamma . J)
__global__ __launch_bounds__(kBlock) void rbf_pap_kernel(const double* __restrict__ P, int64_t np,
                                                         double* __restrict__ part)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double s[3] = {0.0, 0.0, 0.0};
    if (i < np) {
        const double* p = P + i * NF;
        if (p[RSTATIC] == 0)
            for (int c = 0; c < 3; ++c) s[c] = p[RG + c] * p[RJ + c];
    }
    block_sum3(s[0], s[1], s[2], part);
}

// x += alpha p; r -= alpha A p; partial sums of r.r
__global__ __launch_bounds__(kBlock) void rbf_step_kernel(double* __restrict__ P, int64_t np, double a0, double a1,
                                                          double a2, double* __restrict__ part)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    double s[3] = {0.0, 0.0, 0.0};
    if (i < np) {
        double* p = P + i * NF;
        if (p[RSTATIC] == 0) {
            const double a[3] = {a0, a1, a2};
            for (int c = 0; c < 3; ++c) {
                p[RM + c] += a[c] * p[RG + c];
                p[RM + c + 3] -= a[c] * p[RJ + c];
                s[c] = p[RM + c + 3] * p[RM + c + 3];
            }
        }
    }
    block_sum3(s[0], s[1], s[2], part);
}

// p = r + beta p
__global__ __launch_bounds__(kBlock) void rbf_dir_kernel(double* __restrict__ P, int64_t np, double b0, double b1,
                                                         double b2)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    const double b[3] = {b0, b1, b2};
    for (int c = 0; c < 3; ++c) p[RG + c] = p[RM + c + 3] + b[c] * p[RG + c];
}

// Gamma = x
__global__ __launch_bounds__(kBlock) void rbf_final_kernel(double* __restrict__ P, int64_t np)
{
    const int64_t i = (int64_t)blockIdx.x * kBlock + threadIdx.x;
    if (i >= np) return;
    double* p = P + i * NF;
    if (p[RSTATIC] != 0) return;
    for (int c = 0; c < 3; ++c) p[RG + c] = p[RM + c];
}

// sum of the block partials in a fixed order: thread t adds partials
// t, t + kBlock, ... in turn, then a fixed tree over the threads
__global__ __launch_bounds__(kBlock) void rbf_reduce_kernel(const double* __restrict__ part, int nb,
                                                            double* __restrict__ out)
{
    const int t = threadIdx.x;
    double s[3] = {0.0, 0.0, 0.0};
    for (int b = t; b < nb; b += kBlock)
        for (int c = 0; c < 3; ++c) s[c] += part[3 * b + c];
    block_sum3(s[0], s[1], s[2], out);   // one block: out[0..2]
}
#pragma clang fp contract(on)

static int reduce3(vpm_field f, double* part, int nb, double* dout, double out[3])
{
    hipLaunchKernelGGL(rbf_reduce_kernel, dim3(1), dim3(kBlock), 0, f->stream, (const double*)part, nb, dout);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "rbf_reduce");
    if ((e = hipMemcpyAsync(out, dout, 3 * sizeof(double), hipMemcpyDeviceToHost, f->stream)) != hipSuccess)
        return hip_fail(e, "hipMemcpyAsync(rbf sums)");
    if ((e = hipStreamSynchronize(f->stream)) != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    return VPM_OK;
}

#define VPM_RBF_LAUNCH(kern, ...)                                                                 \
    do {                                                                                          \
        hipLaunchKernelGGL(kern, dim3(nb), dim3(kBlock), 0, f->stream, __VA_ARGS__);               \
        hipError_t e_ = hipGetLastError();                                                        \
        if (e_ != hipSuccess) return hip_fail(e_, #kern);                                         \
    } while (0)

// rbf_conjugategradient (viscous.jl:314-478) on the target vorticity already
// in M[7:9]
int launch_rbf_cg(vpm_field f, int kernel, int precision, int itmax, double tol, int iterror, int* iters,
                  double res[3])
{
    const int64_t np = f->np;
    *iters = 0;
    res[0] = res[1] = res[2] = 0.0;
    if (np == 0) return VPM_OK;
    const int nb = (int)((np + kBlock - 1) / kBlock);
    int rc;
    if ((rc = ensure(f->flag, sizeof(double) * (3 * (size_t)nb + 3)))) return rc;
    double* part = (double*)f->flag.p;
    double* dsum = part + 3 * (size_t)nb;
    const double eps2 = 2 * 2.220446049250313e-16;
    VPM_RBF_LAUNCH(rbf_x0_kernel, f->P, np);                              // :345-352
    if ((rc = launch_zeta(f, kernel, precision))) return rc;                // :355
    VPM_RBF_LAUNCH(rbf_r0_kernel, f->P, np, part);                         // :357-368
    double rr0[3], rrs[3], prev[3], pAp[3], alpha[3], beta[3];
    if ((rc = reduce3(f, part, nb, dsum, rr0))) return rc;
    bool flags[3];
    for (int c = 0; c < 3; ++c) {
        rrs[c] = rr0[c];
        flags[c] = std::sqrt(rr0[c]) > tol || std::sqrt(rrs[c] / rr0[c]) > tol;
    }
    int it = 0;
    bool unconverged = false;
    for (it = 1; it <= itmax; ++it) {                                        // :376-447
        if (!(flags[0] || flags[1] || flags[2])) { --it; break; }
        if ((rc = launch_zeta(f, kernel, precision))) return rc;
        VPM_RBF_LAUNCH(rbf_pap_kernel, (const double*)f->P, np, part);
        if ((rc = reduce3(f, part, nb, dsum, pAp))) return rc;
        for (int c = 0; c < 3; ++c) {
            alpha[c] = flags[c] ? rrs[c] / pAp[c] : 0.0;   // Julia's false is a strong zero
            prev[c] = rrs[c];
        }
        VPM_RBF_LAUNCH(rbf_step_kernel, f->P, np, alpha[0], alpha[1], alpha[2], part);
        if ((rc = reduce3(f, part, nb, dsum, rrs))) return rc;
        for (int c = 0; c < 3; ++c) {
            beta[c] = rrs[c] / prev[c];
            if (std::fabs(prev[c]) <= eps2) beta[c] = 1;
        }
        VPM_RBF_LAUNCH(rbf_dir_kernel, f->P, np, beta[0], beta[1], beta[2]);
        for (int c = 0; c < 3; ++c)
            flags[c] = flags[c] && (std::fabs(rr0[c]) <= eps2 ? false : std::sqrt(rrs[c] / rr0[c]) > tol);
        if (it == itmax && (flags[0] || flags[1] || flags[2])) unconverged = true;
    }
    if (it > itmax) it = itmax;
    VPM_RBF_LAUNCH(rbf_final_kernel, f->P, np);                            // :450-455
    hipError_t e = hipStreamSynchronize(f->stream);
    if (e != hipSuccess) return hip_fail(e, "hipStreamSynchronize");
    *iters = it;
    for (int c = 0; c < 3; ++c) res[c] = rr0[c] > 0 ? std::sqrt(rrs[c] / rr0[c]) : 0.0;
    if (unconverged && iterror)
        return fail(VPM_ECONVERGE, "RBF interpolation: maximum number of iterations reached before convergence");
    return VPM_OK;
}

// viscous.jl:202-220 without the timer: zeta field -> target vorticity,
// sigma <- sgm0, then the CG
int launch_core_spreading_reset(vpm_field f, int kernel, int precision, double sgm0, int itmax, double tol,
                                int iterror, int* iters, double res[3])
{
    const int64_t np = f->np;
    *iters = 0;
    res[0] = res[1] = res[2] = 0.0;
    if (np == 0) return VPM_OK;
    const int nb = (int)((np + kBlock - 1) / kBlock);
    int rc;
    if ((rc = launch_zeta(f, kernel, precision))) return rc;                // viscous.jl:204
    VPM_RBF_LAUNCH(rbf_target_kernel, f->P, np, sgm0);                    // :206-214
    return launch_rbf_cg(f, kernel, precision, itmax, tol, iterror, iters, res);
}

}  // namespace vpm
