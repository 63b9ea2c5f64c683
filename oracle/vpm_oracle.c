/*
 * vpm_oracle.c -- FP64 CPU restatement of FLOWVPM's particle-to-particle path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library (see the header).
 * Built with -ffp-contract=off so that every expression rounds as the Julia
 * reference does (Julia does not fuse a*b+c unless asked to).  Operation
 * order inside each expression follows the cited reference line, including
 * Julia's left-to-right evaluation of chained * and /, and x^3 == x*x*x
 * (Base.literal_pow).
 *
 * Pinning (what checks each function; tests/test_oracle.py):
 *   vpmo_erf64         mpmath erf at 50 digits over all five fdlibm branches
 *   vpmo_g_dgdr/zeta   mpmath closed forms; g' = 4 pi q^2 zeta identity
 *   vpmo_g_derivs      mpmath derivative of the closed form g'
 *   vpmo_direct2       mpmath brute force (N<=64, four kernels), 2-body
 *                      singular closed form (scripts/check_fmm.jl:39-71),
 *                      tr J = 0, rotation/scaling/translation invariance
 *   vpmo_direct_ranges equals vpmo_direct2 over the union of the ranges
 *   vpmo_estr_direct   mpmath brute force; zero when all J are equal
 *   vpmo_update_particle_states / vpmo_rk3_step
 *                      invariants of the update's ODE, with planted mistakes
 *                      shown to fail them: |Gamma| sigma^((1-3g)/(f+g))
 *                      per particle (third-order drift over whole steps,
 *                      second-order over one Euler-form call with random J),
 *                      |Gamma|^f sigma with SFS at g = 0, sum Gamma conserved
 *                      by transposed stretching under cVPM with uniform
 *                      sigma (not by the classic scheme); exact free-stream
 *                      advection; third-order RK3 convergence; the
 *                      single-ring translation speed test of
 *                      test/runtests_singlevortexring.jl:22-23,143
 *   vpmo_relax         Pedrizzetti alignment at rlxf = 1, identity at 0,
 *                      corrected-norm preservation
 *   vpmo_core_spreading_rk3  sigma^2 grows by exactly 2 nu dt
 *   vpmo_sfs_before / vpmo_sfs_after
 *                      numerator and denominator rebuilt from independent
 *                      evaluations of the two filter widths; the Lagrangian
 *                      memory C2_k = nume (1 - (1 - rlxf)^k) under a frozen
 *                      M; the coefficient's bounds and clipping rules
 *   vpmo_uj_adjoint    central finite differences of vpmo_loss_local;
 *                      equals vpmo_uj_adjoint_rrules (the reference's
 *                      UJS_pullback restated, a cross-check) to FP64 rounding
 *   vpmo_estr_ranges   equals vpmo_estr_direct over the pair list of the whole
 *                      field minus its static rows, and the sum of its
 *                      sub-range lists
 *   vpmo_core_spreading_reset  converged Gamma reproduces the target
 *                      vorticity; iterated to tight tolerance it equals a dense
 *                      solve of the zeta system
 *   vpmo_zeta_direct   mpmath brute force; one particle gives Gamma zeta(0)/sigma^3
 *   vpmo_zeta_ranges   the single pair (all, all) onto zeroed J equals
 *                      vpmo_zeta_direct; sub-ranges sum to the same
 *   vpmo_update_particle_states_adjoint / vpmo_rk3_step_adjoint
 *                      central finite differences of vpmo_update_particle_states
 *                      and of a loss through whole vpmo_rk3_step calls
 */
#include "vpm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* Row offsets in the 46-row particle record (particlefield.jl:239-252). */
enum {
    RX = 0, RG = 3, RSIG = 6, RVOL = 7, RCIRC = 8, RU = 9, RW = 12, RJ = 15,
    RPSE = 24, RM = 27, RC = 36, RSFS = 39, RSTATIC = 42, RUPREV = 43
};

/* src/FLOWVPM.jl:62-66 */
#define PI_D 3.141592653589793
static double const1(void) { return 1.0 / pow(2.0 * PI_D, 1.5); }
static double const2(void) { return sqrt(2.0 / PI_D); }
static double const3(void) { return 3.0 / (4.0 * PI_D); }
static double const4(void) { return 1.0 / (4.0 * PI_D); }
static double sqr2(void) { return sqrt(2.0); }

/* ---- erf: the fdlibm s_erf.c method as restated by custom_erf64
 *      (src/FLOWVPM_gpu_erf.jl:62-121 coefficients, :159-191 branches). */
static const double e_erx = 8.45062911510467529297e-01;
static const double e_pp[5] = { 1.28379167095512558561e-01, -3.25042107247001499370e-01,
    -2.84817495755985104766e-02, -5.77027029648944159157e-03, -2.37630166566501626084e-05 };
static const double e_qq[5] = { 3.97917223959155352819e-01, 6.50222499887672944485e-02,
    5.08130628187576562776e-03, 1.32494738004321644526e-04, -3.96022827877536812320e-06 };
static const double e_pa[7] = { -2.36211856075265944077e-03, 4.14856118683748331666e-01,
    -3.72207876035701323847e-01, 3.18346619901161753674e-01, -1.10894694282396677476e-01,
    3.54783043256182359371e-02, -2.16637559486879084300e-03 };
static const double e_qa[6] = { 1.06420880400844228286e-01, 5.40397917702171048937e-01,
    7.18286544141962662868e-02, 1.26171219808761642112e-01, 1.36370839120290507362e-02,
    1.19844998467991074170e-02 };
static const double e_ra[8] = { -9.86494403484714822705e-03, -6.93858572707181764372e-01,
    -1.05586262253232909814e+01, -6.23753324503260060396e+01, -1.62396669462573470355e+02,
    -1.84605092906711035994e+02, -8.12874355063065934246e+01, -9.81432934416914548592e+00 };
static const double e_sa[8] = { 1.96512716674392571292e+01, 1.37657754143519042600e+02,
    4.34565877475229228821e+02, 6.45387271733267880336e+02, 4.29008140027567833386e+02,
    1.08635005541779435134e+02, 6.57024977031928170135e+00, -6.04244152148580987438e-02 };
static const double e_rb[7] = { -9.86494292470009928597e-03, -7.99283237680523006574e-01,
    -1.77579549177547519889e+01, -1.60636384855821916062e+02, -6.37566443368389627722e+02,
    -1.02509513161107724954e+03, -4.83519191608651397019e+02 };
static const double e_sb[7] = { 3.03380607434824582924e+01, 3.25792512996573918826e+02,
    1.53672958608443695994e+03, 3.19985821950859553908e+03, 2.55305040643316442583e+03,
    4.74528541206955367215e+02, -2.24409524465858183362e+01 };

/* Julia's sign(): +-1 for nonzero, the argument itself for +-0 and NaN. */
static double jl_sign(double x) { return x > 0 ? 1.0 : (x < 0 ? -1.0 : x); }

double vpmo_erf64(double x)
{
    const double xabs = fabs(x), sgn = jl_sign(x), one = 1.0;
    double val = sgn * one;
    if (xabs < 0.84375) {
        double z = x * x;
        double r = e_pp[0] + z * (e_pp[1] + z * (e_pp[2] + z * (e_pp[3] + z * e_pp[4])));
        double s = one + z * (e_qq[0] + z * (e_qq[1] + z * (e_qq[2] + z * (e_qq[3] + z * e_qq[4]))));
        double y = r / s;
        val = sgn * (xabs + xabs * y);
    } else if (xabs < 1.25) {
        double s = xabs - one;
        double P = e_pa[0] + s * (e_pa[1] + s * (e_pa[2] + s * (e_pa[3] + s * (e_pa[4] + s * (e_pa[5] + s * e_pa[6])))));
        double Q = one + s * (e_qa[0] + s * (e_qa[1] + s * (e_qa[2] + s * (e_qa[3] + s * (e_qa[4] + s * e_qa[5])))));
        val = sgn * (e_erx + P / Q);
    } else if (xabs < 2.857142857142857) {
        double s = one / (x * x);
        double R = e_ra[0] + s * (e_ra[1] + s * (e_ra[2] + s * (e_ra[3] + s * (e_ra[4] + s * (e_ra[5] + s * (e_ra[6] + s * e_ra[7]))))));
        double S = one + s * (e_sa[0] + s * (e_sa[1] + s * (e_sa[2] + s * (e_sa[3] + s * (e_sa[4] + s * (e_sa[5] + s * (e_sa[6] + s * e_sa[7])))))));
        double r = exp(-x * x - 0.5625 + R / S);
        val = sgn * (one - r / xabs);
    } else if (xabs < 6.0) {
        double s = one / (x * x);
        double R = e_rb[0] + s * (e_rb[1] + s * (e_rb[2] + s * (e_rb[3] + s * (e_rb[4] + s * (e_rb[5] + s * e_rb[6])))));
        double S = one + s * (e_sb[0] + s * (e_sb[1] + s * (e_sb[2] + s * (e_sb[3] + s * (e_sb[4] + s * (e_sb[5] + s * e_sb[6]))))));
        double r = exp(-x * x - 0.5625 + R / S);
        val = sgn * (one - r / xabs);
    }
    return val;
}

/* ---- regularization kernels (src/FLOWVPM_kernel.jl:45-84) ---- */
void vpmo_g_dgdr(int kernel, double r, double* g, double* dg)
{
    switch (kernel) {
    case VPMO_SINGULAR:                                   /* kernel.jl:48 */
        *g = 1.0; *dg = 0.0; return;
    case VPMO_GAUSSIAN: {                                 /* kernel.jl:63-66 */
        double aux = exp(-r * r * r);
        *g = 1 - aux; *dg = 3 * r * r * aux; return;
    }
    case VPMO_GAUSSIANERF: {                              /* kernel.jl:54-57 */
        double aux = const2() * r * exp(-r * r / 2);
        *g = vpmo_erf64(r / sqr2()) - aux; *dg = r * aux; return;
    }
    case VPMO_WINCKELMANS: {                              /* kernel.jl:77-84 */
        double aux0 = r * r + 1;
        double aux02 = aux0 * aux0;
        aux0 = aux02 * aux02 * aux0;
        aux0 = sqrt(aux0);
        *g = r * r * r * (r * r + 2.5) / aux0;
        *dg = 7.5 * r * r / (aux0 * (r * r + 1));
        return;
    }
    }
    *g = NAN; *dg = NAN;
}

double vpmo_zeta(int kernel, double r)
{
    switch (kernel) {
    case VPMO_SINGULAR: return r == 0 ? 1.0 : 0.0;                 /* kernel.jl:45 */
    case VPMO_GAUSSIAN: return const3() * exp(-r * r * r);         /* kernel.jl:60 */
    case VPMO_GAUSSIANERF: return const1() * exp(-r * r / 2);      /* kernel.jl:51 */
    case VPMO_WINCKELMANS: {                                       /* kernel.jl:69-74 */
        double temp = r * r + 1;
        double temp2 = temp * temp * temp;
        temp = temp2 * temp2 * temp;
        return const4() * 7.5 / sqrt(temp);
    }
    }
    return NAN;
}

/* g, g' and g'' (the second derivative is not in the reference; it is the
 * derivative of the closed-form dgdr_* of kernel.jl:47,53,62,76). */
void vpmo_g_derivs(int kernel, double q, double* g, double* dg, double* d2g)
{
    vpmo_g_dgdr(kernel, q, g, dg);
    switch (kernel) {
    case VPMO_SINGULAR: *d2g = 0.0; return;
    case VPMO_GAUSSIAN: *d2g = (6 * q - 9 * q * q * q * q) * exp(-q * q * q); return;
    case VPMO_GAUSSIANERF: *d2g = const2() * (2 * q - q * q * q) * exp(-q * q / 2); return;
    case VPMO_WINCKELMANS: {
        double t = q * q + 1;
        *d2g = 7.5 * q * (2 - 5 * q * q) / pow(t, 4.5);
        return;
    }
    }
    *d2g = NAN;
}

/* One source -> one target pair, the body of fmm.jl:113-161.  Adds into
 * acc[0..2] (U) and acc[3..11] (J, column-major). */
static inline void pair_uj(int kernel, double tx, double ty, double tz,
                           const double* src, double* acc)
{
    const double c4 = const4();
    const double gamma_x = src[RG], gamma_y = src[RG + 1], gamma_z = src[RG + 2];
    const double sigma = src[RSIG];
    double dx = tx - src[RX], dy = ty - src[RX + 1], dz = tz - src[RX + 2];
    double r2 = dx * dx + dy * dy + dz * dz;
    if (r2 == 0.0) return;                                   /* fmm.jl:118 */
    double r = sqrt(r2);
    double g_sgm, dg_sgmdr;
    vpmo_g_dgdr(kernel, r / sigma, &g_sgm, &dg_sgmdr);
    double r3inv = 1.0 / (r2 * r);
    double crss1 = -c4 * r3inv * (dy * gamma_z - dz * gamma_y);
    double crss2 = -c4 * r3inv * (dz * gamma_x - dx * gamma_z);
    double crss3 = -c4 * r3inv * (dx * gamma_y - dy * gamma_x);
    acc[0] += g_sgm * crss1;
    acc[1] += g_sgm * crss2;
    acc[2] += g_sgm * crss3;
    double aux = dg_sgmdr / (sigma * r) - 3 * g_sgm / r2;
    double aux2 = -c4 * g_sgm * r3inv;
    acc[3 + 0] += aux * crss1 * dx;
    acc[3 + 1] += aux * crss2 * dx - aux2 * gamma_z;
    acc[3 + 2] += aux * crss3 * dx + aux2 * gamma_y;
    acc[3 + 3] += aux * crss1 * dy + aux2 * gamma_z;
    acc[3 + 4] += aux * crss2 * dy;
    acc[3 + 5] += aux * crss3 * dy - aux2 * gamma_x;
    acc[3 + 6] += aux * crss1 * dz - aux2 * gamma_y;
    acc[3 + 7] += aux * crss2 * dz + aux2 * gamma_x;
    acc[3 + 8] += aux * crss3 * dz;
}

static void add_acc(double* tp, const double* acc)
{
    for (int k = 0; k < 3; k++) tp[RU + k] += acc[k];            /* fmm.jl:171 */
    for (int k = 0; k < 9; k++) tp[RJ + k] += acc[3 + k];        /* fmm.jl:172-175 */
}

void vpmo_direct2(double* Pt, int64_t st, int64_t t0, int64_t t1,
                  const double* Ps, int64_t ss, int64_t s0, int64_t s1,
                  int kernel, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = t0; i < t1; i++) {
        double* tp = Pt + i * st;
        double acc[12] = {0};
        const double tx = tp[RX], ty = tp[RX + 1], tz = tp[RX + 2];
        for (int64_t j = s0; j < s1; j++) pair_uj(kernel, tx, ty, tz, Ps + j * ss, acc);
        add_acc(tp, acc);
    }
}

void vpmo_direct_ranges(double* P, int64_t stride, int kernel,
                        const int64_t* tb, const int64_t* te,
                        const int64_t* sb, const int64_t* se, int64_t npairs)
{
    /* buffer over every particle touched, summed in list order */
    int64_t lo = -1, hi = -1;
    for (int64_t k = 0; k < npairs; k++) {
        if (lo < 0 || tb[k] < lo) lo = tb[k];
        if (te[k] > hi) hi = te[k];
    }
    if (lo < 0 || hi <= lo) return;
    double* buf = (double*)calloc((size_t)(hi - lo) * 12, sizeof(double));
    /* read targets' positions before any write: buffers are added at the end */
    for (int64_t k = 0; k < npairs; k++)
        for (int64_t i = tb[k]; i < te[k]; i++) {
            const double* tp = P + i * stride;
            for (int64_t j = sb[k]; j < se[k]; j++)
                pair_uj(kernel, tp[RX], tp[RX + 1], tp[RX + 2], P + j * stride, buf + (i - lo) * 12);
        }
    for (int64_t i = lo; i < hi; i++) add_acc(P + i * stride, buf + (i - lo) * 12);
    free(buf);
}

static int is_static(const double* p) { return p[RSTATIC] != 0; }

void vpmo_reset_particles(double* P, int64_t np)
{
    for (int64_t i = 0; i < np; i++) {
        double* p = P + i * VPMO_NFIELDS;
        if (p[RSTATIC] == 0)                                   /* particlefield.jl:468 */
            for (int k = RU; k < RPSE + 3; k++) p[k] = 0.0;    /* U, vorticity, J, PSE */
    }
}

void vpmo_reset_particles_sfs(double* P, int64_t np)
{
    for (int64_t i = 0; i < np; i++) {
        double* p = P + i * VPMO_NFIELDS;
        if (p[RSTATIC] == 0)
            for (int k = 0; k < 3; k++) p[RSFS + k] = 0.0;
    }
}

/* subfilterscale_models.jl:16-41 for one (target, source) pair */
static inline void estr_pair(int kernel, int transposed, double* tp, const double* sp, double r)
{
    const double* GS = sp + RG;
    const double* JS = sp + RJ;
    const double* JT = tp + RJ;
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
    double sigma_inv = 1.0 / sp[RSIG];
    double zeta_sgm = vpmo_zeta(kernel, r * sigma_inv) * sigma_inv * sigma_inv * sigma_inv;
    tp[RSFS + 0] += zeta_sgm * S1;
    tp[RSFS + 1] += zeta_sgm * S2;
    tp[RSFS + 2] += zeta_sgm * S3;
}

void vpmo_estr_direct(double* P, int64_t np, int kernel, int transposed, int nthreads)
{
    vpmo_estr_direct_targets(P, np, 0, np, kernel, transposed, nthreads);
}

void vpmo_estr_direct_targets(double* P, int64_t np, int64_t t0, int64_t t1, int kernel, int transposed,
                              int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    /* The SFS rows of a target are only written by that target's iteration,
     * and J/Gamma/X are only read, so targets are independent. */
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = t0; i < t1; i++) {
        double* tp = P + i * VPMO_NFIELDS;
        if (is_static(tp)) continue;                                /* :63 */
        const double tx = tp[RX], ty = tp[RX + 1], tz = tp[RX + 2];
        for (int64_t j = 0; j < np; j++) {
            const double* sp = P + j * VPMO_NFIELDS;
            if (is_static(sp)) continue;                            /* iterator() */
            double dx = sp[RX] - tx, dy = sp[RX + 1] - ty, dz = sp[RX + 2] - tz;
            double r = sqrt(dx * dx + dy * dy + dz * dz);
            estr_pair(kernel, transposed, tp, sp, r);
        }
    }
}

void vpmo_uj_direct(double* P, int64_t np, int kernel, int reset, int reset_sfs,
                    int sfs, int transposed, int nthreads)
{
    if (reset) vpmo_reset_particles(P, np);                 /* UJ.jl:28-30 */
    if (reset_sfs) vpmo_reset_particles_sfs(P, np);         /* UJ.jl:31-33 */
    vpmo_direct2(P, VPMO_NFIELDS, 0, np, P, VPMO_NFIELDS, 0, np, kernel, nthreads);
    if (sfs) vpmo_estr_direct(P, np, kernel, transposed, nthreads);   /* UJ.jl:36 */
}

/* ---- per-particle schemes ---- */
static inline double cube(double x) { return x * x * x; }

static inline void stretching(const double* J, const double* G, int transposed, double S[3])
{
    if (transposed) {                                   /* timeintegration.jl:489-493 */
        S[0] = J[0] * G[0] + J[1] * G[1] + J[2] * G[2];
        S[1] = J[3] * G[0] + J[4] * G[1] + J[5] * G[2];
        S[2] = J[6] * G[0] + J[7] * G[1] + J[8] * G[2];
    } else {                                            /* :494-499 */
        S[0] = J[0] * G[0] + J[3] * G[1] + J[6] * G[2];
        S[1] = J[1] * G[0] + J[4] * G[1] + J[7] * G[2];
        S[2] = J[2] * G[0] + J[5] * G[1] + J[8] * G[2];
    }
}

void vpmo_update_particle_states(double* P, int64_t np, double a, double b, double dt,
                                 const double Uinf[3], double f, double g,
                                 double zeta0, int transposed)
{
    for (int64_t i = 0; i < np; i++) {
        double* p = P + i * VPMO_NFIELDS;
        if (is_static(p)) continue;                                /* :472 */
        double C = p[RC];
        double* M = p + RM;   /* M[k-1] here is M[k] in the reference */
        double* G = p + RG;
        const double* J = p + RJ;
        M[0] = a * M[0] + dt * (p[RU + 0] + Uinf[0]);
        M[1] = a * M[1] + dt * (p[RU + 1] + Uinf[1]);
        M[2] = a * M[2] + dt * (p[RU + 2] + Uinf[2]);
        p[RX + 0] += b * M[0];
        p[RX + 1] += b * M[1];
        p[RX + 2] += b * M[2];
        double MM[3];
        stretching(J, G, transposed, MM);
        double Gnorm2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
        double MM4;
        double sgm = p[RSIG];
        if (Gnorm2 > 0) {                                         /* :503-509 */
            MM4 = (f + g) / (1 + 3 * f) * (MM[0] * G[0] + MM[1] * G[1] + MM[2] * G[2]);
            MM4 -= f / (1 + 3 * f) * (C * p[RSFS] * G[0] + C * p[RSFS + 1] * G[1] + C * p[RSFS + 2] * G[2]) * cube(sgm) / zeta0;
            MM4 /= Gnorm2;
        } else {
            MM4 = 0.0;
        }
        M[3] = a * M[3] + dt * (MM[0] - 3 * MM4 * G[0] - C * p[RSFS + 0] * cube(sgm) / zeta0);
        M[4] = a * M[4] + dt * (MM[1] - 3 * MM4 * G[1] - C * p[RSFS + 1] * cube(sgm) / zeta0);
        M[5] = a * M[5] + dt * (MM[2] - 3 * MM4 * G[2] - C * p[RSFS + 2] * cube(sgm) / zeta0);
        M[7] = a * M[7] - dt * (sgm * MM4);                       /* :518 */
        G[0] += b * M[3];
        G[1] += b * M[4];
        G[2] += b * M[5];
        p[RSIG] += b * M[7];                                      /* :526 */
    }
}

void vpmo_relax(double* P, int64_t np, double rlxf, int corrected)
{
    for (int64_t i = 0; i < np; i++) {
        double* p = P + i * VPMO_NFIELDS;
        if (is_static(p)) continue;                   /* timeintegration.jl:447 */
        const double* J = p + RJ;
        double* G = p + RG;
        double nrmw = sqrt((J[5] - J[7]) * (J[5] - J[7]) +
                           (J[6] - J[2]) * (J[6] - J[2]) +
                           (J[1] - J[3]) * (J[1] - J[3]));
        if (nrmw != 0) {
            double nrmGamma = sqrt(G[0] * G[0] + G[1] * G[1] + G[2] * G[2]);
            double b2 = 0;
            if (corrected)                              /* relaxation.jl:129-131 */
                b2 = 1 - 2 * (1 - rlxf) * rlxf * (1 - (G[0] * (J[5] - J[7]) +
                                                     G[1] * (J[6] - J[2]) +
                                                     G[2] * (J[1] - J[3])) / (nrmGamma * nrmw));
            G[0] = (1 - rlxf) * G[0] + rlxf * nrmGamma * (J[5] - J[7]) / nrmw;
            G[1] = (1 - rlxf) * G[1] + rlxf * nrmGamma * (J[6] - J[2]) / nrmw;
            G[2] = (1 - rlxf) * G[2] + rlxf * nrmGamma * (J[1] - J[3]) / nrmw;
            if (corrected) {
                double sb2 = sqrt(b2);
                G[0] /= sb2; G[1] /= sb2; G[2] /= sb2;
            }
        }
    }
}

int vpmo_core_spreading_rk3(double* P, int64_t np, double nu, double sgm0, double beta,
                            double* t_sgm, double a, double b, double dt)
{
    for (int64_t i = 0; i < np; i++) {
        double* p = P + i * VPMO_NFIELDS;
        if (is_static(p)) continue;                         /* iterator() */
        p[RM + 6] = a * p[RM + 6] + dt * 2 * nu;            /* viscous.jl:173 */
        p[RSIG] = sqrt(p[RSIG] * p[RSIG] + b * p[RM + 6]);  /* viscous.jl:174 */
    }
    if (fabs(b - 8.0 / 15.0) <= 1e-7) {                     /* viscous.jl:178 */
        *t_sgm += dt;                                       /* :191 */
        double beta_cur = sqrt(2 * nu * (*t_sgm) / (sgm0 * sgm0) + 1);
        if (beta_cur >= beta) return 1;                     /* :202 */
    }
    return 0;
}

static int check_bounds(double minC, double maxC)
{
    if (minC < 0 || maxC < 0 || minC > maxC) return -1;     /* subfilterscale.jl:455-461 */
    return 0;
}

int vpmo_sfs_before(double* P, int64_t np, int kernel, int transposed, double alpha,
                    double minC, double maxC, int nthreads)
{
    if (check_bounds(minC, maxC)) return -1;
    for (int64_t i = 0; i < np; i++) {                      /* :465-475 */
        double* p = P + i * VPMO_NFIELDS;
        if (is_static(p)) continue;
        p[RSIG] *= alpha;
    }
    vpmo_uj_direct(P, np, kernel, 1, 1, 1, transposed, nthreads);   /* :478 */
    for (int64_t i = 0; i < np; i++) {                      /* :481-521 */
        double* p = P + i * VPMO_NFIELDS;
        if (is_static(p)) continue;
        for (int k = 0; k < 9; k++) p[RM + k] = 0.0;
        double S[3];
        stretching(p + RJ, p + RG, transposed, S);
        p[RM + 0] = S[0]; p[RM + 1] = S[1]; p[RM + 2] = S[2];
        p[RM + 3] = p[RSFS]; p[RM + 4] = p[RSFS + 1]; p[RM + 5] = p[RSFS + 2];
    }
    for (int64_t i = 0; i < np; i++) {                      /* :526-536 */
        double* p = P + i * VPMO_NFIELDS;
        if (is_static(p)) continue;
        p[RSIG] /= alpha;
    }
    return 0;
}

int vpmo_sfs_after(double* P, int64_t np, int kernel, int transposed, double alpha,
                   double rlxf, double minC, double maxC, int force_positive, int clip)
{
    if (check_bounds(minC, maxC)) return -1;
    for (int64_t i = 0; i < np; i++) {                      /* :559-586 */
        double* p = P + i * VPMO_NFIELDS;
        if (is_static(p)) continue;
        double* M = p + RM;
        const double* J = p + RJ;
        const double* G = p + RG;
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
    }
    const double zeta0 = vpmo_zeta(kernel, 0.0);            /* :590 */
    int status = 0;
    for (int64_t i = 0; i < np; i++) {                      /* :592-656 */
        double* p = P + i * VPMO_NFIELDS;
        if (is_static(p)) continue;
        const double* M = p + RM;
        double* C = p + RC;
        const double* G = p + RG;
        double nume = M[0] * G[0] + M[1] * G[1] + M[2] * G[2];
        nume *= 3 * alpha - 2;
        double deno = M[3] * G[0] + M[4] * G[1] + M[5] * G[2];
        deno /= zeta0 / cube(p[RSIG]);
        if (C[2] == 0) {
            C[2] = deno;
            if (C[2] == 0) C[2] = 2.220446049250313e-16;    /* eps() */
        }
        nume = rlxf * nume + (1 - rlxf) * C[1];
        deno = rlxf * deno + (1 - rlxf) * C[2];
        if (fabs(nume / deno) > maxC) {
            if (fabs(deno) < fabs(C[2])) deno = jl_sign(deno) * fabs(C[2]);
            if (fabs(nume / deno) >= maxC) nume = jl_sign(nume) * fabs(deno) * maxC;
        } else if (fabs(nume / deno) < minC) {
            nume = jl_sign(nume) * fabs(deno) * minC;
        }
        C[1] = nume;
        C[2] = deno;
        C[0] = C[1] / C[2];
        if (isnan(C[0])) status = -2;                       /* :645-652 */
        if (force_positive) C[0] *= jl_sign(C[0]);          /* :655 */
    }
    for (int64_t i = 0; i < np; i++) {                      /* :659-670 */
        double* p = P + i * VPMO_NFIELDS;
        if (is_static(p)) continue;
        for (int k = 0; k < 9; k++) p[RM + k] = 0.0;
    }
    if (status) return status;
    if (clip) {                                             /* :225-247, 287-296 */
        for (int64_t i = 0; i < np; i++) {
            double* p = P + i * VPMO_NFIELDS;
            if (is_static(p)) continue;
            double Cc = p[RC];
            if (Cc * (p[RG] * p[RSFS] + p[RG + 1] * p[RSFS + 1] + p[RG + 2] * p[RSFS + 2]) < 0)
                p[RC] *= 0;
        }
    }
    return 0;
}

int vpmo_rk3_step(double* P, int64_t np, vpmo_settings* s, double dt, int relax_now)
{
    const double zeta0 = vpmo_zeta(s->kernel, 0.0);                 /* :398 */
    const int sfs_on = s->sfs != 0;
    int status = 0;
    for (int64_t i = 0; i < np; i++) {                              /* :401-414 */
        double* p = P + i * VPMO_NFIELDS;
        if (p[RSTATIC] == 0) for (int k = 0; k < 9; k++) p[RM + k] = 0.0;
    }
    const double ab[3][2] = { {0.0, 1.0 / 3.0}, {-5.0 / 9.0, 15.0 / 16.0},
                              {-153.0 / 128.0, 8.0 / 15.0} };       /* :417 */
    for (int st = 0; st < 3; st++) {
        double a = ab[st][0], b = ab[st][1];
        int first = (a == 1 || a == 0);                             /* subfilterscale.jl:207 */
        if (sfs_on && first) {
            int rc = vpmo_sfs_before(P, np, s->kernel, s->transposed, s->sfs_alpha,
                                     s->sfs_minC, s->sfs_maxC, s->nthreads);
            if (rc) return rc;
        }
        vpmo_uj_direct(P, np, s->kernel, 1, sfs_on, sfs_on, s->transposed, s->nthreads);  /* :422 */
        if (sfs_on && first) {
            int rc = vpmo_sfs_after(P, np, s->kernel, s->transposed, s->sfs_alpha, s->sfs_rlxf,
                                    s->sfs_minC, s->sfs_maxC, s->sfs_force_positive, s->sfs_clip);
            if (rc) return rc;
        }
        vpmo_update_particle_states(P, np, a, b, dt, s->Uinf, s->f, s->g, zeta0, s->transposed);
        if (s->viscous == 1) {
            if (vpmo_core_spreading_rk3(P, np, s->nu, s->sgm0, s->beta, &s->t_sgm, a, b, dt)) {
                if (s->rbf_itmax > 0) {                             /* viscous.jl:202-220 */
                    int its;
                    double res[3];
                    if (vpmo_core_spreading_reset(P, np, s->kernel, s->sgm0, s->rbf_itmax, s->rbf_tol,
                                                  s->rbf_iterror, s->nthreads, &its, res))
                        status = 2;
                    s->t_sgm = 0;                                   /* :218 */
                } else {
                    status = 1;
                }
            }
        }
    }
    if (relax_now && s->relax) {                                    /* :437-458 */
        vpmo_reset_particles(P, np);
        vpmo_uj_direct(P, np, s->kernel, 1, 0, 0, s->transposed, s->nthreads);
        vpmo_relax(P, np, s->rlxf, s->relax == 2);
    }
    /* nextstep: U_prev = |U| for every particle (particlefield.jl:443-455) */
    for (int64_t i = 0; i < np; i++) {
        double* p = P + i * VPMO_NFIELDS;
        p[RUPREV] = sqrt(p[RU] * p[RU] + p[RU + 1] * p[RU + 1] + p[RU + 2] * p[RU + 2]);
    }
    return status;
}

/* ---- adjoint ----
 * Pair contribution  l = F (A + C) + H B  with d = x_i - x_j, v = d x Gamma_j,
 * A = ubar . v, B = v . (Jbar d), C = tau . Gamma_j,
 * tau = (Jbar_23 - Jbar_32, Jbar_31 - Jbar_13, Jbar_12 - Jbar_21),
 * F = -c g/r^3, H = F'(r)/r, K = H'(r)/r, F_s = dF/dsigma, H_s = dH/dsigma.
 * (Derived from the forward expressions of fmm.jl:125-161.) */
typedef struct { double F, H, K, Fs, Hs; } pair_coef;

static inline pair_coef coefs(int kernel, double r, double sigma)
{
    const double c = const4();
    double q = r / sigma, g, dg, d2g;
    vpmo_g_derivs(kernel, q, &g, &dg, &d2g);
    double r2 = r * r, r3 = r2 * r, r5 = r3 * r2, r7 = r5 * r2;
    pair_coef k;
    k.F = -c * g / r3;
    k.H = -c * (q * dg - 3 * g) / r5;
    k.K = -c * (q * q * d2g - 7 * q * dg + 15 * g) / r7;
    k.Fs = c * dg * q / (sigma * r3);
    k.Hs = c * q * (q * d2g - 2 * dg) / (sigma * r5);
    return k;
}

static inline void cross3(const double* a, const double* b, double* o)
{
    o[0] = a[1] * b[2] - a[2] * b[1];
    o[1] = a[2] * b[0] - a[0] * b[2];
    o[2] = a[0] * b[1] - a[1] * b[0];
}

/* Jbar flat f = 3b + a is the cotangent of J[a][b]. */
static inline void pair_adjoint(int kernel, const double* ti, const double* sj,
                                const double* ub, const double* Jb,
                                double gd[3], double gG[3], double* gs)
{
    double d[3] = { ti[RX] - sj[RX], ti[RX + 1] - sj[RX + 1], ti[RX + 2] - sj[RX + 2] };
    double r2 = d[0] * d[0] + d[1] * d[1] + d[2] * d[2];
    gd[0] = gd[1] = gd[2] = gG[0] = gG[1] = gG[2] = 0; *gs = 0;
    if (r2 == 0) return;
    double r = sqrt(r2);
    const double* G = sj + RG;
    pair_coef k = coefs(kernel, r, sj[RSIG]);
    double v[3]; cross3(d, G, v);
    double w[3], JTv[3];
    for (int a = 0; a < 3; a++) {
        w[a] = Jb[a] * d[0] + Jb[3 + a] * d[1] + Jb[6 + a] * d[2];
        JTv[a] = Jb[3 * a + 0] * v[0] + Jb[3 * a + 1] * v[1] + Jb[3 * a + 2] * v[2];
    }
    double tau[3] = { Jb[3 * 2 + 1] - Jb[3 * 1 + 2], Jb[3 * 0 + 2] - Jb[3 * 2 + 0],
                      Jb[3 * 1 + 0] - Jb[3 * 0 + 1] };
    double A = ub[0] * v[0] + ub[1] * v[1] + ub[2] * v[2];
    double B = v[0] * w[0] + v[1] * w[1] + v[2] * w[2];
    double C = tau[0] * G[0] + tau[1] * G[1] + tau[2] * G[2];
    double Gxu[3], Gxw[3], uxd[3], wxd[3];
    cross3(G, ub, Gxu); cross3(G, w, Gxw); cross3(ub, d, uxd); cross3(w, d, wxd);
    double s = k.H * (A + C) + k.K * B;
    for (int a = 0; a < 3; a++) {
        gd[a] = s * d[a] + k.F * Gxu[a] + k.H * (Gxw[a] + JTv[a]);
        gG[a] = k.F * (uxd[a] + tau[a]) + k.H * wxd[a];
    }
    *gs = k.Fs * (A + C) + k.Hs * B;
}

void vpmo_uj_adjoint(const double* P, int64_t np, int kernel,
                     const double* Ubar, const double* Jbar,
                     double* Xbar, double* Gbar, double* Sbar, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    /* target role: d(l)/d(x_i) = +gd */
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = 0; i < np; i++) {
        const double* ti = P + i * VPMO_NFIELDS;
        double ax[3] = {0, 0, 0};
        for (int64_t j = 0; j < np; j++) {
            double gd[3], gG[3], gs;
            pair_adjoint(kernel, ti, P + j * VPMO_NFIELDS, Ubar + 3 * i, Jbar + 9 * i, gd, gG, &gs);
            ax[0] += gd[0]; ax[1] += gd[1]; ax[2] += gd[2];
        }
        Xbar[3 * i] = ax[0]; Xbar[3 * i + 1] = ax[1]; Xbar[3 * i + 2] = ax[2];
    }
    /* source role: d(l)/d(x_j) = -gd, plus Gamma_j and sigma_j */
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t j = 0; j < np; j++) {
        const double* sj = P + j * VPMO_NFIELDS;
        double ax[3] = {0, 0, 0}, aG[3] = {0, 0, 0}, as = 0;
        for (int64_t i = 0; i < np; i++) {
            double gd[3], gG[3], gs;
            pair_adjoint(kernel, P + i * VPMO_NFIELDS, sj, Ubar + 3 * i, Jbar + 9 * i, gd, gG, &gs);
            for (int a = 0; a < 3; a++) { ax[a] -= gd[a]; aG[a] += gG[a]; }
            as += gs;
        }
        for (int a = 0; a < 3; a++) { Xbar[3 * j + a] += ax[a]; Gbar[3 * j + a] = aG[a]; }
        Sbar[j] = as;
    }
}

/* Summation scale of vpmo_uj_adjoint: the same sums with every pair
 * contribution replaced by its absolute value (componentwise).  An FP32 pair
 * evaluation is accurate to a few units of FP32 rounding per term, so its
 * error in an output is bounded by a small multiple of this scale, however
 * much the terms cancel. */
void vpmo_uj_adjoint_scale(const double* P, int64_t np, int kernel,
                           const double* Ubar, const double* Jbar,
                           double* Xmag, double* Gmag, double* Smag, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = 0; i < np; i++) {
        const double* ti = P + i * VPMO_NFIELDS;
        double ax[3] = {0, 0, 0};
        for (int64_t j = 0; j < np; j++) {
            double gd[3], gG[3], gs;
            pair_adjoint(kernel, ti, P + j * VPMO_NFIELDS, Ubar + 3 * i, Jbar + 9 * i, gd, gG, &gs);
            for (int a = 0; a < 3; a++) ax[a] += fabs(gd[a]);
        }
        for (int a = 0; a < 3; a++) Xmag[3 * i + a] = ax[a];
    }
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t j = 0; j < np; j++) {
        const double* sj = P + j * VPMO_NFIELDS;
        double ax[3] = {0, 0, 0}, aG[3] = {0, 0, 0}, as = 0;
        for (int64_t i = 0; i < np; i++) {
            double gd[3], gG[3], gs;
            pair_adjoint(kernel, P + i * VPMO_NFIELDS, sj, Ubar + 3 * i, Jbar + 9 * i, gd, gG, &gs);
            for (int a = 0; a < 3; a++) { ax[a] += fabs(gd[a]); aG[a] += fabs(gG[a]); }
            as += fabs(gs);
        }
        for (int a = 0; a < 3; a++) { Xmag[3 * j + a] += ax[a]; Gmag[3 * j + a] = aG[a]; }
        Smag[j] = as;
    }
}

static long double pair_loss(int kernel, const double* ti, const double* sj,
                             const double* ub, const double* Jb)
{
    double acc[12] = {0};
    pair_uj(kernel, ti[RX], ti[RX + 1], ti[RX + 2], sj, acc);
    long double l = 0;
    for (int a = 0; a < 3; a++) l += (long double)ub[a] * acc[a];
    for (int f = 0; f < 9; f++) l += (long double)Jb[f] * acc[3 + f];
    return l;
}

double vpmo_loss_local(const double* P, int64_t np, int kernel,
                       const double* Ubar, const double* Jbar, int64_t k)
{
    long double L = 0;
    const double* pk = P + k * VPMO_NFIELDS;
    for (int64_t j = 0; j < np; j++)
        if (j != k) L += pair_loss(kernel, pk, P + j * VPMO_NFIELDS, Ubar + 3 * k, Jbar + 9 * k);
    for (int64_t i = 0; i < np; i++)
        if (i != k) L += pair_loss(kernel, P + i * VPMO_NFIELDS, pk, Ubar + 3 * i, Jbar + 9 * i);
    return (double)L;
}

/* ---- adjoint of the time step ----
 * Pullback of update_particle_states for ReformulatedVPM without SFS
 * (timeintegration.jl:467-528 with C S = 0; the pullback the reference writes
 * as state_pullback in rrules.jl:436-555, here derived from the update
 * itself).  With m1 = Mbar1' + b Xbar', m23 = Mbar23' + b sigmabar',
 * m2 = Mbar2' + b Gammabar', MM = stretching(J, Gamma), G2 = |Gamma|^2,
 * MM4 = A (MM . Gamma)/G2, A = (f + g)/(1 + 3 f):
 *   Xbar = Xbar',  Mbar1 = a m1,  Ubar = dt m1,
 *   Mbar23 = a m23,  sigmabar = sigmabar' - dt MM4 m23,
 *   Mbar2 = a m2,  MMbar = dt m2 + MM4bar A Gamma/G2,
 *   MM4bar = -dt sigma m23 - 3 dt (Gamma . m2),
 *   Gammabar = Gammabar' - 3 dt MM4 m2 + MM4bar (A MM - 2 MM4 Gamma)/G2
 *              + dMM/dGamma^T MMbar,   Jbar = dMM/dJ^T MMbar.
 * (rrules.jl:521 keeps only the eta = zeta term of d(MM4 Gamma_zeta)/dJ;
 * the full sum over zeta is used here.)
 * B[i] is the cotangent record of particle i (46 rows, same layout as P):
 * on entry the rows X, Gamma, sigma, M of the updated state; on exit those
 * rows for the state before the update plus Ubar in the U rows and Jbar in
 * the J rows.  P holds the state before the update (with its U and J). */
void vpmo_update_particle_states_adjoint(const double* P, double* B, int64_t np, double a, double b,
                                         double dt, double f, double g, int transposed)
{
    const double A = (f + g) / (1 + 3 * f);
    for (int64_t i = 0; i < np; i++) {
        const double* p = P + i * VPMO_NFIELDS;
        double* q = B + i * VPMO_NFIELDS;
        for (int k = 0; k < 3; k++) q[RU + k] = 0.0;
        for (int k = 0; k < 9; k++) q[RJ + k] = 0.0;
        if (is_static(p)) continue;
        const double* G = p + RG;
        const double* J = p + RJ;
        const double sgm = p[RSIG];
        double MM[3];
        stretching(J, G, transposed, MM);
        const double G2 = G[0] * G[0] + G[1] * G[1] + G[2] * G[2];
        const double MM4 = G2 > 0 ? A * (MM[0] * G[0] + MM[1] * G[1] + MM[2] * G[2]) / G2 : 0.0;
        double m1[3], m2[3];
        for (int k = 0; k < 3; k++) {
            m1[k] = q[RM + k] + b * q[RX + k];
            m2[k] = q[RM + 3 + k] + b * q[RG + k];
        }
        const double m23 = q[RM + 7] + b * q[RSIG];
        double MM4b = -dt * sgm * m23 - 3 * dt * (G[0] * m2[0] + G[1] * m2[1] + G[2] * m2[2]);
        double MMb[3], Gb[3];
        for (int k = 0; k < 3; k++) {
            q[RM + k] = a * m1[k];
            q[RU + k] = dt * m1[k];
            q[RM + 3 + k] = a * m2[k];
            MMb[k] = dt * m2[k];
            Gb[k] = q[RG + k] - 3 * dt * MM4 * m2[k];
        }
        q[RM + 7] = a * m23;
        q[RSIG] = q[RSIG] - dt * MM4 * m23;
        if (G2 > 0) {
            for (int k = 0; k < 3; k++) {
                MMb[k] += MM4b * A * G[k] / G2;
                Gb[k] += MM4b * (A * MM[k] - 2 * MM4 * G[k]) / G2;
            }
        }
        /* MM[e] = sum_x J(x, e) G[x] (transposed) or J(e, x) G[x]; J(r, c) = J[3 c + r] */
        for (int e = 0; e < 3; e++)
            for (int x = 0; x < 3; x++) {
                const int idx = transposed ? 3 * e + x : 3 * x + e;
                q[RJ + idx] += MMb[e] * G[x];
                Gb[x] += J[idx] * MMb[e];
            }
        for (int k = 0; k < 3; k++) q[RG + k] = Gb[k];
    }
}

/* ---- the reference's own update pullback: state_pullback (rrules.jl:436-547)
 * A literal restatement (transposed scheme, the only one it covers) for the
 * record layout of vpmo_update_particle_states_adjoint: B holds on entry the
 * cotangents of the updated X, Gamma, sigma and registers M1 = M[0:3],
 * M2 = M[3:6], M23 = M[7], and on exit those of the state before the update,
 * with Ubar in the U rows and Jbar in the J rows.  C and SFS come from P.
 * As written, the reference differentiates the -3 MM4 Gamma_zeta term of the
 * Gamma register only for zeta = eta, in Jbar (:521) and in Gammabar (:524);
 * full != 0 uses the whole sum over zeta there instead.  Its Gammabar also
 * reads M2outbar[xi] of the first particle (:526); that is kept, so calls
 * with np > 1 reproduce it.  Cross-check only. */
void vpmo_update_particle_states_pullback_rrules(const double* P, double* B, int64_t np, double a, double b,
                                                 double dt, double f, double g, double zeta0, int full)
{
    const double A = (f + g) / (1 + 3 * f);
    const double Bc = f / (1 + 3 * f);
    const double B2 = Bc / zeta0;
    const double M2out0[3] = {B[RM + 3], B[RM + 4], B[RM + 5]};   /* M2outbar[1:3] (particle 1) */
    for (int64_t i = 0; i < np; i++) {
        const double* p = P + i * VPMO_NFIELDS;
        double* q = B + i * VPMO_NFIELDS;
        const double* G = p + RG;
        const double* Jv = p + RJ;                 /* J_mat[xi][eta] = Jv[xi + 3 eta] */
        const double sg = p[RSIG];
        const double Ci = p[RC];
        const double* S = p + RSFS;
        double M1out[3], xout[3], M2out[3], Gout[3];
        for (int k = 0; k < 3; k++) {
            M1out[k] = q[RM + k]; xout[k] = q[RX + k]; M2out[k] = q[RM + 3 + k]; Gout[k] = q[RG + k];
        }
        const double M23out = q[RM + 7], sout = q[RSIG];
        double MM[4] = {0, 0, 0, 0}, G2 = 0.0, dMM4dJ[3][3];
        for (int e = 0; e < 3; e++) {                                      /* :484-492 */
            G2 += G[e] * G[e];
            for (int x = 0; x < 3; x++) MM[e] += Jv[x + 3 * e] * G[x];
        }
        for (int e = 0; e < 3; e++) {                                      /* :493-498 */
            MM[3] += (A * MM[e] * G[e] - B2 * Ci * S[e] * G[e] * sg * sg * sg) / G2;
            for (int x = 0; x < 3; x++) dMM4dJ[x][e] = 1 / G2 * A * G[x] * G[e];
        }
        double dMM4dC = 0.0, dMM4dsigma = 0.0;
        const double sigmabar_term = M23out + b * sout;                    /* :501-502 */
        double gdot = 0.0;                                                 /* sum_zeta Gamma_zeta m2_zeta */
        for (int e = 0; e < 3; e++) gdot += G[e] * (M2out[e] + b * Gout[e]);
        double Gin[3], Jb[9], M1in[3], Ub[3], M2in[3], Sb[3], sin_ = 0.0;
        for (int e = 0; e < 3; e++) {                                      /* :503-533 */
            double dMM4dGamma = -B2 * Ci * S[e] * sg * sg * sg - 2 * MM[3] * G[e] + A * MM[e];
            for (int x = 0; x < 3; x++) dMM4dGamma += A * Jv[e + 3 * x] * G[x];
            dMM4dGamma /= G2;
            dMM4dC += -B2 * S[e] * G[e] * sg * sg * sg / G2;
            const double dMM4dS = -B2 * Ci * G[e] * sg * sg * sg / G2;
            dMM4dsigma += -3 * B2 * Ci * S[e] * G[e] * sg * sg / G2;
            const double xbar_term = M1out[e] + b * xout[e];
            M1in[e] = a * xbar_term;
            Ub[e] = dt * xbar_term;
            const double gammabar_term = M2out[e] + b * Gout[e];
            M2in[e] = a * gammabar_term;
            const double gm = full ? gdot : G[e] * gammabar_term;          /* Gamma_eta m2_eta (:521, :524) */
            for (int x = 0; x < 3; x++)
                Jb[x + 3 * e] = dt * G[x] * gammabar_term - 3 * dt * dMM4dJ[x][e] * gm
                                - dt * sg * dMM4dJ[x][e] * sigmabar_term;
            sin_ -= 3 * dt * (G[e] * dMM4dsigma + Ci * S[e] * sg * sg / zeta0) * gammabar_term;
            Gin[e] = Gout[e] - 3 * dt * (dMM4dGamma * gm + MM[3] * gammabar_term)
                     - dt * sg * dMM4dGamma * sigmabar_term;
            for (int x = 0; x < 3; x++) Gin[e] += dt * Jv[e + 3 * x] * (M2out0[x] + b * Gout[x]);
            Sb[e] = -dt * (3 * G[e] * dMM4dS + Ci * sg * sg * sg / zeta0) * gammabar_term
                    - dt * sg * dMM4dS * sigmabar_term;
        }
        (void)sin_;                                                        /* overwritten at :535 */
        const double sin_final = sout - dt * (MM[3] + sg * dMM4dsigma) * sigmabar_term;
        (void)Sb;
        for (int k = 0; k < 3; k++) {
            q[RX + k] = xout[k];                                           /* :480 */
            q[RG + k] = Gin[k];
            q[RM + k] = M1in[k];
            q[RM + 3 + k] = M2in[k];
            q[RU + k] = Ub[k];
        }
        for (int k = 0; k < 9; k++) q[RJ + k] = Jb[k];
        q[RM + 7] = a * sigmabar_term;
        q[RSIG] = sin_final;
    }
}

/* Pullback of one vpmo_rk3_step (no SFS, inviscid, no relaxation): B holds
 * the cotangent records of the state after the step and is replaced by those
 * of the state before it (rows X, Gamma, sigma; M rows are zero because the
 * step starts from M = 0).  The forward step is recomputed from P0 with the
 * three stage states (after each stage's UJ) kept in memory.  Returns -1 for
 * a setting whose adjoint is not restated. */
int vpmo_rk3_step_adjoint(const double* P0, int64_t np, const vpmo_settings* s, double dt, double* B)
{
    if (s->sfs != 0 || s->viscous != 0) return -1;
    const size_t rec = (size_t)np * VPMO_NFIELDS;
    double* tape = (double*)malloc(sizeof(double) * rec * 4);
    if (!tape) return -2;
    double* P = tape + 3 * rec;
    memcpy(P, P0, sizeof(double) * rec);
    const double zeta0 = vpmo_zeta(s->kernel, 0.0);
    for (int64_t i = 0; i < np; i++) {
        double* p = P + i * VPMO_NFIELDS;
        if (p[RSTATIC] == 0) for (int k = 0; k < 9; k++) p[RM + k] = 0.0;
    }
    const double ab[3][2] = { {0.0, 1.0 / 3.0}, {-5.0 / 9.0, 15.0 / 16.0}, {-153.0 / 128.0, 8.0 / 15.0} };
    for (int st = 0; st < 3; st++) {
        vpmo_uj_direct(P, np, s->kernel, 1, 0, 0, s->transposed, s->nthreads);
        memcpy(tape + st * rec, P, sizeof(double) * rec);
        vpmo_update_particle_states(P, np, ab[st][0], ab[st][1], dt, s->Uinf, s->f, s->g, zeta0, s->transposed);
    }
    double* Ub = (double*)malloc(sizeof(double) * (size_t)np * 19);
    if (!Ub) { free(tape); return -2; }
    double* Jb = Ub + 3 * np;
    double* Xb = Jb + 9 * np;
    double* Gb = Xb + 3 * np;
    double* Sb = Gb + 3 * np;
    for (int st = 2; st >= 0; st--) {
        const double* S = tape + st * rec;
        vpmo_update_particle_states_adjoint(S, B, np, ab[st][0], ab[st][1], dt, s->f, s->g, s->transposed);
        for (int64_t i = 0; i < np; i++) {
            const double* q = B + i * VPMO_NFIELDS;
            for (int k = 0; k < 3; k++) Ub[3 * i + k] = q[RU + k];
            for (int k = 0; k < 9; k++) Jb[9 * i + k] = q[RJ + k];
        }
        vpmo_uj_adjoint(S, np, s->kernel, Ub, Jb, Xb, Gb, Sb, s->nthreads);
        for (int64_t i = 0; i < np; i++) {
            double* q = B + i * VPMO_NFIELDS;
            for (int k = 0; k < 3; k++) {
                q[RX + k] += Xb[3 * i + k];
                q[RG + k] += Gb[3 * i + k];
            }
            q[RSIG] += Sb[i];
        }
    }
    for (int64_t i = 0; i < np; i++) {
        double* q = B + i * VPMO_NFIELDS;
        for (int k = 0; k < 3; k++) q[RU + k] = 0.0;
        for (int k = 0; k < 9; k++) q[RJ + k] = 0.0;
        if (P0[i * VPMO_NFIELDS + RSTATIC] == 0) for (int k = 0; k < 9; k++) q[RM + k] = 0.0;
    }
    free(Ub);
    free(tape);
    return 0;
}

/* ---- zeta_direct (viscous.jl:486-515) ----
 * J[1:3] of every particle (static included) is reset, then every target i
 * gets sum over every source j (i itself included: no r == 0 skip) of
 * Gamma_j zeta(r/sigma_j)/sigma_j^3. */
void vpmo_zeta_direct(double* P, int64_t np, int kernel, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
    for (int64_t i = 0; i < np; i++)                                 /* :487-489 */
        for (int k = 0; k < 3; k++) P[i * VPMO_NFIELDS + RJ + k] = 0.0;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = 0; i < np; i++) {
        double* pi = P + i * VPMO_NFIELDS;
        double acc[3] = {0, 0, 0};
        for (int64_t j = 0; j < np; j++) {
            const double* pj = P + j * VPMO_NFIELDS;
            const double dX1 = pi[RX] - pj[RX];                      /* :500-502 */
            const double dX2 = pi[RX + 1] - pj[RX + 1];
            const double dX3 = pi[RX + 2] - pj[RX + 2];
            const double r = sqrt(dX1 * dX1 + dX2 * dX2 + dX3 * dX3);
            const double s = pj[RSIG];
            const double zeta_sgm = 1 / (s * s * s) * vpmo_zeta(kernel, r / s);   /* :505 */
            acc[0] += pj[RG + 0] * zeta_sgm;                        /* :507-509 */
            acc[1] += pj[RG + 1] * zeta_sgm;
            acc[2] += pj[RG + 2] * zeta_sgm;
        }
        for (int k = 0; k < 3; k++) pi[RJ + k] = acc[k];
    }
}

/* ---- the reference's own CPU gradient: UJS_pullback (rrules.jl:183-326) --
 * A literal restatement of the ChainRules pullback of fmm.direct!, for a
 * field acting on itself (the target and source cotangents of each particle
 * are summed).  Cross-check only: the reference marks parts of it as
 * inexact (rrules.jl:233, 237), so it is compared with vpmo_uj_adjoint (which
 * central differences pin) and its differences are reported, not forced.
 * eps3(a,x,y) = (x cross y)_a and lc(a,b,c) is the Levi-Civita symbol
 * (rrules.jl:22-26); J-bar of target i is read column-major, Jb(c,b) =
 * Jbar[9 i + c + 3 b] (0-based; rrules.jl:262). */
static double eps3(int a, const double* x, const double* y)
{
    if (a == 0) return x[1] * y[2] - x[2] * y[1];
    if (a == 1) return x[2] * y[0] - x[0] * y[2];
    return x[0] * y[1] - x[1] * y[0];
}
static double lc(int a, int b, int c)
{
    if (a == b || b == c || c == a) return 0.0;
    return ((b - a + 3) % 3 == 1) ? 1.0 : -1.0;
}

void vpmo_uj_adjoint_rrules(const double* P, int64_t np, int kernel,
                            const double* Ubar, const double* Jbar,
                            double* Xbar, double* Gbar, double* Sbar)
{
    const double c4 = 1 / (4 * PI_D);
    for (int64_t k = 0; k < 3 * np; k++) { Xbar[k] = 0.0; Gbar[k] = 0.0; }
    for (int64_t k = 0; k < np; k++) Sbar[k] = 0.0;
    /* contributions from U-bar (rrules.jl:213-243) */
    for (int64_t j = 0; j < np; j++) {
        const double* sj = P + j * VPMO_NFIELDS;
        const double* G = sj + RG;
        const double sg = sj[RSIG];
        for (int64_t i = 0; i < np; i++) {
            const double* ti = P + i * VPMO_NFIELDS;
            const double* Ub = Ubar + 3 * i;
            double dx[3];
            for (int e = 0; e < 3; e++) dx[e] = ti[RX + e] - sj[RX + e];
            const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            if (!(r > 0)) continue;
            double g, dg, d2g;
            vpmo_g_derivs(kernel, r / sg, &g, &dg, &d2g);
            const double A = g / (r * r * r);
            const double B = dg / (r * r * r * r * sg) - 3 * g / (r * r * r * r * r);
            for (int e = 0; e < 3; e++) {
                double x_term = -A * eps3(e, Ub, G);
                for (int a = 0; a < 3; a++) {
                    double GU[3] = {G[0] * Ub[a], G[1] * Ub[a], G[2] * Ub[a]};
                    x_term += B * dx[e] * eps3(a, dx, GU);
                }
                x_term *= -c4;
                Xbar[3 * i + e] += x_term;
                Xbar[3 * j + e] -= x_term;
            }
            for (int a = 0; a < 3; a++) {
                Sbar[j] += c4 * dg / (r * r * sg * sg) * eps3(a, dx, G) * Ub[a];
                Gbar[3 * j + a] += c4 * g / (r * r * r) * eps3(a, dx, Ub);
            }
        }
    }
    /* contributions from J-bar (rrules.jl:252-314) */
    for (int64_t i = 0; i < np; i++) {
        const double* ti = P + i * VPMO_NFIELDS;
        const double* Jb = Jbar + 9 * i;
        for (int64_t j = 0; j < np; j++) {
            const double* sj = P + j * VPMO_NFIELDS;
            const double* G = sj + RG;
            const double sg = sj[RSIG];
            double dx[3];
            for (int e = 0; e < 3; e++) dx[e] = ti[RX + e] - sj[RX + e];
            const double r = sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
            if (!(r > 0.0)) continue;
            double g, dg, ddg;
            vpmo_g_derivs(kernel, r / sg, &g, &dg, &ddg);
            const double alpha = dg / (sg * r) - 3 * g / (r * r);
            const double beta = -c4 * g / (r * r * r);
            double crss[3];
            for (int a = 0; a < 3; a++) crss[a] = -c4 / (r * r * r) * eps3(a, dx, G);
            for (int a = 0; a < 3; a++) {
                for (int b = 0; b < 3; b++) {
                    double sigma_term = 0.0;
                    for (int c = 0; c < 3; c++) {
                        sigma_term += c4 * dg * lc(a, b, c) * G[c] / (r * r);
                        double gamma_term = -beta * lc(a, b, c) * Jb[c + 3 * b];
                        double xyz_term = 0.0;
                        for (int d = 0; d < 3; d++) {
                            xyz_term += (dx[b] * lc(c, a, d) + dx[a] * lc(c, b, d)) * G[d];
                            gamma_term += alpha * c4 / (r * r * r) * lc(a, c, d) * dx[c] * dx[b] * Jb[d + 3 * b];
                        }
                        xyz_term *= -alpha * c4 / r;
                        xyz_term += (ddg / (sg * sg) - 7 * dg / (r * sg) + 15 * g / (r * r)) * crss[c] * dx[b] * dx[a];
                        xyz_term *= Jb[c + 3 * b] / (r * r);
                        Gbar[3 * j + a] += gamma_term;
                        Xbar[3 * i + a] += xyz_term;
                        Xbar[3 * j + a] -= xyz_term;
                    }
                    sigma_term += (-ddg / sg + 2 * dg / r) * crss[a] * dx[b];
                    Sbar[j] += Jb[a + 3 * b] / (sg * sg) * sigma_term;
                    Xbar[3 * i + a] += Jb[b + 3 * a] * alpha * crss[b];
                    Xbar[3 * j + a] -= Jb[b + 3 * a] * alpha * crss[b];
                }
            }
        }
    }
}

/* ---- zeta_fmm's pair loop over a direct list (viscous.jl:523-558) ----
 * For every (target range, source range) pair in list order, every target i
 * of the range gets, over the sources j in order,
 * Gamma_j zeta(r_ij/sigma_j)/sigma_j^3 added onto J[1:3] (:541-554).  No
 * reset (zeta_fmm zeroes nothing) and no static skip.  The reference's loop
 * writes the particles of its *source* branch with the strengths of its
 * *target* branch; a caller restating it passes those as targets and
 * sources here. */
void vpmo_zeta_ranges(double* P, int64_t stride, int kernel,
                      const int64_t* tb, const int64_t* te,
                      const int64_t* sb, const int64_t* se, int64_t npairs)
{
    for (int64_t k = 0; k < npairs; k++)
        for (int64_t i = tb[k]; i < te[k]; i++) {
            double* pi = P + i * stride;
            for (int64_t j = sb[k]; j < se[k]; j++) {
                const double* pj = P + j * stride;
                const double dX1 = pi[RX] - pj[RX];                  /* :545-547 */
                const double dX2 = pi[RX + 1] - pj[RX + 1];
                const double dX3 = pi[RX + 2] - pj[RX + 2];
                const double r = sqrt(dX1 * dX1 + dX2 * dX2 + dX3 * dX3);
                const double s = pj[RSIG];
                const double zeta_sgm = 1 / (s * s * s) * vpmo_zeta(kernel, r / s);   /* :550 */
                pi[RJ + 0] += pj[RG + 0] * zeta_sgm;                 /* :552-554 */
                pi[RJ + 1] += pj[RG + 1] * zeta_sgm;
                pi[RJ + 2] += pj[RG + 2] * zeta_sgm;
            }
        }
}

/* ---- Estr_fmm! over a direct list (subfilterscale_models.jl:157-188) ----
 * For every (target range, source range) pair in list order, every source j
 * and every target i of the ranges: Estr_direct(i, j) (:16-41) onto SFS_i.
 * No static particle is skipped and SFS is not reset (unlike Estr_direct!). */
void vpmo_estr_ranges(double* P, int64_t stride, int kernel, int transposed,
                      const int64_t* tb, const int64_t* te,
                      const int64_t* sb, const int64_t* se, int64_t npairs)
{
    for (int64_t k = 0; k < npairs; k++)
        for (int64_t j = sb[k]; j < se[k]; j++) {
            const double* sp = P + j * stride;
            for (int64_t i = tb[k]; i < te[k]; i++) {
                double* tp = P + i * stride;
                const double dx = sp[RX] - tp[RX], dy = sp[RX + 1] - tp[RX + 1], dz = sp[RX + 2] - tp[RX + 2];
                const double r = sqrt(dx * dx + dy * dy + dz * dz);
                estr_pair(kernel, transposed, tp, sp, r);
            }
        }
}

/* ---- E_str summation scale ----
 * s_i = sum over the pairs of Estr_direct! (subfilterscale_models.jl:16-41,
 * :63-66 static skips) of zeta_sigma_j(r) |S_ij|, with S_ij the pair's
 * (J_i - J_j)^T Gamma_j (or the classic form).  Whatever the cancellation in
 * E_str_i = sum zeta S_ij, an evaluation whose pair terms carry a relative
 * rounding error u is off by at most a small multiple of u s_i. */
static inline double estr_pair_abs(int kernel, int transposed, const double* tp, const double* sp, double r)
{
    double T[VPMO_NFIELDS];
    for (int c = 0; c < VPMO_NFIELDS; c++) T[c] = tp[c];
    T[RSFS] = T[RSFS + 1] = T[RSFS + 2] = 0.0;
    estr_pair(kernel, transposed, T, sp, r);
    return sqrt(T[RSFS] * T[RSFS] + T[RSFS + 1] * T[RSFS + 1] + T[RSFS + 2] * T[RSFS + 2]);
}

void vpmo_estr_scale(const double* P, int64_t np, int kernel, int transposed, double* s, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for schedule(static) num_threads(nthreads) if (nthreads > 1)
    for (int64_t i = 0; i < np; i++) {
        const double* tp = P + i * VPMO_NFIELDS;
        double acc = 0.0;
        if (!is_static(tp)) {
            for (int64_t j = 0; j < np; j++) {
                const double* sp = P + j * VPMO_NFIELDS;
                if (is_static(sp)) continue;
                const double dx = sp[RX] - tp[RX], dy = sp[RX + 1] - tp[RX + 1], dz = sp[RX + 2] - tp[RX + 2];
                acc += estr_pair_abs(kernel, transposed, tp, sp, sqrt(dx * dx + dy * dy + dz * dz));
            }
        }
        s[i] = acc;
    }
}

/* the same over the pairs of vpmo_estr_ranges (accumulated onto s) */
/* The U summation scale over the pairs of a direct list: s_i accumulates
 * |g(r/sigma_j) crss_ij| (the 2-norm of the pair's U term, fmm.jl:125-138)
 * for every pair of vpmo_direct_ranges, coincident pairs skipped as there.
 * Any evaluation whose terms carry relative errors e_j is off by at most
 * max_j e_j times s_i, however much the U sum cancels. */
void vpmo_u_scale_ranges(const double* P, int64_t stride, int kernel,
                         const int64_t* tb, const int64_t* te, const int64_t* sb, const int64_t* se,
                         int64_t npairs, double* s)
{
    const double c4 = const4();
    for (int64_t k = 0; k < npairs; k++)
        for (int64_t i = tb[k]; i < te[k]; i++) {
            const double* tp = P + i * stride;
            for (int64_t j = sb[k]; j < se[k]; j++) {
                const double* sp = P + j * stride;
                const double dx = tp[RX] - sp[RX], dy = tp[RX + 1] - sp[RX + 1], dz = tp[RX + 2] - sp[RX + 2];
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (r2 == 0.0) continue;
                const double r = sqrt(r2);
                double g, dg;
                vpmo_g_dgdr(kernel, r / sp[RSIG], &g, &dg);
                const double cx = dy * sp[RG + 2] - dz * sp[RG + 1];
                const double cy = dz * sp[RG] - dx * sp[RG + 2];
                const double cz = dx * sp[RG + 1] - dy * sp[RG];
                s[i] += fabs(g) * c4 / (r2 * r) * sqrt(cx * cx + cy * cy + cz * cz);
            }
        }
}

void vpmo_estr_scale_ranges(const double* P, int64_t stride, int kernel, int transposed,
                            const int64_t* tb, const int64_t* te, const int64_t* sb, const int64_t* se,
                            int64_t npairs, double* s)
{
    for (int64_t k = 0; k < npairs; k++)
        for (int64_t j = sb[k]; j < se[k]; j++) {
            const double* sp = P + j * stride;
            for (int64_t i = tb[k]; i < te[k]; i++) {
                const double* tp = P + i * stride;
                const double dx = sp[RX] - tp[RX], dy = sp[RX + 1] - tp[RX + 1], dz = sp[RX + 2] - tp[RX + 2];
                s[i] += estr_pair_abs(kernel, transposed, tp, sp, sqrt(dx * dx + dy * dy + dz * dz));
            }
        }
}

/* ---- core-spreading reset with RBF interpolation ----
 * viscous.jl:202-220 (target vorticity from the zeta field, sigma <- sgm0)
 * and rbf_conjugategradient (:314-478), with zeta = zeta_direct.  Vectors
 * and sums run over the non-static particles (iterator(pfield),
 * particlefield.jl:368-374); zeta_direct over all of them.  alpha is 0 when
 * a component has converged (Julia's false is a strong zero).  Returns 0, or
 * 1 when itmax is reached unconverged with iterror set; *iters and res[3]
 * (sqrt(rr/rr0)) report the iteration count and residuals. */
int vpmo_core_spreading_reset(double* P, int64_t np, int kernel, double sgm0, int itmax, double tol,
                              int iterror, int nthreads, int* iters, double res[3])
{
    const double eps2 = 2 * 2.220446049250313e-16;
    vpmo_zeta_direct(P, np, kernel, nthreads);                      /* :204 */
    for (int64_t n = 0; n < np; n++) {                              /* :206-214 */
        double* p = P + n * VPMO_NFIELDS;
        if (is_static(p)) continue;
        for (int i = 0; i < 3; i++) p[RM + 6 + i] = p[RJ + i];
        p[RSIG] = sgm0;
    }
    double rr0[3] = {0, 0, 0}, rrs[3], prev[3], pAp[3], alpha[3], beta[3];
    int flags[3];
    for (int64_t n = 0; n < np; n++) {                              /* :345-352 */
        double* p = P + n * VPMO_NFIELDS;
        if (is_static(p)) continue;
        for (int i = 0; i < 3; i++) {
            p[RM + i] = p[RM + 6 + i] * p[RVOL];
            p[RG + i] = p[RM + i];
        }
    }
    vpmo_zeta_direct(P, np, kernel, nthreads);                      /* :355 */
    for (int64_t n = 0; n < np; n++) {                              /* :357-368 */
        double* p = P + n * VPMO_NFIELDS;
        if (is_static(p)) continue;
        for (int i = 0; i < 3; i++) {
            p[RM + 3 + i] = p[RM + 6 + i] - p[RJ + i];
            p[RG + i] = p[RM + 3 + i];
            rr0[i] += p[RM + 3 + i] * p[RM + 3 + i];
        }
    }
    for (int i = 0; i < 3; i++) {
        rrs[i] = rr0[i];
        flags[i] = sqrt(rr0[i]) > tol || sqrt(rrs[i] / rr0[i]) > tol;
    }
    int it = 0, status = 0;
    for (it = 1; it <= itmax; it++) {                               /* :376-447 */
        if (!(flags[0] || flags[1] || flags[2])) { it--; break; }
        vpmo_zeta_direct(P, np, kernel, nthreads);
        for (int i = 0; i < 3; i++) pAp[i] = 0;
        for (int64_t n = 0; n < np; n++) {
            const double* p = P + n * VPMO_NFIELDS;
            if (is_static(p)) continue;
            for (int i = 0; i < 3; i++) pAp[i] += p[RG + i] * p[RJ + i];
        }
        for (int i = 0; i < 3; i++) {
            alpha[i] = flags[i] ? rrs[i] / pAp[i] : 0.0;
            prev[i] = rrs[i];
            rrs[i] = 0;
        }
        for (int64_t n = 0; n < np; n++) {
            double* p = P + n * VPMO_NFIELDS;
            if (is_static(p)) continue;
            for (int i = 0; i < 3; i++) {
                p[RM + i] += alpha[i] * p[RG + i];
                p[RM + i + 3] -= alpha[i] * p[RJ + i];
                rrs[i] += p[RM + i + 3] * p[RM + i + 3];
            }
        }
        for (int i = 0; i < 3; i++) {
            beta[i] = rrs[i] / prev[i];
            if (fabs(prev[i]) <= eps2) beta[i] = 1;
        }
        for (int64_t n = 0; n < np; n++) {
            double* p = P + n * VPMO_NFIELDS;
            if (is_static(p)) continue;
            for (int i = 0; i < 3; i++) p[RG + i] = p[RM + i + 3] + beta[i] * p[RG + i];
        }
        for (int i = 0; i < 3; i++)
            flags[i] = flags[i] && (fabs(rr0[i]) <= eps2 ? 0 : sqrt(rrs[i] / rr0[i]) > tol);
        if (it == itmax && (flags[0] || flags[1] || flags[2]) && iterror) status = 1;
    }
    if (it > itmax) it = itmax;
    for (int64_t n = 0; n < np; n++) {                              /* :450-455 */
        double* p = P + n * VPMO_NFIELDS;
        if (is_static(p)) continue;
        for (int i = 0; i < 3; i++) p[RG + i] = p[RM + i];
    }
    if (iters) *iters = it;
    if (res) for (int i = 0; i < 3; i++) res[i] = rr0[i] > 0 ? sqrt(rrs[i] / rr0[i]) : 0.0;
    return status;
}
