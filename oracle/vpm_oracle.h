/*
 * vpm_oracle.h -- FP64 CPU restatement of FLOWVPM's particle-to-particle path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP library is
 * compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never links it.
 *
 * Particle storage follows the reference exactly: a particle-major FP64
 * matrix, one particle every `stride` (= 46) doubles, rows
 * (0-based) X 0-2, Gamma 3-5, sigma 6, vol 7, circulation 8, U 9-11,
 * vorticity 12-14, J 15-23, PSE 24-26, M 27-35, C 36-38, SFS 39-41,
 * static 42, U_prev 43   (src/FLOWVPM_particlefield.jl:11,239-252).
 * J is column-major: J[3*b + a] = du_a/dx_b (src/FLOWVPM_fmm.jl:160).
 * All index ranges are 0-based half-open [begin, end).
 */
#ifndef VPM_ORACLE_H
#define VPM_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { VPMO_SINGULAR = 0, VPMO_GAUSSIAN = 1, VPMO_GAUSSIANERF = 2, VPMO_WINCKELMANS = 3 };
enum { VPMO_NFIELDS = 46 };

/* scalar functions ------------------------------------------------------- */
double vpmo_erf64(double x);                           /* gpu_erf.jl:159-191 */
void   vpmo_g_dgdr(int kernel, double q, double* g, double* dg);  /* kernel.jl:45-84 */
double vpmo_zeta(int kernel, double q);                /* kernel.jl:45,51,60,69 */
void   vpmo_g_derivs(int kernel, double q, double* g, double* dg, double* d2g);

/* U/J --------------------------------------------------------------------- */
/* fmm.direct! (fmm.jl:102-168) + buffer_to_target_system! (fmm.jl:170-176):
 * per target, contributions of sources s0..s1-1 are summed in source order
 * into a zeroed buffer, then the buffer is added onto U and J of the target. */
void vpmo_direct2(double* Pt, int64_t stride_t, int64_t t0, int64_t t1,
                  const double* Ps, int64_t stride_s, int64_t s0, int64_t s1,
                  int kernel, int nthreads);
/* Leaf-pair list form: pairs k=0..npairs-1 of (target range, source range),
 * accumulated per target in list order into one buffer, then added. */
void vpmo_direct_ranges(double* P, int64_t stride, int kernel,
                        const int64_t* tb, const int64_t* te,
                        const int64_t* sb, const int64_t* se, int64_t npairs);
void vpmo_reset_particles(double* P, int64_t np);      /* particlefield.jl:464-490 */
void vpmo_reset_particles_sfs(double* P, int64_t np);  /* particlefield.jl:492-507 */
/* UJ_direct (UJ.jl:21-37) */
void vpmo_uj_direct(double* P, int64_t np, int kernel, int reset, int reset_sfs,
                    int sfs, int transposed, int nthreads);
/* Estr_direct! (subfilterscale_models.jl:16-92) */
void vpmo_estr_direct(double* P, int64_t np, int kernel, int transposed, int nthreads);
/* E_str summation scale s_i = sum zeta |S_ij| over Estr_direct!'s pairs,
 * or (accumulated onto s) over the pairs of vpmo_estr_ranges */
void vpmo_estr_scale(const double* P, int64_t np, int kernel, int transposed, double* s, int nthreads);
void vpmo_estr_scale_ranges(const double* P, int64_t stride, int kernel, int transposed,
                            const int64_t* tb, const int64_t* te, const int64_t* sb, const int64_t* se,
                            int64_t npairs, double* s);
/* the U summation scale s_i = sum |U term| over the pairs of vpmo_direct_ranges */
void vpmo_u_scale_ranges(const double* P, int64_t stride, int kernel,
                         const int64_t* tb, const int64_t* te, const int64_t* sb, const int64_t* se,
                         int64_t npairs, double* s);
/* the same for targets [t0, t1) only (sources: all np) */
void vpmo_estr_direct_targets(double* P, int64_t np, int64_t t0, int64_t t1, int kernel, int transposed,
                              int nthreads);

/* per-particle FP64 schemes ------------------------------------------------ */
/* update_particle_states, rVPM (timeintegration.jl:463-532) */
void vpmo_update_particle_states(double* P, int64_t np, double a, double b, double dt,
                                 const double Uinf[3], double f, double g,
                                 double zeta0, int transposed);
/* relax_pedrizzetti / relax_correctedpedrizzetti (relaxation.jl:62-81,117-142) */
void vpmo_relax(double* P, int64_t np, double rlxf, int corrected);
/* CoreSpreading, rungekutta3 branch (viscous.jl:167-180).  Advances the
 * timer on the last stage and returns 1 when the RBF reset would be
 * triggered (viscous.jl:188-202), 0 otherwise. */
int vpmo_core_spreading_rk3(double* P, int64_t np, double nu, double sgm0, double beta,
                            double* t_sgm, double a, double b, double dt);
/* dynamic SFS, pseudo-3-level procedure (subfilterscale.jl:447-673) and
 * backscatter clipping (subfilterscale.jl:225-247,287-296).  Return 0, or a
 * negative code: -1 invalid bounds, -2 NaN coefficient. */
int vpmo_sfs_before(double* P, int64_t np, int kernel, int transposed, double alpha,
                    double minC, double maxC, int nthreads);
int vpmo_sfs_after(double* P, int64_t np, int kernel, int transposed, double alpha,
                   double rlxf, double minC, double maxC, int force_positive, int clip);

typedef struct {
    int32_t kernel;          /* VPMO_* */
    int32_t transposed;      /* 1: transposed stretching */
    double  f, g;            /* ReformulatedVPM parameters */
    double  Uinf[3];
    int32_t sfs;             /* 0 none, 1 dynamic pseudo-3-level */
    int32_t sfs_force_positive, sfs_clip;
    int32_t relax;           /* 0 none, 1 pedrizzetti, 2 corrected */
    double  sfs_alpha, sfs_rlxf, sfs_minC, sfs_maxC;
    double  rlxf;
    int32_t viscous;         /* 0 inviscid, 1 core spreading */
    int32_t nthreads;
    double  nu, sgm0, beta;
    double  t_sgm;           /* core-spreading timer, updated in place */
    int32_t rbf_itmax;       /* CoreSpreading itmax; 0: no reset (status 1) */
    int32_t rbf_iterror;     /* CoreSpreading iterror */
    double  rbf_tol;         /* CoreSpreading tol */
} vpmo_settings;

/* One rungekutta3 step for ReformulatedVPM (timeintegration.jl:388-461) with
 * UJ = UJ_direct.  relax_now mirrors run_vpm!'s relax flag (utils.jl:94-96).
 * When the core growth reaches beta in the last stage, the reset
 * (vpmo_core_spreading_reset with the rbf_* settings) runs there and t_sgm
 * returns to 0 (viscous.jl:202-220).  Returns 0, a negative SFS code, 1 when
 * a reset was due with rbf_itmax == 0 (not performed), or 2 when the reset's
 * CG did not converge with rbf_iterror set (the step is completed). */
int vpmo_rk3_step(double* P, int64_t np, vpmo_settings* s, double dt, int relax_now);

/* adjoint ------------------------------------------------------------------ */
/* Gradient of L = sum_i Ubar_i . U_i + sum_i sum_f Jbar_i[f] J_i[f] with
 * respect to X (3N), Gamma (3N), sigma (N), U and J from one UJ_direct of the
 * whole field.  Closed-form pair pullback; outputs are overwritten. */
void vpmo_uj_adjoint(const double* P, int64_t np, int kernel,
                     const double* Ubar, const double* Jbar,
                     double* Xbar, double* Gbar, double* Sbar, int nthreads);
/* The same sums with |contribution| per pair and component: the summation
 * scale that bounds the rounding error of any evaluation of the adjoint. */
void vpmo_uj_adjoint_scale(const double* P, int64_t np, int kernel,
                           const double* Ubar, const double* Jbar,
                           double* Xmag, double* Gmag, double* Smag, int nthreads);
/* The part of L that depends on particle k (all pairs with k as target or as
 * source), summed in long double; for finite differences. */
double vpmo_loss_local(const double* P, int64_t np, int kernel,
                       const double* Ubar, const double* Jbar, int64_t k);
/* Estr_fmm! over a direct list (subfilterscale_models.jl:157-188): the E_str
 * pair (:16-41) for every (target range, source range), list order, onto
 * SFS; no static skip, no reset.  0-based half-open ranges, rows of stride. */
void vpmo_estr_ranges(double* P, int64_t stride, int kernel, int transposed,
                      const int64_t* tb, const int64_t* te,
                      const int64_t* sb, const int64_t* se, int64_t npairs);
/* Core-spreading reset (viscous.jl:202-220) with rbf_conjugategradient
 * (:314-478) and zeta_direct; returns 1 when unconverged with iterror. */
int vpmo_core_spreading_reset(double* P, int64_t np, int kernel, double sgm0, int itmax, double tol,
                              int iterror, int nthreads, int* iters, double res[3]);
/* The reference's own update pullback state_pullback (rrules.jl:436-547),
 * restated literally in the record layout of
 * vpmo_update_particle_states_adjoint (transposed scheme).  full != 0 uses
 * the whole sum over zeta where the reference keeps zeta = eta (:521, :524).
 * Cross-check only. */
void vpmo_update_particle_states_pullback_rrules(const double* P, double* B, int64_t np, double a, double b,
                                                 double dt, double f, double g, double zeta0, int full);
/* The reference's own pullback UJS_pullback (rrules.jl:183-326), restated
 * literally for a field acting on itself; a cross-check of vpmo_uj_adjoint
 * (the reference marks parts of it inexact).  Outputs are overwritten. */
void vpmo_uj_adjoint_rrules(const double* P, int64_t np, int kernel, const double* Ubar, const double* Jbar,
                            double* Xbar, double* Gbar, double* Sbar);
/* zeta_direct (viscous.jl:486-515): J[1:3] of every particle (static
 * included) <- sum over all particles j (self included) of
 * Gamma_j zeta(r/sigma_j)/sigma_j^3. */
void vpmo_zeta_direct(double* P, int64_t np, int kernel, int nthreads);
/* zeta_fmm's pair loop (viscous.jl:523-558) over a direct list: per
 * (target range, source range), J[1:3] of each target += Gamma_j
 * zeta(r/sigma_j)/sigma_j^3 over the sources; no reset, no static skip.
 * 0-based half-open ranges, rows of stride. */
void vpmo_zeta_ranges(double* P, int64_t stride, int kernel, const int64_t* tb, const int64_t* te,
                      const int64_t* sb, const int64_t* se, int64_t npairs);
/* Pullback of update_particle_states (ReformulatedVPM, no SFS;
 * timeintegration.jl:467-528, rrules.jl:405-555).  B: cotangent records
 * (46 rows per particle) of the updated state in, of the state before the
 * update out, plus Ubar/Jbar in the U/J rows.  P: the state before. */
void vpmo_update_particle_states_adjoint(const double* P, double* B, int64_t np, double a, double b,
                                         double dt, double f, double g, int transposed);
/* Pullback of one vpmo_rk3_step (no SFS, inviscid, no relaxation); B as
 * above for the whole step.  Returns 0, or -1 for an unsupported setting. */
int vpmo_rk3_step_adjoint(const double* P0, int64_t np, const vpmo_settings* s, double dt, double* B);

#ifdef __cplusplus
}
#endif
#endif
