/*
 * flowvpm_hip.h -- C ABI of the MI355X (gfx950) particle-to-particle library.
 *
 * Drop-in boundary for FLOWVPM's P2P path.  The caller keeps its particle
 * matrix exactly as ParticleField stores it: Float64, particle-major, one
 * 46-double record per particle (src/FLOWVPM_particlefield.jl:11,134), rows
 * (0-based) X 0-2, Gamma 3-5, sigma 6, vol 7, circulation 8, U 9-11,
 * vorticity 12-14, J 15-23, PSE 24-26, M 27-35, C 36-38, SFS 39-41,
 * static 42, U_prev 43 (particlefield.jl:239-252).  J is column-major,
 * J[3*b + a] = du_a/dx_b (src/FLOWVPM_fmm.jl:160).
 *
 * A vpm_field is a device-resident copy of one particle matrix on one GPU;
 * it persists across calls, so a time step never re-uploads state.  Every
 * function returns a vpm_status; on failure vpm_last_error() (thread-local)
 * describes it.  No function keeps a caller pointer past its return.  Calls
 * on one vpm_field are serialized by an internal mutex; distinct fields are
 * independent.  Index ranges are 0-based and half-open.
 */
#ifndef FLOWVPM_HIP_H
#define FLOWVPM_HIP_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define VPM_NFIELDS 46
#define VPM_ABI_VERSION 1

typedef enum {
    VPM_OK = 0,
    VPM_EINVAL = -1,    /* bad argument (null pointer, bad id, bad range)      */
    VPM_ERANGE = -2,    /* particle count above the field's capacity           */
    VPM_ENOMEM = -3,    /* device or host allocation failed                    */
    VPM_EDEVICE = -4,   /* no such device or a HIP runtime error               */
    VPM_ENAN = -5,      /* NaN in the dynamic SFS coefficient (subfilterscale.jl:645-652) */
    VPM_ERESET = -6,    /* core-spreading reset required (viscous.jl:202) and rbf_itmax == 0 */
    VPM_ECONVERGE = -7  /* RBF conjugate gradient unconverged at itmax (viscous.jl:421-431) */
} vpm_status;

/* Regularization kernels (src/FLOWVPM.jl:129-133, src/FLOWVPM_kernel.jl:45-84) */
enum { VPM_KERNEL_SINGULAR = 0, VPM_KERNEL_GAUSSIAN = 1,
       VPM_KERNEL_GAUSSIANERF = 2, VPM_KERNEL_WINCKELMANS = 3 };

/* Pair-arithmetic precision.  F32C: FP32 pair arithmetic, compensated
 * per-target sums, results folded into the FP64 state.  F64: the reference's
 * FP64 pair arithmetic (fmm.jl:113-161 with custom_erf64). */
enum { VPM_PREC_F32C = 0, VPM_PREC_F64 = 1 };

/* flags of vpm_uj_direct: the keywords of UJ_direct (src/FLOWVPM_UJ.jl:21-25),
 * and two of this library's own:
 *  VPM_UJ_NO_SORT  evaluate in storage order (no spatial sort);
 *  VPM_UJ_F64_FAR  with VPM_PREC_F64 only: FP64 pair arithmetic with one
 *                  rsqrt per pair and, for Gaussian-erf (q >= 10) and
 *                  Gaussian (q^2 >= 13), the far form F' = r^-3, H'' = r^-5
 *                  where the kernel's g is 1 and q g' is 0 to FP64 rounding;
 *                  near pairs take the reference's g and dg.  It agrees with
 *                  the reference arithmetic to FP64 rounding of the sums.
 *                  With VPM_UJ_SFS, E_str is evaluated in FP64 over the
 *                  pairs within q = 10 (Gaussian-erf) or q^2 = 14 (Gaussian),
 *                  where zeta has fallen below 2e-22 of zeta(0).
 *                  Not taken by vpm_uj_direct_noalloc (VPM_EINVAL). */
enum { VPM_UJ_RESET = 1, VPM_UJ_RESET_SFS = 2, VPM_UJ_SFS = 4,
       VPM_UJ_NO_SORT = 8, VPM_UJ_F64_FAR = 16 };

typedef struct {
    int32_t kernel;       /* VPM_KERNEL_*                                    */
    int32_t precision;    /* VPM_PREC_*                                      */
    int32_t transposed;   /* 1: transposed stretching (ParticleField.transposed) */
    int32_t reserved;
} vpm_config;

/* Settings of one rungekutta3 step (src/FLOWVPM_timeintegration.jl:388-461) */
typedef struct {
    vpm_config uj;
    double f, g;               /* ReformulatedVPM{f,g} (src/FLOWVPM_formulation.jl:31-37) */
    double Uinf[3];            /* pfield.Uinf(t) */
    int32_t sfs;               /* 0 NoSFS, 1 DynamicSFS pseudo-3-level (subfilterscale.jl:172-268) */
    int32_t sfs_force_positive;
    int32_t sfs_clip;          /* clipping_backscatter */
    int32_t relax;             /* 0 none, 1 pedrizzetti, 2 correctedpedrizzetti */
    double sfs_alpha, sfs_rlxf, sfs_minC, sfs_maxC;
    double rlxf;
    int32_t viscous;           /* 0 Inviscid, 1 CoreSpreading (viscous.jl:65-223) */
    int32_t uj_flags;          /* 0, or VPM_UJ_F64_FAR (with uj.precision VPM_PREC_F64):
                                  the step's U/J evaluations take the FP64 far form */
    double nu, sgm0, beta;
    double t_sgm;              /* core-spreading timer; updated in place */
    /* CoreSpreading's RBF reset (viscous.jl:202-220, 314-478): itmax, tol,
     * iterror of the CoreSpreading scheme (viscous.jl:63-101).  rbf_itmax == 0
     * leaves the reset to the caller (the step returns VPM_ERESET). */
    int32_t rbf_itmax;
    int32_t rbf_iterror;
    double rbf_tol;
} vpm_step_config;

typedef struct vpm_field_s* vpm_field;

/* ---- library ---------------------------------------------------------- */
const char* vpm_version(void);
const char* vpm_last_error(void);
int vpm_device_count(int* count);

/* ---- particle fields: ParticleField storage (particlefield.jl:67-141) --- */
/* Allocate a field for up to maxparticles particles on `device`
 * (1 <= maxparticles < 2^31). */
int vpm_create(int device, int64_t maxparticles, vpm_field* out);
int vpm_destroy(vpm_field f);
/* Copy np whole 46-row records in (np <= maxparticles).  Rejects sigma <= 0
 * or non-finite X, Gamma, sigma (the reference divides by sigma, fmm.jl:122). */
int vpm_set_particles(vpm_field f, const double* particles, int64_t np);
/* Copy rows [row0, row1) of the first np records out, into a 46-stride array
 * (other rows of `particles` are left untouched). */
int vpm_get_particles(vpm_field f, double* particles, int64_t np, int32_t row0, int32_t row1);
/* add_particle (particlefield.jl:167-205) for n whole 46-row records, appended
 * after the current ones; the checks of vpm_set_particles apply, and the
 * maxparticles overflow is VPM_ERANGE. */
int vpm_add_particles(vpm_field f, const double* particles, int64_t n);
/* remove_particle (particlefield.jl:401-419), 0-based i: the last particle
 * moves into slot i, the last slot's U, vorticity, J and PSE are zeroed,
 * np decreases by one. */
int vpm_remove_particle(vpm_field f, int64_t i);
int vpm_get_np(vpm_field f, int64_t* np);
int vpm_synchronize(vpm_field f);

/* ---- boundary #1: the UJ functor ---------------------------------------- */
/* UJ_direct(pfield; reset, reset_sfs, sfs) (src/FLOWVPM_UJ.jl:21-37):
 * reset zeroes U, vorticity, J, PSE of non-static particles
 * (particlefield.jl:464-490); reset_sfs zeroes SFS of non-static particles
 * (:492-507); then all-pairs U/J accumulate (fmm.jl:102-176); sfs then adds
 * E_str (Estr_direct!, subfilterscale_models.jl:43-92). */
int vpm_uj_direct(vpm_field f, const vpm_config* cfg, uint32_t flags);
/* vpm_uj_direct without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same kernels and results as vpm_uj_direct; every scratch buffer (sort
 * keys, staged sources, partial sums) comes from one 64 MiB device array per
 * GPU that the library holds until the process ends.  np above what fits
 * (about 3.4e5 particles for F32C, 4.9e5 for F64, with one source split)
 * returns VPM_ERANGE; fewer source splits than vpm_uj_direct takes are used
 * when its own do not fit.  VPM_UJ_SFS and VPM_UJ_F64_FAR are rejected (VPM_EINVAL).  Calls on
 * one GPU share the array and are serialized by the library (overlapping
 * ones would race on it). */
int vpm_uj_direct_noalloc(vpm_field f, const vpm_config* cfg, uint32_t flags);
/* UJ_direct(source, target) (src/FLOWVPM_UJ.jl:48-50): every particle of
 * `source` acts on every particle of `target`; accumulates, no reset.  Both
 * fields must live on the same device. */
int vpm_uj_direct_fields(vpm_field target, vpm_field source, const vpm_config* cfg);

/* ---- boundary #2: FastMultipole leaf-pair P2P ---------------------------- */
/* fmm.direct!(target_buffer, target_index, switch, source_system,
 * source_buffer, source_index) (src/FLOWVPM_fmm.jl:102-168), batched: for
 * k < npairs, sources [sb[k], se[k]) act on targets [tb[k], te[k]) of the
 * same field, accumulated into U and J like buffer_to_target_system!
 * (fmm.jl:170-176).  Host arrays; ranges must lie within [0, np). */
int vpm_p2p_ranges(vpm_field f, const vpm_config* cfg,
                   const int64_t* tb, const int64_t* te,
                   const int64_t* sb, const int64_t* se, int64_t npairs);
/* Estr_fmm! over a direct list (src/FLOWVPM_subfilterscale_models.jl:157-188):
 * for k < npairs, the E_str pair (:16-41) of every source in [sb[k], se[k])
 * on every target in [tb[k], te[k]), accumulated onto SFS.  As in the
 * reference, static particles are not skipped and SFS is not reset. */
int vpm_estr_ranges(vpm_field f, const vpm_config* cfg,
                    const int64_t* tb, const int64_t* te,
                    const int64_t* sb, const int64_t* se, int64_t npairs);
/* zeta_fmm's pair loop over a direct list (src/FLOWVPM_viscous.jl:523-558):
 * for k < npairs, J[1:3] of every target in [tb[k], te[k]) +=
 * Gamma_j zeta(r/sigma_j)/sigma_j^3 over the sources j in [sb[k], se[k]).
 * As in the reference, J is not reset and static particles are not skipped.
 * The reference's loop writes the particles of its source branch with the
 * strengths of its target branch: pass those as tb/te and sb/se here. */
int vpm_zeta_ranges(vpm_field f, const vpm_config* cfg,
                    const int64_t* tb, const int64_t* te,
                    const int64_t* sb, const int64_t* se, int64_t npairs);

/* A direct list planned once and served to several list evaluations:
 * UJ_fmm hands one direct_list to the near field (fmm.direct! per pair, with
 * nearfield_device, src/FLOWVPM_UJ.jl:90-101) and then to Estr_fmm!
 * (src/FLOWVPM_UJ.jl:125, subfilterscale_models.jl:157-188); zeta_fmm
 * (viscous.jl:523-558) walks such a list too.  vpm_plan_ranges checks and
 * plans the pairs (the arguments of vpm_p2p_ranges) for field f and keeps
 * the plan on the device; vpm_p2p_plan, vpm_estr_plan and vpm_zeta_plan then
 * run vpm_p2p_ranges, vpm_estr_ranges and vpm_zeta_ranges on it with no host
 * planning and no list upload, with the same results bit for bit (FP64 runs
 * its block kernels over the plan's items).  A plan is valid for its field
 * while the field holds the np it was made for (VPM_EINVAL otherwise).
 * Destroying the field releases a plan's device memory; the plan must still
 * be passed to vpm_plan_destroy. */
typedef struct vpm_plan_s* vpm_plan;
int vpm_plan_ranges(vpm_field f, const int64_t* tb, const int64_t* te,
                    const int64_t* sb, const int64_t* se, int64_t npairs, vpm_plan* out);
int vpm_plan_destroy(vpm_plan p);
int vpm_p2p_plan(vpm_field f, const vpm_config* cfg, vpm_plan p);
int vpm_estr_plan(vpm_field f, const vpm_config* cfg, vpm_plan p);
int vpm_zeta_plan(vpm_field f, const vpm_config* cfg, vpm_plan p);

/* ---- terms built from J ------------------------------------------------- */
/* Estr_direct! (src/FLOWVPM_subfilterscale_models.jl:16-92): SFS += sum over
 * non-static sources of zeta_sigma(r) (Gamma_s . grad)(U_t - U_s), for
 * non-static targets. */
int vpm_estr_direct(vpm_field f, const vpm_config* cfg);
int vpm_reset_particles(vpm_field f);          /* particlefield.jl:464-490 */
int vpm_reset_particles_sfs(vpm_field f);      /* particlefield.jl:492-507 */

/* ---- per-particle FP64 schemes ------------------------------------------ */
/* update_particle_states for ReformulatedVPM (timeintegration.jl:463-532) */
int vpm_update_particle_states(vpm_field f, double a, double b, double dt,
                               const double Uinf[3], double fr, double gr,
                               double zeta0, int32_t transposed);
/* relax_pedrizzetti / relax_correctedpedrizzetti (relaxation.jl:62-81,117-142) */
int vpm_relax(vpm_field f, double rlxf, int32_t corrected);
/* CoreSpreading sigma update, rungekutta3 branch (viscous.jl:170-175) */
int vpm_core_spreading(vpm_field f, double nu, double a, double b, double dt);
/* dynamicprocedure_pseudo3level_beforeUJ (subfilterscale.jl:447-539),
 * including its test-filter UJ + E_str evaluation */
int vpm_sfs_dynamic_before(vpm_field f, const vpm_config* cfg, double alpha,
                           double minC, double maxC);
/* dynamicprocedure_pseudo3level_afterUJ (subfilterscale.jl:541-673) followed
 * by clipping_backscatter when clip != 0 (subfilterscale.jl:225-247).
 * Returns VPM_ENAN when a coefficient is NaN (subfilterscale.jl:645-652). */
int vpm_sfs_dynamic_after(vpm_field f, const vpm_config* cfg, double alpha, double rlxf,
                          double minC, double maxC, int32_t force_positive, int32_t clip);
/* One rungekutta3 step for ReformulatedVPM with UJ = UJ_direct
 * (timeintegration.jl:388-461) and nextstep's U_prev update
 * (particlefield.jl:443-455).  relax_now mirrors run_vpm!'s relax flag
 * (utils.jl:94-96).  When CoreSpreading's core growth reaches beta in the
 * last stage (viscous.jl:189-220), the reset runs there as
 * vpm_core_spreading_reset with zeta = zeta_direct and the rbf_* settings,
 * and t_sgm returns to 0; with rbf_itmax == 0 the step completes without it
 * and returns VPM_ERESET.  VPM_ECONVERGE: the reset's CG did not converge
 * (the step is complete, Gamma holds the last iterate). */
int vpm_rk3_step(vpm_field f, vpm_step_config* s, double dt, int32_t relax_now);

/* zeta_direct (src/FLOWVPM_viscous.jl:486-515): J[1:3] of every particle,
 * static ones included, <- sum over every particle j (the target itself
 * included) of Gamma_j zeta(r/sigma_j)/sigma_j^3.  The other J rows are left
 * as they are.  The singular kernel always runs in FP64. */
int vpm_zeta_direct(vpm_field f, const vpm_config* cfg);

/* CoreSpreading reset (viscous.jl:202-220): target vorticity M[7:9] <- zeta
 * field (zeta_direct), sigma <- sgm0, then rbf_conjugategradient
 * (viscous.jl:314-478): Gamma of the non-static particles solves
 * sum_j Gamma_j zeta(r_ij/sgm0)/sgm0^3 = target, per component, by CG from
 * the guess target*vol, stopping when sqrt(rr/rr0) <= tol on all three
 * components or after itmax iterations.  M[1:6] hold the solution and the
 * residual, J[1:3] the last zeta field, as in the reference.  Dot products
 * are FP64 with a fixed summation order (deterministic).  *iters and res
 * (sqrt(rr/rr0) per component; either may be null) report the iteration.
 * Returns VPM_ECONVERGE when itmax is reached unconverged and iterror != 0
 * (Gamma then holds the last iterate; the reference raises). */
int vpm_core_spreading_reset(vpm_field f, const vpm_config* cfg, double sgm0, int32_t itmax, double tol,
                             int32_t iterror, int32_t* iters, double* res);
/* rbf_conjugategradient(pfield, cs) alone (viscous.jl:314-478): the CG of
 * vpm_core_spreading_reset on the target vorticity already in M[7:9], with
 * the core sizes as they are. */
int vpm_rbf_conjugategradient(vpm_field f, const vpm_config* cfg, int32_t itmax, double tol, int32_t iterror,
                              int32_t* iters, double* res);

/* ---- adjoint (replaces src/FLOWVPM_rrules.jl:179-330, UJS_pullback) ------ */
/* For L = sum_i Ubar_i . U_i + sum_i sum_k Jbar_i[k] J_i[k] with U, J from one
 * all-pairs evaluation of the field (fmm.jl:102-168), writes dL/dX (3 np),
 * dL/dGamma (3 np) and dL/dsigma (np), particle-major.  Host arrays;
 * Ubar is 3 np, Jbar 9 np in the J storage order.  cfg->precision selects
 * the fused FP32 pair kernel with FP64 partial sums (VPM_PREC_F32C) or FP64
 * pair arithmetic (VPM_PREC_F64).  With VPM_PREC_F32C and every Jbar zero (a
 * loss on U alone) a kernel without the Jbar terms runs, with the same
 * results bit for bit. */
int vpm_uj_adjoint(vpm_field f, const vpm_config* cfg, const double* Ubar, const double* Jbar,
                   double* Xbar, double* Gbar, double* Sbar);

/* ---- adjoint of the time step (replaces the ReverseDiff pass through
 *      rungekutta3, timeintegration.jl:388-461, with the pullbacks of
 *      rrules.jl:179-330 (pairs) and rrules.jl:344-555 (update)) ----------- */
/* A tape holds the three stage records of one rk3 step of a field: the
 * particle state after each stage's UJ evaluation (46 doubles per particle
 * and stage, device memory), i.e. per-stage checkpoints of X, Gamma, sigma
 * with their U and J. */
typedef struct vpm_tape_s* vpm_tape;
int vpm_tape_create(vpm_field f, vpm_tape* out);
/* Destroying a field releases the device memory of its tapes; such a tape
 * must still be passed to vpm_tape_destroy, and any other use of it returns
 * VPM_EINVAL. */
int vpm_tape_destroy(vpm_tape t);
/* vpm_rk3_step (ReformulatedVPM, no SFS, inviscid, no relaxation) that also
 * records its stage states into t (overwriting what t held). */
int vpm_rk3_step_taped(vpm_field f, vpm_step_config* s, double dt, vpm_tape t);
/* Pullback of the step recorded in t.  bar: np x 46 cotangent records
 * (host, particle record layout).  On entry the rows X, Gamma, sigma and M of
 * the state after the step; on exit the cotangents of the state before it
 * (X, Gamma, sigma; M rows zero).  The U/J adjoint runs at the precision of
 * the recorded step.  Leaves the field holding the state before the step
 * (with the U and J of its first stage). */
int vpm_rk3_step_adjoint(vpm_tape t, double* bar);

/* ---- measurement -------------------------------------------------------- */
/* When enabled, the library brackets each all-pairs kernel launch with HIP
 * events on the field's stream.  vpm_timing returns the summed duration (ms)
 * and the launch count of one kernel class since the last reset:
 * 0 U/J FP32c, 1 U/J FP64, 2 E_str, 3 adjoint, 4 zeta. */
int vpm_profile(vpm_field f, int32_t enable);
int vpm_timing(vpm_field f, int32_t which, double* total_ms, int64_t* count);

#ifdef __cplusplus
}
#endif
#endif
