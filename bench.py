"""Benchmark of the P2P U/J path on one MI355X.

Workload (the configuration the headline metric is quoted on): the round-jet
column of N = 1,048,576 particles, one all-pairs U+J evaluation (UJ_direct,
Gaussian-erf kernel) per step, FP32 pair arithmetic with compensated sums,
particle state resident in HBM.  Also measured in the same run:
  * the FP64 GPU path (the reference's FP64 pair arithmetic) on the same input,
  * the FP32c-vs-FP64 per-particle relative error of U and J,
  * the FP32 adjoint of L = sum |U|^2 on the same field and its time
    relative to the forward evaluation,
  * the FP64 CPU oracle on a bounded sample of targets (host cores).

    python bench.py [--steps K] [--warmup W] [--gpus 1]

Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "UJ pair-interactions/s FP32 vs FP64 at ≤1e-5 rel err; adjoint/forward ratio"
FLOP_PER_PAIR = {"gaussianerf": 88, "winckelmans": 97, "gaussian": 86, "singular": 79}
PEAK_FP32 = 157.3e12
PEAK_FP64 = 78.6e12
# HBM bytes per U/J launch from two rocprofv3 --pmc runs of this script
# (FETCH_SIZE scaled by 2.0 and WRITE_SIZE by 1.0, the factors a 256 MiB
# device copy reads back), keyed by (N, kernel)
TRAFFIC_BYTES = {(1 << 20, "gaussianerf"): (9.0304e6 * 2.0 + 9.1750e5) * 1024}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def _oracle():
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as orc
    orc.lib()
    return orc


def host_threads():
    """Threads of the CPU legs: OMP_NUM_THREADS when it is set (the host
    cores this job owns next to its GPU), else every CPU this process may run
    on; nproc and os.cpu_count() are recorded beside it."""
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
    return (omp or aff), {"nproc": aff, "cpu_count": os.cpu_count(), "OMP_NUM_THREADS": omp or None}


def time_cpu(run, n_total, target_s):
    """Times run(nt) (FP64 oracle over targets [0, nt)) on a sample sized by
    a short calibration to about target_s seconds; returns (targets/s, nt, s)."""
    nt = 16
    while True:                                   # calibration: at least 0.3 s
        t0 = time.perf_counter()
        run(nt)
        dt = time.perf_counter() - t0
        if dt > 0.3 or nt * 2 > n_total:
            break
        nt = min(n_total, nt * 4)
    nt = max(1, min(n_total, int(nt * target_s / max(dt, 1e-3))))
    t0 = time.perf_counter()
    Q = run(nt)
    dt = time.perf_counter() - t0
    return nt / dt, nt, dt, Q


def cpu_uj_rate(P, kernel, nthreads, target_s):
    """FP64 oracle U+J (fmm.direct! restated, OpenMP over targets) on a
    bounded sample of targets against all sources: pairs/s."""
    orc = _oracle()
    n = P.shape[0]

    def run(nt):
        Q = P.copy()
        orc.direct(Q, kernel, t_range=(0, nt), nthreads=nthreads)
        return Q

    tps, nt, dt, Q = time_cpu(run, n, target_s)
    return tps * n, nt, dt, Q


def cpu_baseline(P, kernel, target_s=12.0):
    """FP64 oracle (reference-algorithm restatement) on a bounded sample of
    targets against all sources, OpenMP over targets on the host threads."""
    nthreads, info = host_threads()
    rate, nt, dt, Q = cpu_uj_rate(P, kernel, nthreads, target_s)
    n = P.shape[0]
    out = {"value": rate, "unit": "pair-interactions/s", "cores": nthreads, "kind": "port",
           "sample": "%d targets x %d sources of the workload, FP64 oracle (oracle/vpm_oracle.c), "
                     "%.1f s" % (nt, n, dt)}
    out.update(info)
    return out, Q[:nt], nt


def copy_calibration(mib):
    """One device-to-device copy of `mib` MiB through the HIP runtime: a
    plain copy kernel of known traffic (reads and writes mib MiB), used to
    scale the FETCH_SIZE / WRITE_SIZE counters of a profiled run."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    nbytes = C.c_size_t(int(mib) << 20)
    a, b = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(a), nbytes) == 0 and hip.hipMalloc(C.byref(b), nbytes) == 0
    assert hip.hipMemset(a, 1, nbytes) == 0
    assert hip.hipDeviceSynchronize() == 0
    assert hip.hipMemcpy(b, a, nbytes, 3) == 0          # hipMemcpyDeviceToDevice
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)


def rollout_adjoint(vpm, P, kernel, nsteps=3, dt=0.005):
    """Forward + adjoint through nsteps rk3 steps on the jet (rVPM, no SFS,
    inviscid, no relaxation), per-stage checkpoints on the device,
    L = sum |U_i|^2 at the final state."""
    api = vpm.api
    n = P.shape[0]
    pf = api.ParticleField(n, kernel=kernel, formulation=api.rVPM, relaxation=api.norelaxation,
                           precision="f32c")
    api.add_particles(pf, P)
    tm = {}
    log("bench: rollout forward+adjoint, %d rk3 steps" % nsteps)
    L, bar = api.kinetic_energy_rollout_gradient(pf, dt, nsteps, timings=tm)
    pf.close()
    out = {"N": n, "steps": nsteps, "dt": dt, "formulation": "rVPM (f=0, g=1/5), no SFS, inviscid",
           "loss": "sum |U_i|^2 after the last step", "L": L,
           "forward_ms": tm["forward_ms"], "adjoint_ms": tm["adjoint_ms"],
           "ratio": tm["adjoint_ms"] / tm["forward_ms"],
           "forward_step_ms": tm["forward_step_ms"], "final_uj_ms": tm["final_uj_ms"],
           "final_adjoint_ms": tm["final_adjoint_ms"], "step_adjoint_ms": tm["step_adjoint_ms"],
           "uj_evaluations": 3 * nsteps + 1, "checkpoint_bytes": tm["tape_bytes"],
           "grad_norms": {"X": float(np.linalg.norm(bar[:, 0:3])), "Gamma": float(np.linalg.norm(bar[:, 3:6])),
                          "sigma": float(np.linalg.norm(bar[:, 6]))}}
    log("bench: rollout forward %.0f ms, adjoint %.0f ms" % (tm["forward_ms"], tm["adjoint_ms"]))
    return out


def leaf_list(vpm, P, kernel, allpairs_rate, leaf=64, reps=3):
    """The near-field service of an FMM (FastMultipole's direct list through
    fmm.direct!, src/FLOWVPM_fmm.jl:102-168) on the jet: leaves of `leaf`
    Morton-consecutive particles, direct pairs within 3 median leaf radii,
    one vpm_p2p_ranges call per evaluation.  Then Estr_fmm! over the same
    list (src/FLOWVPM_subfilterscale_models.jl:157-188), one vpm_estr_ranges
    call per evaluation, with the J the list evaluation left, and zeta_fmm's
    pair loop (src/FLOWVPM_viscous.jl:523-558), one vpm_zeta_ranges call."""
    import ctypes as C
    api = vpm.api
    Q = vpm.workloads.morton_order(P)
    n = Q.shape[0]
    tb, te, sb, se = vpm.workloads.leaf_direct_list(Q, leaf)
    pairs = float(np.sum((te - tb) * (se - sb)))
    f = api.Field(n, 0)
    f.set_particles(Q)
    cfg = api.make_config(kernel, "f32c", True)
    arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (tb, te, sb, se)]
    ptr = [a.ctypes.data_as(vpm._lib._i64p) for a in arrs]
    L = vpm._lib.lib()
    vpm._lib.check(L.vpm_p2p_ranges(f.handle, C.byref(cfg), *ptr, len(tb)))
    f.profile(True)
    t0 = time.perf_counter()
    for _ in range(reps):
        vpm._lib.check(L.vpm_p2p_ranges(f.handle, C.byref(cfg), *ptr, len(tb)))
    wall = (time.perf_counter() - t0) / reps
    ms, cnt = f.timing(0)
    vpm._lib.check(L.vpm_estr_ranges(f.handle, C.byref(cfg), *ptr, len(tb)))
    t0 = time.perf_counter()
    for _ in range(reps):
        vpm._lib.check(L.vpm_estr_ranges(f.handle, C.byref(cfg), *ptr, len(tb)))
    ewall = (time.perf_counter() - t0) / reps
    ems, ecnt = f.timing(2)
    vpm._lib.check(L.vpm_zeta_ranges(f.handle, C.byref(cfg), *ptr, len(tb)))
    t0 = time.perf_counter()
    for _ in range(reps):
        vpm._lib.check(L.vpm_zeta_ranges(f.handle, C.byref(cfg), *ptr, len(tb)))
    zwall = (time.perf_counter() - t0) / reps
    zms, zcnt = f.timing(4)
    # the same list planned once (vpm_plan_ranges) and served to U/J, E_str
    # and zeta, as UJ_fmm hands one direct_list to the near field and to
    # Estr_fmm!: call times without the host plan and the list upload
    t0 = time.perf_counter()
    plan = f.plan_ranges(*arrs)
    plan_ms = 1e3 * (time.perf_counter() - t0)
    planned = {}
    for name, fn, which in (("uj", lambda: plan.p2p(kernel, "f32c"), 0),
                            ("estr", lambda: plan.estr(kernel, "f32c", True), 2),
                            ("zeta", lambda: plan.zeta(kernel, "f32c"), 4)):
        fn()
        f.profile(True)
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        call = 1e3 * (time.perf_counter() - t0) / reps
        kms, kc = f.timing(which)
        planned[name] = {"call_ms": call, "kernel_ms": kms / max(kc, 1), "call_over_kernel": call / (kms / max(kc, 1))}
    # the reference's own precision on the same plan: FP64 pair arithmetic,
    # each target summing its list in order (one kernel call each)
    planned_f64 = {}
    for name, fn, which in (("uj", lambda: plan.p2p(kernel, "f64"), 1),
                            ("estr", lambda: plan.estr(kernel, "f64", True), 2),
                            ("zeta", lambda: plan.zeta(kernel, "f64"), 4)):
        fn()
        f.profile(True)
        fn()
        kms, kc = f.timing(which)
        f.profile(False)
        planned_f64[name] = {"kernel_ms": kms / max(kc, 1), "pairs_per_s": pairs / (kms / max(kc, 1) / 1e3)}
    plan.close()
    f.close()
    out = {"N": n, "leaf": leaf, "list_pairs": len(tb), "particle_pairs": pairs, "sources_per_target": pairs / n,
           "call_ms": 1e3 * wall, "kernel_ms": ms / cnt, "pairs_per_s_kernel": pairs / (ms / cnt / 1e3),
           "pairs_per_s_call": pairs / wall, "fraction_of_allpairs_rate": pairs / (ms / cnt / 1e3) / allpairs_rate,
           "estr_call_ms": 1e3 * ewall, "estr_kernel_ms": ems / ecnt,
           "estr_pairs_per_s_kernel": pairs / (ems / ecnt / 1e3),
           "zeta_call_ms": 1e3 * zwall, "zeta_kernel_ms": zms / zcnt,
           "zeta_pairs_per_s_kernel": pairs / (zms / zcnt / 1e3),
           "planned": dict(planned, plan_ms=plan_ms), "planned_f64": planned_f64}
    log("bench: leaf list %s" % out)
    return out


def cfg1_cloud(vpm, n=65536, reps=5, cpu=True):
    """Random cloud N = 65,536 (X ~ U[0,1]^3, Gamma ~ N(0, 1e-3^2), sigma ~
    U[0.02, 0.04], seed 1): one U+J evaluation per kernel in FP32c, FP64 and
    FP64 with the far form on the GPU, throughput and per-output error of
    FP32c (and of the far form) against FP64."""
    api = vpm.api
    P = vpm.workloads.random_cloud(n, seed=1)
    f = api.Field(n, 0)
    f.set_particles(P)
    out = {"N": n}
    for kernel in ("gaussianerf", "winckelmans"):
        r = {}
        rows = {}
        for prec in ("f32c", "f64", "f64far"):
            f.uj_direct(kernel, prec)
            f.synchronize()
            t0 = time.perf_counter()
            k = reps if prec == "f32c" else 1
            for _ in range(k):
                f.uj_direct(kernel, prec)
            f.synchronize()
            dt = (time.perf_counter() - t0) / k
            r["%s_pairs_per_s" % prec] = float(n) * n / dt
            Q = np.zeros_like(P)
            f.get_particles(Q, rows=(9, 24))
            rows[prec] = Q
        A, R = rows["f32c"], rows["f64"]
        S32, S64 = vpm.metrics.stretching(A[:, 15:24], P[:, 3:6]), vpm.metrics.stretching(R[:, 15:24], P[:, 3:6])
        sc = np.linalg.norm(R[:, 15:24], axis=1) * np.linalg.norm(P[:, 3:6], axis=1)
        es = np.linalg.norm(S32 - S64, axis=1) / sc
        r["error_f32c_vs_f64"] = {
            "U": vpm.metrics.err_stats(A[:, 9:12], R[:, 9:12]),
            "J": vpm.metrics.err_stats(A[:, 15:24], R[:, 15:24]),
            "dGamma_dt": vpm.metrics.err_stats(S32, S64),
            "dGamma_dt_vs_J_Gamma_scale": {"max": float(es.max()), "rms": float(np.sqrt(np.mean(es * es)))}}
        r["error_f64far_vs_f64"] = {"U": vpm.metrics.err_stats(rows["f64far"][:, 9:12], R[:, 9:12]),
                                    "J": vpm.metrics.err_stats(rows["f64far"][:, 15:24], R[:, 15:24])}
        if cpu:
            nthreads, info = host_threads()
            rate, nt, dt, Q = cpu_uj_rate(P, kernel, nthreads, 3.0)
            r["cpu"] = {"pairs_per_s": rate, "threads": nthreads, "kind": "port",
                        "sample": "%d targets x %d sources, FP64 oracle, %.1f s" % (nt, n, dt)}
            r["cpu"].update(info)
            r["gpu_f32c_over_cpu"] = r["f32c_pairs_per_s"] / rate
            r["error_f64_vs_cpu_oracle_sample"] = {"U": vpm.metrics.err_stats(R[:nt, 9:12], Q[:nt, 9:12]),
                                                   "J": vpm.metrics.err_stats(R[:nt, 15:24], Q[:nt, 15:24])}
        out[kernel] = r
        log("bench: cfg1 %s %s" % (kernel, r))
    f.close()
    return out


def cfg2_rollout(vpm, nsteps=10, dt=0.01, cpu=True):
    """Two leapfrogging rings, N = 262,144: nsteps rk3 steps of rVPM with the
    dynamic SFS model (SFS_Cd_twolevel_nobackscatter), core spreading
    (nu = 1e-4) and Pedrizzetti relaxation every step; FP32c pair passes, and
    FP64 with the far form, against the FP64 rollout."""
    api = vpm.api
    P = vpm.workloads.leapfrog_cfg2()
    n = P.shape[0]
    sg0 = float(P[0, 6])
    res = {"N": n, "steps": nsteps, "dt": dt}
    finals = {}
    for prec in ("f32c", "f64", "f64far"):
        pf = api.ParticleField(n, kernel="gaussianerf", formulation=api.rVPM, relaxation=api.pedrizzetti,
                               SFS=api.SFS_Cd_twolevel_nobackscatter, viscous=api.CoreSpreading(1e-4, sg0),
                               precision=prec)
        api.add_particles(pf, P)
        log("bench: cfg2 rollout %s" % prec)
        t0 = time.perf_counter()
        api.run_vpm(pf, dt, nsteps)
        res["%s_s" % prec] = time.perf_counter() - t0
        finals[prec] = pf.particles[:n].copy()
        pf.close()
    A, R = finals["f32c"], finals["f64"]
    res["error_f32c_vs_f64"] = {name: vpm.metrics.err_stats(A[:, sl], R[:, sl])
                                for name, sl in (("X", slice(0, 3)), ("Gamma", slice(3, 6)),
                                                 ("sigma", slice(6, 7)))}
    res["error_f64far_vs_f64"] = {name: vpm.metrics.err_stats(finals["f64far"][:, sl], R[:, sl])
                                  for name, sl in (("X", slice(0, 3)), ("Gamma", slice(3, 6)),
                                                   ("sigma", slice(6, 7)))}
    log("bench: cfg2 errors %s; far form %.2f s" % (res["error_f32c_vs_f64"], res["f64far_s"]))
    if cpu:
        # The reference's step with this model evaluates U+J 5 times (3
        # stages, the test filter of stage 1, the relaxation's UJ) and E_str
        # 4 times (rungekutta3, timeintegration.jl:388-461; pseudo3level
        # before/after UJ, subfilterscale.jl:447-673): 50 U+J and 40 E_str
        # all-pairs passes over 10 steps.  One pass of each is timed on a
        # target sample of the initial state and the rollout extrapolated.
        orc = _oracle()
        nthreads, info = host_threads()
        uj_rate, nt, dt, _ = cpu_uj_rate(P, "gaussianerf", nthreads, 4.0)
        Q = P.copy()
        g = api.Field(n, 0)
        g.set_particles(Q)
        g.uj_direct("gaussianerf", "f64")
        g.get_particles(Q)
        g.close()

        def run(m):
            Z = Q.copy()
            Z[:, 39:42] = 0.0
            orc.estr_direct(Z, "gaussianerf", True, nthreads=nthreads, t_range=(0, m))
            return Z

        etps, ent, edt, _ = time_cpu(run, n, 4.0)
        e_rate = etps * n
        pass_pairs = float(n) * n
        est = nsteps * (5 * pass_pairs / uj_rate + 4 * pass_pairs / e_rate)
        res["cpu"] = {"uj_pairs_per_s": uj_rate, "estr_pairs_per_s": e_rate, "threads": nthreads, "kind": "port",
                      "uj_sample": "%d targets x %d sources, %.1f s" % (nt, n, dt),
                      "estr_sample": "%d targets x %d sources, %.1f s" % (ent, n, edt),
                      "rollout_s_extrapolated": est,
                      "extrapolation": "%d steps x (5 U+J + 4 E_str all-pairs passes) at the sampled rates; "
                                       "the O(N) per-particle work is left out" % nsteps}
        res["cpu"].update(info)
        res["gpu_f32c_over_cpu_extrapolated"] = est / res["f32c_s"]
        log("bench: cfg2 CPU %s" % res["cpu"])
    return res


def gpu_uj_time(f, kernel, prec, reps):
    """Call and kernel time (ms) of vpm_uj_direct on a resident field."""
    f.uj_direct(kernel, prec)                     # warmup (sort, staging)
    f.profile(True)
    f.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        f.uj_direct(kernel, prec)
    f.synchronize()
    wall = (time.perf_counter() - t0) / reps
    kms, cnt = f.timing(0 if prec == "f32c" else 1)
    f.profile(False)
    return 1e3 * wall, kms / max(cnt, 1)


def cfg0_ring(vpm, kernel="gaussianerf"):
    """Config 0: one vortex ring, N = 8,000 (Nphi = 200 x 40), one UJ_direct
    U+J evaluation (the reference's CPU-runnable case): GPU F32C and FP64,
    and the FP64 oracle over every target on the host threads."""
    api = vpm.api
    orc = _oracle()
    nthreads, info = host_threads()
    P = vpm.workloads.ring_cfg0()
    n = P.shape[0]
    pairs = float(n) * n
    f = api.Field(n, 0)
    f.set_particles(P)
    out = {"N": n, "kernel": kernel}
    rows = {}
    for prec, reps in (("f32c", 50), ("f64", 10)):
        call_ms, kern_ms = gpu_uj_time(f, kernel, prec, reps)
        Q = np.zeros_like(P)
        f.get_particles(Q, rows=(9, 24))
        rows[prec] = Q
        out[prec] = {"call_ms": call_ms, "kernel_ms": kern_ms, "pairs_per_s_call": pairs / (call_ms / 1e3),
                     "pairs_per_s_kernel": pairs / (kern_ms / 1e3)}
    f.close()
    best = None
    for _ in range(3):
        R = P.copy()
        t0 = time.perf_counter()
        orc.direct(R, kernel, nthreads=nthreads)
        dt = time.perf_counter() - t0
        best = dt if best is None else min(best, dt)
    out["cpu"] = {"s": best, "pairs_per_s": pairs / best, "threads": nthreads, "kind": "port",
                  "sample": "every target, best of 3"}
    out["cpu"].update(info)
    for prec in ("f32c", "f64"):
        A = rows[prec]
        out[prec]["error_vs_cpu_oracle"] = {"U": vpm.metrics.err_stats(A[:, 9:12], R[:, 9:12]),
                                            "J": vpm.metrics.err_stats(A[:, 15:24], R[:, 15:24])}
    out["gpu_f32c_over_cpu"] = out["f32c"]["pairs_per_s_call"] / out["cpu"]["pairs_per_s"]
    log("bench: cfg0 %s" % out)
    return out


def cfg3_adjoint(vpm, kernel="gaussianerf", cpu=True):
    """Config 3: random cloud N = 32,768 (seed 33), weights w ~ N(0,1)^3, W ~
    N(0,1)^(3x3) (seed 34), L = sum w.U + sum W:J.  F32C forward and adjoint
    kernel times on the GPU and their ratio; the FP64 CPU gradient (the
    oracle's adjoint, which equals the reference's own UJS_pullback,
    rrules.jl:183-326, to FP64 rounding) timed on the host threads; the GPU
    gradient's error against it, per particle, relative to the summation
    scale of each output."""
    api = vpm.api
    orc = _oracle()
    nthreads, info = host_threads()
    n = 32768
    P = vpm.workloads.random_cloud(n, seed=33)
    w, W = vpm.workloads.adjoint_weights(n, seed=34)
    f = api.Field(n, 0)
    f.set_particles(P)
    fwd_call, fwd_kern = gpu_uj_time(f, kernel, "f32c", 20)
    f.uj_adjoint(w, W, kernel, "f32c")            # warmup
    f.profile(True)
    f.synchronize()
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        got = f.uj_adjoint(w, W, kernel, "f32c")
    adj_call = 1e3 * (time.perf_counter() - t0) / reps
    ams, acnt = f.timing(3)
    f.profile(False)
    f.close()
    out = {"N": n, "kernel": kernel, "loss": "sum w.U + sum W:J", "forward_kernel_ms": fwd_kern,
           "forward_call_ms": fwd_call, "adjoint_kernel_ms": ams / max(acnt, 1), "adjoint_call_ms": adj_call,
           "ratio_kernel": ams / max(acnt, 1) / fwd_kern, "ratio_call": adj_call / fwd_call}
    if not cpu:
        log("bench: cfg3 %s" % out)
        return out
    t0 = time.perf_counter()
    ref = orc.uj_adjoint(P, kernel, w, W, nthreads=nthreads)
    cpu_s = time.perf_counter() - t0
    scale = orc.uj_adjoint_scale(P, kernel, w, W, nthreads=nthreads)
    errs = {}
    for name, a, r, m in zip(("X", "Gamma", "sigma"), got, ref, scale):
        a, r, m = (np.asarray(x).reshape(n, -1) for x in (a, r, m))
        den = np.linalg.norm(m, axis=1)
        e = np.linalg.norm(a - r, axis=1) / np.maximum(den, 1e-6 * den.max())
        errs[name] = {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e * e)))}
    out.update({"cpu_gradient_s": cpu_s, "cpu_threads": nthreads,
                "cpu_kind": "port (oracle/vpm_oracle.c vpmo_uj_adjoint)",
                "gpu_adjoint_over_cpu": cpu_s * 1e3 / adj_call,
                "error_vs_cpu_gradient_per_summation_scale": errs})
    out.update({"cpu_" + k: v for k, v in info.items()})
    log("bench: cfg3 %s" % out)
    return out


def core_spreading_reset(vpm, P, kernel, itmax=15, tol=1e-3):
    """The CoreSpreading reset on the jet (viscous.jl:202-220, 314-478):
    zeta field -> target, sigma <- sgm0 (the jet's own core size), then the
    RBF conjugate gradient with CoreSpreading's defaults (itmax 15, tol
    1e-3; the iteration count and residuals are reported, not required).
    Also one zeta_direct evaluation alone (F32C, kernel time)."""
    api = vpm.api
    n = P.shape[0]
    g = api.Field(n, 0)
    g.set_particles(P)
    g.zeta_direct(kernel, "f32c")                      # warmup (sort, staging)
    g.profile(True)
    g.synchronize()
    t0 = time.perf_counter()
    g.zeta_direct(kernel, "f32c")
    tz = time.perf_counter() - t0
    kz, cz = g.timing(4)
    g.set_particles(P)
    sgm0 = float(np.median(P[:, 6]))
    g.profile(True)
    t0 = time.perf_counter()
    its, res = g.core_spreading_reset(sgm0, kernel, "f32c", itmax=itmax, tol=tol, iterror=False)
    tr = time.perf_counter() - t0
    kr, cr = g.timing(4)
    g.profile(False)
    g.close()
    out = {"N": n, "zeta_kernel_ms": kz / max(cz, 1), "zeta_call_ms": 1e3 * tz,
           "zeta_pairs_per_s": float(n) * n / (kz / 1e3 / max(cz, 1)),
           "reset_ms": 1e3 * tr, "reset_iterations": its, "reset_residuals": [float(x) for x in res],
           "reset_zeta_evaluations": cr, "reset_zeta_kernel_ms": kr, "sgm0": sgm0,
           "itmax": itmax, "tol": tol}
    log("bench: core-spreading reset %s" % out)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--kernel", default="gaussianerf")
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-f64", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-rollout", action="store_true",
                    help="skip forward+adjoint through 3 RK3 steps on the jet")
    ap.add_argument("--no-leaf", action="store_true",
                    help="skip the FMM leaf-pair list (vpm_p2p_ranges) on the jet")
    ap.add_argument("--no-reset", action="store_true",
                    help="skip zeta_direct and the CoreSpreading RBF reset on the jet")
    ap.add_argument("--no-cfg1", action="store_true",
                    help="skip the N=65,536 random-cloud accuracy/throughput report")
    ap.add_argument("--no-cfg2", action="store_true",
                    help="skip the 10-step leapfrog-rings rollout, FP32c vs FP64")
    ap.add_argument("--no-cfg0", action="store_true", help="skip the N=8,000 ring (GPU and CPU)")
    ap.add_argument("--no-cfg3", action="store_true",
                    help="skip the N=32,768 adjoint/forward ratio and the CPU gradient")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device copy of this size (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: one GPU only; --gpus %d is not supported (the all-pairs path would shard "
              "targets across GPUs with an all-gather of the source records)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from __graft_entry__ import load_package
    vpm = load_package()
    api = vpm.api
    kernel = args.kernel
    log("bench: generating jet column N=%d" % args.n)
    P = vpm.workloads.jet_column(args.n)
    n = P.shape[0]
    pairs = float(n) * float(n)

    f = api.Field(n, 0)
    f.set_particles(P)
    f.profile(True)
    log("bench: warmup %d" % args.warmup)
    for _ in range(args.warmup):
        f.uj_direct(kernel, "f32c")
    f.profile(True)   # reset counters
    f.synchronize()
    log("bench: timing %d F32C steps" % args.steps)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        f.uj_direct(kernel, "f32c")
    f.synchronize()
    wall = time.perf_counter() - t0
    ms_step = 1e3 * wall / args.steps
    kms, kcount = f.timing(0)
    k_avg_s = kms / 1e3 / max(kcount, 1)
    value = pairs / (wall / args.steps)
    achieved = pairs * FLOP_PER_PAIR[kernel] / k_avg_s
    out32 = np.zeros_like(P)
    f.get_particles(out32, rows=(9, 24))
    log("bench: F32C %.1f ms/step, kernel %.1f ms, %.3e pairs/s" % (ms_step, 1e3 * k_avg_s, value))

    res = {
        "metric": METRIC, "value": value, "unit": "pair-interactions/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": ms_step, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "jet column N=%d, one UJ_direct U+J evaluation, %s kernel, FP32 pair "
                               "arithmetic with compensated sums, state resident in HBM" % (n, kernel),
                   "N": n, "kernel": kernel, "pairs_per_step": pairs, "flop_per_pair": FLOP_PER_PAIR[kernel]},
        "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP32 / 1e12, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP32, "traffic": TRAFFIC_BYTES.get((n, kernel)),
                     "kernel": "uj_f32c_kernel<%s>" % kernel, "avg_launch_ms": 1e3 * k_avg_s,
                     "traffic_source": "profiles/uj_f32c_gaussianerf_jet_n1048576.md (rocprofv3 FETCH_SIZE x2.0 + "
                                       "WRITE_SIZE x1.0, copy-kernel calibrated)"},
    }

    if not args.no_f64:
        log("bench: FP64 GPU evaluation")
        f.profile(True)
        t0 = time.perf_counter()
        f.uj_direct(kernel, "f64")
        f.synchronize()
        w64 = time.perf_counter() - t0
        k64, _ = f.timing(1)
        out64 = np.zeros_like(P)
        f.get_particles(out64, rows=(9, 24))
        eu = vpm.metrics.err_stats(out32[:, 9:12], out64[:, 9:12])
        ej = vpm.metrics.err_stats(out32[:, 15:24], out64[:, 15:24])
        S32, S64 = vpm.metrics.stretching(out32[:, 15:24], P[:, 3:6]), vpm.metrics.stretching(out64[:, 15:24], P[:, 3:6])
        es = vpm.metrics.err_stats(S32, S64)
        # dGamma/dt = J^T Gamma nearly vanishes in the column's core; against
        # its summation scale |J| |Gamma| the error shows the kernel, not the data
        sc = np.linalg.norm(out64[:, 15:24], axis=1) * np.linalg.norm(P[:, 3:6], axis=1)
        esc = np.linalg.norm(S32 - S64, axis=1) / np.maximum(sc, 1e-300)
        res["f64"] = {"value": pairs / w64, "unit": "pair-interactions/s", "ms_per_step": 1e3 * w64,
                      "kernel_ms": k64, "speedup_f32c_over_f64": w64 / (wall / args.steps),
                      "frac_of_fp64_peak": pairs * FLOP_PER_PAIR[kernel] / (k64 / 1e3) / PEAK_FP64}
        res["accuracy_f32c_vs_f64_gpu"] = {"U": eu, "J": ej, "dGamma_dt": es,
                                           "dGamma_dt_vs_J_Gamma_scale": {"max": float(esc.max()),
                                                                          "rms": float(np.sqrt(np.mean(esc * esc)))}}
        log("bench: FP64 %.1f ms; err U %s J %s dGamma/dt %s" % (1e3 * w64, eu, ej, es))

        # FP64 with the far form (VPM_UJ_F64_FAR): one rsqrt per pair, g = 1
        # past q = 10, the F32C kernel's spatial order
        log("bench: FP64 far-form GPU evaluation (warmup + 1)")
        f.profile(False)
        f.uj_direct(kernel, "f64far")
        f.synchronize()
        f.profile(True)
        t0 = time.perf_counter()
        f.uj_direct(kernel, "f64far")
        f.synchronize()
        w64f = time.perf_counter() - t0
        k64f, _ = f.timing(1)
        out64f = np.zeros_like(P)
        f.get_particles(out64f, rows=(9, 24))
        res["f64far"] = {"value": pairs / w64f, "unit": "pair-interactions/s", "ms_per_step": 1e3 * w64f,
                         "kernel_ms": k64f, "speedup_f32c_over_f64far": w64f / (wall / args.steps),
                         "speedup_f64far_over_f64": w64 / w64f,
                         "frac_of_fp64_peak": pairs * FLOP_PER_PAIR[kernel] / (k64f / 1e3) / PEAK_FP64,
                         "error_vs_f64": {"U": vpm.metrics.err_stats(out64f[:, 9:12], out64[:, 9:12]),
                                          "J": vpm.metrics.err_stats(out64f[:, 15:24], out64[:, 15:24])},
                         "arithmetic": "FP64 pairs, one rsqrt each, far form F' = r^-3, H'' = r^-5 past q = 10 "
                                       "(g = 1 and q g' = 0 to FP64 rounding), reference g and dg nearer"}
        log("bench: FP64 far form %.1f ms (kernel %.1f ms); vs FP64 U %s" % (1e3 * w64f, k64f,
                                                                             res["f64far"]["error_vs_f64"]["U"]))
    f.profile(False)

    if not args.no_adjoint:
        def adjoint_run(Ubar, Jbar):
            f.uj_adjoint(Ubar, Jbar, kernel, "f32c")            # warm-up
            f.profile(True)
            f.synchronize()
            t0 = time.perf_counter()
            f.uj_adjoint(Ubar, Jbar, kernel, "f32c")
            ta = time.perf_counter() - t0
            ka, _ = f.timing(3)
            f.profile(False)
            return {"N": n, "precision": "f32c", "adjoint_kernel_ms": ka, "forward_kernel_ms": 1e3 * k_avg_s,
                    "ratio_kernel": ka / (1e3 * k_avg_s), "adjoint_call_ms": 1e3 * ta,
                    "forward_call_ms": ms_step, "ratio_call": 1e3 * ta / ms_step}

        # adjoint of the kinetic-energy loss L = sum |U_i|^2 through the same
        # U/J evaluation: Ubar = 2 U, Jbar = 0 (a loss on U alone takes the
        # kernel without the Jbar terms), on the resident jet field
        log("bench: F32C adjoint, L = sum |U|^2 (warmup + 1)")
        r = adjoint_run(2.0 * out32[:, 9:12], np.zeros((n, 9)))
        r.update({"loss": "sum |U_i|^2", "kernel": "adj_f32c_kernel<%s, Jbar = 0>" % kernel})
        res["adjoint"] = r
        log("bench: adjoint kernel %.1f ms (call %.1f ms) vs forward kernel %.1f ms" %
            (r["adjoint_kernel_ms"], r["adjoint_call_ms"], 1e3 * k_avg_s))
        # a loss on U and J: L = sum |U_i|^2 + sum |J_i|^2 (Ubar = 2 U, Jbar = 2 J)
        log("bench: F32C adjoint, L = sum |U|^2 + sum |J|^2 (warmup + 1)")
        r = adjoint_run(2.0 * out32[:, 9:12], 2.0 * out32[:, 15:24])
        r.update({"loss": "sum |U_i|^2 + sum |J_i|^2", "kernel": "adj_f32c_kernel<%s>" % kernel})
        res["adjoint_general"] = r
        log("bench: adjoint (U and J) kernel %.1f ms vs forward kernel %.1f ms" %
            (r["adjoint_kernel_ms"], 1e3 * k_avg_s))

    if not args.no_rollout:
        res["rollout_adjoint"] = rollout_adjoint(vpm, P, kernel)

    if not args.no_leaf:
        res["fmm_leaf_list"] = leaf_list(vpm, P, kernel, value)
        res["fmm_leaf_list_16"] = leaf_list(vpm, P, kernel, value, leaf=16)
        # FMM's default ncrit = 10 (src/FLOWVPM_particlefield.jl:60) gives
        # leaves of at most 10 particles
        res["fmm_leaf_list_8"] = leaf_list(vpm, P, kernel, value, leaf=8)

    if not args.no_reset:
        res["core_spreading_reset"] = core_spreading_reset(vpm, P, kernel)

    if not args.no_cfg0:
        res["cfg0_ring"] = cfg0_ring(vpm)

    if not args.no_cfg1:
        res["cfg1_cloud"] = cfg1_cloud(vpm, cpu=not args.no_cpu)

    if not args.no_cfg2:
        res["cfg2_rollout"] = cfg2_rollout(vpm, cpu=not args.no_cpu)

    if not args.no_cfg3:
        res["cfg3_adjoint"] = cfg3_adjoint(vpm, cpu=not args.no_cpu)

    if not args.no_cpu:
        log("bench: CPU baseline")
        cb, ref_rows, nt = cpu_baseline(P, kernel)
        res["cpu_baseline"] = cb
        eu = vpm.metrics.err_stats(out32[:nt, 9:12], ref_rows[:, 9:12])
        ej = vpm.metrics.err_stats(out32[:nt, 15:24], ref_rows[:, 15:24])
        es = vpm.metrics.err_stats(vpm.metrics.stretching(out32[:nt, 15:24], P[:nt, 3:6]),
                                   vpm.metrics.stretching(ref_rows[:, 15:24], P[:nt, 3:6]))
        res["accuracy_f32c_vs_cpu_oracle"] = {"targets": nt, "U": eu, "J": ej, "dGamma_dt": es}
    f.close()
    if args.copy_mib:
        copy_calibration(args.copy_mib)
        res["copy_calibration_mib"] = args.copy_mib
    print(json.dumps(res))


if __name__ == "__main__":
    main()
