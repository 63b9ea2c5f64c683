"""The FP64 far-form U/J kernel (vpm_uj_direct with VPM_PREC_F64 and
VPM_UJ_F64_FAR, precision "f64far" in the Python API), through the C-ABI.

It evaluates the pair of fmm.jl:113-161 in FP64 with one rsqrt per pair and,
past q = 10 (Gaussian-erf) or q^2 = 13 (Gaussian), the far form g = 1,
q g' = 0, exact there to FP64 rounding (1 - g < 5e-21, q g' < 7e-19); near
pairs take the reference's g and dg.  So it is held to the bounds of the
reference-arithmetic FP64 path: 1e-12 per particle against the oracle and
the mpmath goldens (3e-7 / 3e-5 for the 16 golden particles at q = 1e-4,
where the reference's own FP64 formula loses digits), and 1e-12 against
that path on the jet column, where most pairs take the far form (at 2^20,
U to its summation scale, as the jet's U sums cancel).  At the
far threshold itself the pairs between two clusters agree with the
reference arithmetic to a few ulps on either side."""
import ctypes as C

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu
KERNELS = ["singular", "gaussian", "gaussianerf", "winckelmans"]
F64_TOL = 1e-12


def _say(*a):
    print(*a, flush=True)


def _uj_err(vpm, A, R):
    return (vpm.metrics.err_stats(A[:, 9:12], R[:, 9:12]), vpm.metrics.err_stats(A[:, 15:24], R[:, 15:24]))


@pytest.mark.parametrize("kernel", KERNELS)
def test_golden_uj_f64far(vpm, kernel):
    _say("golden f64far", kernel)
    g = golden("uj_%s.json" % kernel)
    P = vpm.fields.make_particles(g["X"], g["Gamma"], g["sigma"])
    vpm.UJ_hip(P, kernel=kernel, precision="f64far")
    U, J = np.array(g["U"]), np.array(g["J"])
    eu = vpm.metrics.err_stats(P[:, 9:12], U)
    ej = vpm.metrics.err_stats(P[:, 15:24], J)
    _say("  U", eu, "J", ej)
    loose = {"gaussian": 3e-5, "gaussianerf": 3e-7}.get(kernel, 1e-11)
    assert eu["max"] < loose and ej["max"] < loose
    eu2 = vpm.metrics.err_stats(P[16:, 9:12], U[16:])
    ej2 = vpm.metrics.err_stats(P[16:, 15:24], J[16:])
    assert eu2["max"] < F64_TOL and ej2["max"] < F64_TOL


def test_two_body_check_fmm_f64far(vpm):
    g = golden("check_fmm.json")
    P = vpm.fields.make_particles(g["X"], g["Gamma"], g["sigma"])
    vpm.UJ_hip(P, kernel="singular", precision="f64far")
    eu = vpm.metrics.err_stats(P[:, 9:12], g["U"])
    ej = vpm.metrics.err_stats(P[:, 15:24], g["J"])
    assert eu["max"] < 1e-13 and ej["max"] < 1e-13, (eu, ej)


@pytest.mark.parametrize("kernel", KERNELS)
def test_cloud_vs_oracle_and_reference_arithmetic(vpm, orc, kernel):
    P = vpm.workloads.random_cloud(3000, seed=11)
    ref = P.copy()
    orc.uj_direct(ref, kernel)
    A, B = P.copy(), P.copy()
    vpm.UJ_hip(A, kernel=kernel, precision="f64far")
    vpm.UJ_hip(B, kernel=kernel, precision="f64")
    eu, ej = _uj_err(vpm, A, ref)
    eu2, ej2 = _uj_err(vpm, A, B)
    _say("cloud N=3000", kernel, "f64far vs oracle U", eu, "J", ej, "vs f64 U", eu2, "J", ej2)
    assert eu["max"] < F64_TOL and ej["max"] < F64_TOL
    assert eu2["max"] < F64_TOL and ej2["max"] < F64_TOL


def test_edge_sizes_coincident_and_storage_order(vpm, orc):
    """Sizes around the wave and tile widths, coincident particles (skipped,
    fmm.jl:118), and the storage-order form (no spatial sort)."""
    for n in (1, 2, 63, 65, 255, 257, 700):
        P = vpm.workloads.random_cloud(n, seed=n)
        if n > 8:
            P[n // 2, 0:3] = P[1, 0:3]           # a coincident pair
        for kernel in ("gaussianerf", "gaussian", "winckelmans"):
            ref = P.copy()
            orc.uj_direct(ref, kernel)
            for sort in (True, False):
                A = P.copy()
                vpm.UJ_hip(A, kernel=kernel, precision="f64far", sort=sort)
                assert np.all(np.isfinite(A[:, 9:24]))
                if n == 1:
                    assert np.all(A[:, 9:24] == 0)
                    continue
                eu, ej = _uj_err(vpm, A, ref)
                assert eu["max"] < F64_TOL and ej["max"] < F64_TOL, (n, kernel, sort, eu, ej)


@pytest.mark.parametrize("kernel,qf", [("gaussianerf", 10.0), ("gaussian", 13.0 ** 0.5)])
def test_far_threshold_pairs(vpm, kernel, qf):
    """Two clusters of 64 coincident particles (coincident pairs are skipped,
    so each wave's target box is one point) at q just inside and just
    outside the far threshold and well on either side of it: every pair
    between the clusters takes the near form or the far form, against the
    reference arithmetic to a few ulps."""
    sig = 0.02
    rng = np.random.default_rng(5)
    G = rng.normal(size=(128, 3))
    e = np.array([1.0, -1.0, 1.0]) / 3 ** 0.5
    for q in (qf * (1 - 1e-9), qf * (1 + 1e-9), 0.5 * qf, 3 * qf):
        X = np.vstack([np.tile([0.1, 0.2, 0.3], (64, 1)), np.tile(np.array([0.1, 0.2, 0.3]) + q * sig * e, (64, 1))])
        P = vpm.fields.make_particles(X, G, np.full(128, sig))
        A, B = P.copy(), P.copy()
        vpm.UJ_hip(A, kernel=kernel, precision="f64far")
        vpm.UJ_hip(B, kernel=kernel, precision="f64")
        eu, ej = _uj_err(vpm, A, B)
        _say(kernel, "q/q_far %.9f" % (q / qf), "U", eu["max"], "J", ej["max"])
        assert np.all(np.abs(B[:, 9:12]).max(axis=1) > 0)
        assert eu["max"] < 1e-14 and ej["max"] < 1e-14, (q, eu, ej)


def test_jet_column_far_form(vpm, orc):
    """The jet column at 2^16 particles (sigma = 0.015 at a spacing of ~0.036:
    almost every pair beyond q = 10), Gaussian-erf: every particle against
    the reference-arithmetic FP64 path, a sample against the oracle."""
    P = vpm.workloads.jet_column(1 << 16)
    A, B = P.copy(), P.copy()
    vpm.UJ_hip(A, kernel="gaussianerf", precision="f64far")
    vpm.UJ_hip(B, kernel="gaussianerf", precision="f64")
    eu, ej = _uj_err(vpm, A, B)
    idx = np.arange(0, P.shape[0], 97)
    sub = P[idx].copy()
    sub[:, 9:24] = 0
    orc.direct(sub, "gaussianerf", source=P)
    eu2, ej2 = _uj_err(vpm, A[idx], sub)
    _say("jet 2^16 f64far vs f64 U", eu, "J", ej, "vs oracle sample U", eu2, "J", ej2)
    assert eu["max"] < F64_TOL and ej["max"] < F64_TOL
    assert eu2["max"] < F64_TOL and ej2["max"] < F64_TOL


def test_full_size_jet_f64far(vpm, orc):
    """The 2^20 jet column, Gaussian-erf.  The far-form path sums in another
    order than the reference-arithmetic path (sorted sources, the 14 sums of
    the F32C kernel), and the jet's U sums cancel, so U is held to its
    summation scale s_i = sum_j |U term_ij| (`oracle.u_scale_ranges`) on four
    windows of 64 targets against the oracle, as is the reference-arithmetic
    path.  Each side sums 2^20 terms (the oracle in one sequential chain),
    whose rounding walks to ~sqrt(2^20) eps = 1.1e-13 of the scale (both GPU
    paths measured 1.0e-13): |dU_i| <= 5e-13 s_i, and the far-form path no
    worse than 1.5 times the reference-arithmetic one.  Over every particle:
    J within 1e-12 of the reference-arithmetic path, U within 1e-13 RMS."""
    kern = "gaussianerf"
    P = vpm.workloads.jet_column(1 << 20)
    n = P.shape[0]
    A, B = P.copy(), P.copy()
    vpm.UJ_hip(A, kernel=kern, precision="f64far")
    vpm.UJ_hip(B, kernel=kern, precision="f64")
    eu, ej = _uj_err(vpm, A, B)
    _say("jet 2^20 f64far vs f64 U", eu, "J", ej)
    assert eu["rms"] < 1e-13 and ej["max"] < F64_TOL, (eu, ej)
    worst = {"f64far": 0.0, "f64": 0.0}
    for t0 in (0, n // 3, n // 2 + 12345, n - 64):
        t1 = t0 + 64
        ref = P[t0:t1].copy()
        ref[:, 9:24] = 0
        orc.direct(ref, kern, source=P)
        s = orc.u_scale_ranges(P, kern, [t0], [t1], [0], [n])[t0:t1]
        for name, X in (("f64far", A), ("f64", B)):
            e = np.linalg.norm(X[t0:t1, 9:12] - ref[:, 9:12], axis=1)
            worst[name] = max(worst[name], float((e / s).max()))
            ejw = vpm.metrics.err_stats(X[t0:t1, 15:24], ref[:, 15:24])
            assert ejw["max"] < F64_TOL, (name, t0, ejw)
    _say("jet 2^20 U err/scale vs oracle windows: f64far %.2e, f64 %.2e" % (worst["f64far"], worst["f64"]))
    assert worst["f64far"] < 5e-13 and worst["f64"] < 5e-13, worst
    assert worst["f64far"] <= 1.5 * worst["f64"], worst


def test_f64far_with_sfs_and_accumulate(vpm):
    """UJ_direct's other flags with the far form: E_str (sfs) runs in FP64 on
    the far-form J, and reset=False accumulates onto what is there."""
    P = vpm.workloads.random_cloud(2000, seed=21)
    A, B = P.copy(), P.copy()
    vpm.UJ_hip(A, kernel="gaussianerf", precision="f64far", sfs=True, reset_sfs=True)
    vpm.UJ_hip(B, kernel="gaussianerf", precision="f64", sfs=True, reset_sfs=True)
    eu, ej = _uj_err(vpm, A, B)
    es = vpm.metrics.err_stats(A[:, 39:42], B[:, 39:42])
    _say("sfs: U", eu["max"], "J", ej["max"], "E_str", es)
    assert eu["max"] < F64_TOL and ej["max"] < F64_TOL and es["max"] < 1e-10, (eu, ej, es)
    C2 = A.copy()
    vpm.UJ_hip(C2, kernel="gaussianerf", precision="f64far", reset=False)
    assert np.allclose(C2[:, 9:12], 2 * A[:, 9:12], rtol=1e-13, atol=0), np.abs(C2[:, 9:12] - 2 * A[:, 9:12]).max()


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("transposed", [True, False])
def test_estr_f64far_vs_oracle(vpm, orc, kernel, transposed):
    """E_str under the far form (UJ with sfs): FP64 pairs within q_cut only
    (Gaussian-erf q = 10, Gaussian q^2 = 14, where zeta has fallen to 2e-22 of
    zeta(0); Winckelmans every pair; singular the coincident pairs), in the
    U/J kernels' spatial order, on a cloud with coincident and static
    particles.  Against the oracle's Estr_direct! on the same J, per particle
    to 1e-12 of the summation scale s_i = sum_j zeta |S_ij|, floored at 1e-6
    of its largest value: a particle whose every pair lies near or past q_cut
    (isolated in this sparse cloud: the Gaussian's zeta(3.7) is e^-52 of
    zeta(0)) has an E_str of that size, which the cut drops."""
    P = vpm.workloads.random_cloud(3000, seed=31)
    P[100, 0:3] = P[7, 0:3]                          # coincident pairs
    P[2000, 0:3] = P[7, 0:3]
    P[[5, 50, 500], 42] = 1.0                        # static particles
    A = P.copy()
    vpm.UJ_hip(A, kernel=kernel, precision="f64far", sfs=True, reset_sfs=True, transposed=transposed)
    R = A.copy()
    R[:, 39:42] = 0.0
    orc.estr_direct(R, kernel, transposed=transposed)
    scale = orc.estr_scale(A, kernel, transposed=transposed)
    e = np.linalg.norm(A[:, 39:42] - R[:, 39:42], axis=1)
    ratio = e / np.maximum(scale, 1e-6 * scale.max()) if scale.max() > 0 else e
    _say("estr f64far", kernel, transposed, "err/scale max %.2e" % ratio.max())
    assert np.all(A[[5, 50, 500], 39:42] == 0)
    assert ratio.max() < 1e-12, ratio.max()


def test_estr_f64far_edge_sizes(vpm, orc):
    """The far-cut E_str at sizes around the wave and tile widths."""
    for n in (1, 2, 63, 65, 257):
        P = vpm.workloads.random_cloud(n, seed=40 + n)
        P[:, 6] *= 3.0                                   # overlapping cores
        A = P.copy()
        vpm.UJ_hip(A, kernel="gaussianerf", precision="f64far", sfs=True, reset_sfs=True)
        R = A.copy()
        R[:, 39:42] = 0.0
        orc.estr_direct(R, "gaussianerf", transposed=True)
        scale = orc.estr_scale(A, "gaussianerf", transposed=True)
        e = np.linalg.norm(A[:, 39:42] - R[:, 39:42], axis=1)
        if n == 1:
            assert np.all(A[:, 39:42] == 0)
            continue
        assert (e / np.maximum(scale, 1e-6 * scale.max())).max() < 1e-12, n


def test_estr_f64far_jet_column(vpm, orc):
    """The jet column at 2^16 particles, Gaussian-erf (most source tiles
    beyond q_cut of a block): E_str under the far form against the
    reference-arithmetic FP64 E_str on the same J, per particle to 1e-12 of
    the summation scale floored as above (oracle, on a sample of targets)."""
    P = vpm.workloads.jet_column(1 << 16)
    A = P.copy()
    vpm.UJ_hip(A, kernel="gaussianerf", precision="f64far", sfs=True, reset_sfs=True)
    B = A.copy()
    B[:, 39:42] = 0.0
    vpm.Estr_hip(B, kernel="gaussianerf", precision="f64", transposed=True)
    idx = np.arange(0, P.shape[0], 61)
    t = [np.asarray(idx, dtype=np.int64), np.asarray(idx + 1, dtype=np.int64),
         np.zeros(len(idx), dtype=np.int64), np.full(len(idx), P.shape[0], dtype=np.int64)]
    scale = orc.estr_scale_ranges(A, "gaussianerf", *t)[idx]
    e = np.linalg.norm(A[idx, 39:42] - B[idx, 39:42], axis=1)
    ratio = e / np.maximum(scale, 1e-6 * scale.max())
    _say("estr f64far jet 2^16: err/scale max %.2e" % ratio.max())
    assert ratio.max() < 1e-12


def test_f64far_argument_checks(vpm):
    """VPM_UJ_F64_FAR needs VPM_PREC_F64, in vpm_uj_direct and in the step's
    uj_flags (which take no other bit), and the no-allocation entry point does
    not take it: VPM_EINVAL before any device work."""
    api = vpm.api
    P = vpm.workloads.random_cloud(100, seed=3)
    f = api.Field(100, 0)
    try:
        f.set_particles(P)
        for prec, fn in (("f32c", api.lib().vpm_uj_direct), ("f64", api.lib().vpm_uj_direct_noalloc)):
            cfg = api.make_config("gaussianerf", prec, True)
            rc = fn(f.handle, C.byref(cfg), api._lib.UJ_RESET | api._lib.UJ_F64_FAR)
            msg = api.lib().vpm_last_error()
            _say(prec, rc, msg)
            assert rc == api._lib.VPM_EINVAL and b"VPM_UJ_F64_FAR" in msg, (rc, msg)
        for prec, flags, what in (("f32c", api._lib.UJ_F64_FAR, b"VPM_UJ_F64_FAR"), ("f64", 4, b"uj_flags")):
            s = api._lib.StepConfig()
            s.uj = api.make_config("gaussianerf", prec, True)
            s.f, s.g, s.relax, s.uj_flags = 0.0, 0.2, 0, flags
            rc = api.lib().vpm_rk3_step(f.handle, C.byref(s), 0.01, 0)
            msg = api.lib().vpm_last_error()
            assert rc == api._lib.VPM_EINVAL and what in msg, (rc, msg)
        assert api.make_config("gaussianerf", "f64far", True).precision == api._lib.PRECISIONS["f64"]
    finally:
        f.close()


def test_rk3_steps_f64far_vs_f64(vpm, orc):
    """Two rungekutta3 steps of the config-2 physics (rVPM, the two-level
    dynamic SFS, CoreSpreading, Pedrizzetti relaxation) on a reduced leapfrog
    pair, with the step's U/J in the FP64 far form (ParticleField precision
    "f64far"), against the oracle's steps: the FP64 bound of the
    reference-arithmetic step, 1e-10 on the state and 1e-8 on C."""
    api = vpm.api
    P = vpm.workloads.leapfrog_cfg2(Nphi=64)                # 8,192 particles
    sg0 = float(P[0, 6])
    pf = api.ParticleField(P.shape[0], kernel="gaussianerf", formulation=api.rVPM, relaxation=api.pedrizzetti,
                           SFS=api.SFS_Cd_twolevel_nobackscatter, viscous=api.CoreSpreading(1e-4, sg0),
                           precision="f64far")
    api.add_particles(pf, P)
    s = orc.make_settings(kernel="gaussianerf", f=0.0, g=0.2, relax=1, rlxf=0.3, sfs=1, sfs_alpha=0.999,
                          sfs_rlxf=0.005, sfs_minC=0.0, sfs_maxC=1.0, sfs_force_positive=True,
                          sfs_clip=True, viscous=1, nu=1e-4, sgm0=sg0, beta=1.5)
    R = P.copy()
    for step in range(2):
        assert orc.rk3_step(R, s, 0.01, relax_now=True) == 0
        api.nextstep(pf, 0.01, relax=True)
    A = pf.particles[:pf.np]
    e = {name: vpm.metrics.err_stats(A[:, sl], R[:, sl])["max"]
         for name, sl in (("X", slice(0, 3)), ("Gamma", slice(3, 6)), ("sigma", slice(6, 7)))}
    Ce = vpm.metrics.err_stats(A[:, 36:39], R[:, 36:39])["max"]
    _say("rk3 f64far", e, "C", Ce)
    assert max(e.values()) < 1e-10 and Ce < 1e-8, (e, Ce)
