"""GPU parity of the U/J, E_str and leaf-pair paths, through the C-ABI.

Tolerances (per-particle relative error, see flowvpm.jl_amd/metrics.py):
  F32C vs oracle/golden: max <= 1e-5 (the stated FP32c bound)
  F64  vs oracle:        max <= 1e-12 (same FP64 arithmetic, other summation order)
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu
KERNELS = ["singular", "gaussian", "gaussianerf", "winckelmans"]
F32_TOL = 1e-5
F64_TOL = 1e-12
# E_str in FP32 is held to its summation scale s_i = sum_j zeta |S_ij|
# (oracle.estr_scale): |E32_i - E64_i| <= ESTR_SCALE_TOL s_i per particle.
# Budget: a pair term carries <= ~30 eps of relative rounding (zeta =
# b exp2(a r^2): the FP32 argument rounding times |a r^2| <= 35 near the
# cut, plus a, b and the products), FP32 chains of <= 32 terms add <= 32 eps
# of the partial sums, and the factored form J'_i^T A - B sums two parts
# that can each reach a few times s_i: <= ~180 eps s_i = 1.1e-5 s_i; the
# bound keeps a factor 4.5 above that.  On random J the sums cancel by
# factors of 10-100, so the plain relative error is only reported (its RMS
# is held to ESTR_RMS_TOL).
ESTR_SCALE_TOL = 5e-5
ESTR_RMS_TOL = 1e-5


def _estr_ratio(A, ref, scale):
    """|E32_i - E64_i| / max(s_i, 1e-6 max_k s_k): the scale takes the same
    floor as the relative error (metrics.py).  Below it sit particles whose
    every pair lies near or past the kernel's cut (q = 7 Gaussian-erf, 3
    Gaussian), where zeta is below FP32 resolution of the self term and the
    F32C kernel drops what the FP64 sum still adds."""
    e = np.linalg.norm(A - ref, axis=1)
    den = np.maximum(scale, 1e-6 * scale.max())
    return np.where(den > 0, e / np.where(den > 0, den, 1.0), np.where(e > 0, np.inf, 0.0))


def _estr_check(vpm, A, ref, scale, what):
    ratio = _estr_ratio(A, ref, scale)
    st = vpm.metrics.err_stats(A, ref)
    _say(what, "err/scale max %.2e" % ratio.max(), "relative", st)
    assert ratio.max() < ESTR_SCALE_TOL and st["rms"] < ESTR_RMS_TOL, (ratio.max(), st)


def _say(*a):
    print(*a, flush=True)


def _uj_err(vpm, A, R):
    eu = vpm.metrics.err_stats(A[:, 9:12], R[:, 9:12])
    ej = vpm.metrics.err_stats(A[:, 15:24], R[:, 15:24])
    return eu, ej


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_golden_uj(vpm, kernel, precision):
    _say("golden", kernel, precision)
    g = golden("uj_%s.json" % kernel)
    P = vpm.fields.make_particles(g["X"], g["Gamma"], g["sigma"])
    vpm.UJ_hip(P, kernel=kernel, precision=precision)
    U, J = np.array(g["U"]), np.array(g["J"])
    eu = vpm.metrics.err_stats(P[:, 9:12], U)
    ej = vpm.metrics.err_stats(P[:, 15:24], J)
    _say("  U", eu, "J", ej)
    if precision == "f32c":
        # series forms keep FP32 accuracy at q = 1e-4 (the reference's FP64
        # formula itself loses digits there for the Gaussian kernels)
        assert eu["max"] < F32_TOL and ej["max"] < F32_TOL
    else:
        loose = {"gaussian": 3e-5, "gaussianerf": 3e-7}.get(kernel, 1e-11)
        assert eu["max"] < loose and ej["max"] < loose
        eu2 = vpm.metrics.err_stats(P[16:, 9:12], U[16:])
        ej2 = vpm.metrics.err_stats(P[16:, 15:24], J[16:])
        assert eu2["max"] < 1e-12 and ej2["max"] < 1e-12


def test_two_body_check_fmm(vpm):
    g = golden("check_fmm.json")
    P = vpm.fields.make_particles(g["X"], g["Gamma"], g["sigma"])
    for prec, tol in (("f64", 1e-13), ("f32c", 1e-6)):
        A = P.copy()
        vpm.UJ_hip(A, kernel="singular", precision=prec)
        eu = vpm.metrics.err_stats(A[:, 9:12], g["U"])
        ej = vpm.metrics.err_stats(A[:, 15:24], g["J"])
        assert eu["max"] < tol and ej["max"] < tol, (prec, eu, ej)


@pytest.mark.parametrize("kernel", KERNELS)
def test_cloud_vs_oracle(vpm, orc, kernel):
    P = vpm.workloads.random_cloud(3000, seed=11)
    ref = P.copy()
    orc.uj_direct(ref, kernel)
    for prec in ("f64", "f32c"):
        _say("cloud N=3000", kernel, prec)
        A = P.copy()
        vpm.UJ_hip(A, kernel=kernel, precision=prec)
        eu, ej = _uj_err(vpm, A, ref)
        _say("  U", eu, "J", ej)
        tol = F64_TOL if prec == "f64" else F32_TOL
        assert eu["max"] < tol and ej["max"] < tol


def test_ring_cfg0(vpm, orc):
    """Config 0: the N=8,000 ring, Gaussian-erf, one U+J evaluation."""
    P = vpm.workloads.ring_cfg0()
    ref = P.copy()
    orc.uj_direct(ref, "gaussianerf")
    for prec in ("f64", "f32c"):
        A = P.copy()
        vpm.UJ_hip(A, kernel="gaussianerf", precision=prec)
        eu, ej = _uj_err(vpm, A, ref)
        _say("ring cfg0", prec, "U", eu, "J", ej)
        tol = F64_TOL if prec == "f64" else F32_TOL
        assert eu["max"] < tol and ej["max"] < tol


def test_sort_and_storage_order_agree(vpm, orc):
    P = vpm.workloads.random_cloud(1500, seed=2)
    ref = P.copy()
    orc.uj_direct(ref, "gaussianerf")
    A, B = P.copy(), P.copy()
    vpm.UJ_hip(A, kernel="gaussianerf", sort=True)
    vpm.UJ_hip(B, kernel="gaussianerf", sort=False)
    for X in (A, B):
        eu, ej = _uj_err(vpm, X, ref)
        assert eu["max"] < F32_TOL and ej["max"] < F32_TOL


def test_edge_sizes(vpm, orc):
    for n in (1, 2, 63, 255, 257, 700):
        P = vpm.workloads.random_cloud(n, seed=n)
        ref = P.copy()
        orc.uj_direct(ref, "winckelmans")
        for prec, tol in (("f64", F64_TOL), ("f32c", F32_TOL)):
            A = P.copy()
            vpm.UJ_hip(A, kernel="winckelmans", precision=prec)
            if n == 1:
                assert np.all(A[:, 9:24] == 0)
                continue
            eu, ej = _uj_err(vpm, A, ref)
            assert eu["max"] < tol and ej["max"] < tol, (n, prec, eu, ej)


def test_reset_accumulate_static_coincident(vpm, orc):
    """UJ_direct semantics: reset skips static particles; reset=false
    accumulates; coincident pairs are skipped (fmm.jl:118)."""
    P = vpm.workloads.random_cloud(600, seed=5)
    P[7, 0:3] = P[3, 0:3]                  # coincident pair
    P[[10, 20], 42] = 1.0                  # static
    P[:, 9:27] = 0.25                      # stale U/J/vorticity/PSE
    for prec, tol in (("f64", F64_TOL), ("f32c", F32_TOL)):
        ref = P.copy()
        orc.uj_direct(ref, "gaussianerf")
        orc.uj_direct(ref, "gaussianerf", reset=False)
        A = P.copy()
        vpm.UJ_hip(A, kernel="gaussianerf", precision=prec)
        vpm.UJ_hip(A, kernel="gaussianerf", precision=prec, reset=False)
        eu, ej = _uj_err(vpm, A, ref)
        assert eu["max"] < tol and ej["max"] < tol, (prec, eu, ej)
        np.testing.assert_array_equal(A[:, 12:15], ref[:, 12:15])   # vorticity rows
        np.testing.assert_array_equal(A[:, 24:27], ref[:, 24:27])   # PSE rows


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("transposed", [True, False])
def test_estr_vs_oracle(vpm, orc, kernel, transposed):
    P = vpm.workloads.random_cloud(2000, seed=8)
    P[:, 6] *= 3.0                         # overlapping cores: E_str is local
    P[[5, 6], 42] = 1.0
    orc.uj_direct(P, kernel)               # J from the oracle, shared by both
    ref = P.copy()
    orc.estr_direct(ref, kernel, transposed=transposed)
    scale = orc.estr_scale(P, kernel, transposed=transposed)
    for prec in ("f64", "f32c"):
        A = P.copy()
        vpm.Estr_hip(A, kernel=kernel, precision=prec, transposed=transposed)
        if prec == "f64" or kernel == "singular":
            st = vpm.metrics.err_stats(A[:, 39:42], ref[:, 39:42])
            _say("estr", kernel, transposed, prec, st)
            assert st["max"] < F64_TOL, st
        else:
            _estr_check(vpm, A[:, 39:42], ref[:, 39:42], scale, "estr %s %s %s" % (kernel, transposed, prec))
        assert np.all(A[[5, 6], 39:42] == P[[5, 6], 39:42])


def test_estr_golden(vpm, orc):
    g = golden("estr.json")
    for kernel in KERNELS:
        for tr in (True, False):
            P = vpm.fields.make_particles(g["X"], g["Gamma"], g["sigma"], static=g["static"])
            P[:, 15:24] = np.array(g["J"])
            scale = orc.estr_scale(P, kernel, transposed=tr)
            vpm.Estr_hip(P, kernel=kernel, precision="f32c", transposed=tr)
            ref = np.array(g["%s_%s" % (kernel, "T" if tr else "C")])
            if kernel == "singular":      # FP64 kernel (zeta is a delta)
                assert vpm.metrics.err_stats(P[:, 39:42], ref)["max"] < 1e-12
            else:
                _estr_check(vpm, P[:, 39:42], ref, scale, "estr golden %s %s" % (kernel, tr))


def test_uj_with_sfs_flag(vpm, orc):
    P = vpm.workloads.random_cloud(1200, seed=4)
    P[:, 6] *= 3.0
    ref = P.copy()
    orc.uj_direct(ref, "gaussianerf", reset_sfs=True, sfs=True)
    A = P.copy()
    vpm.UJ_hip(A, kernel="gaussianerf", reset_sfs=True, sfs=True)
    _say("uj+sfs vs oracle: SFS", vpm.metrics.err_stats(A[:, 39:42], ref[:, 39:42]))
    eu, ej = _uj_err(vpm, A, ref)
    assert eu["max"] < F32_TOL and ej["max"] < F32_TOL
    # E_str of this call inherits the FP32c error of J; the E_str kernel
    # itself is checked on the J it was given
    ref2 = A.copy()
    ref2[:, 39:42] = 0.0
    orc.estr_direct(ref2, "gaussianerf")
    _estr_check(vpm, A[:, 39:42], ref2[:, 39:42], orc.estr_scale(ref2, "gaussianerf"), "uj+sfs E_str")


@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_p2p_ranges_vs_oracle(vpm, orc, precision):
    """Leaf-pair lists, including overlapping target ranges and repeats."""
    P = vpm.workloads.random_cloud(900, seed=6)
    rng = np.random.default_rng(3)
    cuts = np.sort(np.concatenate([[0, 900], rng.choice(np.arange(1, 900), 20, replace=False)]))
    leaves = list(zip(cuts[:-1], cuts[1:]))
    tb, te, sb, se = [], [], [], []
    for (a, b) in leaves:
        for (c, d) in leaves:
            if rng.random() < 0.4:
                tb.append(a); te.append(b); sb.append(c); se.append(d)
    tb += [5, 100]; te += [300, 120]; sb += [0, 40]; se += [50, 41]     # overlaps
    ref = P.copy()
    orc.direct_ranges(ref, "gaussianerf", tb, te, sb, se)
    A = P.copy()
    vpm.p2p_ranges(A, tb, te, sb, se, kernel="gaussianerf", precision=precision)
    eu, ej = _uj_err(vpm, A, ref)
    tol = F64_TOL if precision == "f64" else F32_TOL
    assert eu["max"] < tol and ej["max"] < tol, (eu, ej)


def test_ranges_reject_out_of_bounds(vpm):
    P = vpm.workloads.random_cloud(100, seed=1)
    with pytest.raises(vpm.VPMError):
        vpm.p2p_ranges(P, [0], [101], [0], [10])
    with pytest.raises(vpm.VPMError):
        vpm.p2p_ranges(P, [5], [3], [0], [10])


def test_two_field_form(vpm, orc):
    S = vpm.workloads.random_cloud(700, seed=9)
    T = vpm.workloads.random_cloud(500, seed=10)
    T[:, 9:24] = 0.5
    ref = T.copy()
    orc.direct(ref, "gaussianerf", source=S)
    A = T.copy()
    vpm.UJ_hip_fields(S, A, kernel="gaussianerf")
    eu, ej = _uj_err(vpm, A, ref)
    assert eu["max"] < F32_TOL and ej["max"] < F32_TOL


def test_rejects_bad_sigma(vpm):
    P = vpm.workloads.random_cloud(10, seed=1)
    P[3, 6] = 0.0
    with pytest.raises(vpm.VPMError):
        vpm.UJ_hip(P)
    P[3, 6] = np.nan
    with pytest.raises(vpm.VPMError):
        vpm.UJ_hip(P)


def test_deterministic(vpm):
    P = vpm.workloads.random_cloud(5000, seed=12)
    A, B = P.copy(), P.copy()
    vpm.UJ_hip(A)
    vpm.UJ_hip(B)
    np.testing.assert_array_equal(A, B)


def test_cfg1_full_size(vpm, orc):
    """Config 1 at full size (N=65,536), Gaussian-erf and Winckelmans: FP64
    against the oracle on a sample of 1,024 targets, and FP32c against that
    FP64 result on every one of the 65,536 targets (U and J within 1e-5 per
    particle).  The cloud's U sums cancel by up to ~2,400x at some particles
    (sum |term| ~ 28 against |U| ~ 0.011), which is what the near-pair folds
    of the F32C kernel are for.  dGamma/dt = J^T Gamma is reported against
    the FP64 result and against its scale |J| |Gamma|."""
    P = vpm.workloads.random_cloud(65536, seed=1)
    idx = np.arange(0, 65536, 64)
    for kernel in ("gaussianerf", "winckelmans"):
        ref = P.copy()
        ref[:, 9:24] = 0
        sub = ref[idx].copy()
        orc.direct(sub, kernel, source=P)
        B = P.copy()
        vpm.UJ_hip(B, kernel=kernel, precision="f64")
        eu, ej = _uj_err(vpm, B[idx], sub)
        _say("cfg1", kernel, "f64 vs oracle sample U", eu, "J", ej)
        assert eu["max"] < F64_TOL and ej["max"] < F64_TOL
        A = P.copy()
        vpm.UJ_hip(A, kernel=kernel, precision="f32c")
        assert np.all(np.isfinite(A[:, 9:24]))
        eu, ej = _uj_err(vpm, A, B)
        G = P[:, 3:6]
        S32, S64 = vpm.metrics.stretching(A[:, 15:24], G), vpm.metrics.stretching(B[:, 15:24], G)
        es = vpm.metrics.err_stats(S32, S64)
        scale = np.linalg.norm(B[:, 15:24], axis=1) * np.linalg.norm(G, axis=1)
        es_scaled = float((np.linalg.norm(S32 - S64, axis=1) / scale).max())
        _say("cfg1", kernel, "f32c vs f64, all targets: U", eu, "J", ej, "dGamma/dt", es,
             "dGamma/dt vs |J||Gamma| max %.2e" % es_scaled)
        assert eu["max"] < F32_TOL and ej["max"] < F32_TOL, (eu, ej)
        assert es_scaled < F32_TOL


@pytest.mark.parametrize("leaf", [5, 8, 16, 32, 64, 200])
@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_p2p_ranges_leaf_lists(vpm, orc, leaf, precision):
    """FMM-style direct lists: target ranges are leaves (identical or
    disjoint), each target leaf's pairs listed together (the linear-time
    plan), then the same pairs with one target leaf split into two runs (the
    general interval plan).  Leaves of up to 64 take the wave kernel (leaves
    of <= 8, 16, 32 share each wave's lanes 8, 4, 2 ways), leaves of 200 the
    block kernel."""
    n = 1500
    P = vpm.workloads.random_cloud(n, seed=26)
    P[:, 6] *= 3.0
    starts = np.arange(0, n, leaf)
    ends = np.minimum(starts + leaf, n)
    C = np.array([P[a:b, 0:3].mean(0) for a, b in zip(starts, ends)])
    tb, te, sb, se = [], [], [], []
    for i in range(len(starts)):
        for j in range(len(starts)):
            if np.linalg.norm(C[i] - C[j]) < 0.35:
                tb.append(starts[i]); te.append(ends[i]); sb.append(starts[j]); se.append(ends[j])
    for lists in ((tb, te, sb, se), (tb + tb[:1], te + te[:1], sb + sb[:1], se + se[:1])):
        ref = P.copy()
        orc.direct_ranges(ref, "gaussianerf", *lists)
        A = P.copy()
        vpm.p2p_ranges(A, *lists, kernel="gaussianerf", precision=precision)
        eu, ej = _uj_err(vpm, A, ref)
        tol = F64_TOL if precision == "f64" else F32_TOL
        assert eu["max"] < tol and ej["max"] < tol, (leaf, eu, ej)


@pytest.mark.parametrize("leaf", [5, 16])
@pytest.mark.parametrize("kernel", ["gaussian", "winckelmans", "singular"])
def test_p2p_ranges_small_leaves_every_kernel(vpm, orc, leaf, kernel):
    """The lane-shared list kernels (8 and 4 ways) for the kernels other than
    Gaussian-erf, F32C against the oracle: the near/far step choice, the
    padded slots and the stream reads do not depend on the kernel's form."""
    n = 1500
    P = vpm.workloads.random_cloud(n, seed=27)
    P[:, 6] *= 3.0
    starts = np.arange(0, n, leaf)
    ends = np.minimum(starts + leaf, n)
    C = np.array([P[a:b, 0:3].mean(0) for a, b in zip(starts, ends)])
    tb, te, sb, se = [], [], [], []
    for i in range(len(starts)):
        for j in range(len(starts)):
            if np.linalg.norm(C[i] - C[j]) < 0.35:
                tb.append(starts[i]); te.append(ends[i]); sb.append(starts[j]); se.append(ends[j])
    ref = P.copy()
    orc.direct_ranges(ref, kernel, tb, te, sb, se)
    A = P.copy()
    vpm.p2p_ranges(A, tb, te, sb, se, kernel=kernel, precision="f32c")
    eu, ej = _uj_err(vpm, A, ref)
    _say("p2p_ranges leaves", leaf, kernel, "U", eu, "J", ej)
    assert eu["max"] < F32_TOL and ej["max"] < F32_TOL, (leaf, kernel, eu, ej)


@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_p2p_ranges_ragged_leaves(vpm, orc, precision):
    """Leaves of uneven size (3 to 40 particles), lists with empty source
    ranges and a source range listed twice for one leaf: the wave kernel's
    source stream crosses range ends inside a 64-source slot, skips empty
    ranges, and counts a repeated range twice, as the oracle does."""
    n = 1200
    P = vpm.workloads.random_cloud(n, seed=31)
    P[:, 6] *= 3.0
    rng = np.random.default_rng(32)
    cuts = [0]
    while cuts[-1] < n:
        cuts.append(min(n, cuts[-1] + int(rng.integers(3, 41))))
    starts, ends = np.array(cuts[:-1]), np.array(cuts[1:])
    C = np.array([P[a:b, 0:3].mean(0) for a, b in zip(starts, ends)])
    tb, te, sb, se = [], [], [], []
    for i in range(len(starts)):
        for j in range(len(starts)):
            if np.linalg.norm(C[i] - C[j]) < 0.3:
                tb.append(starts[i]); te.append(ends[i]); sb.append(starts[j]); se.append(ends[j])
        tb.append(starts[i]); te.append(ends[i]); sb.append(starts[i]); se.append(starts[i])   # empty range
        if i % 7 == 0:                                                                           # listed twice
            tb.append(starts[i]); te.append(ends[i]); sb.append(starts[i]); se.append(ends[i])
    ref = P.copy()
    orc.direct_ranges(ref, "gaussianerf", tb, te, sb, se)
    A = P.copy()
    vpm.p2p_ranges(A, tb, te, sb, se, kernel="gaussianerf", precision=precision)
    eu, ej = _uj_err(vpm, A, ref)
    tol = F64_TOL if precision == "f64" else F32_TOL
    assert eu["max"] < tol and ej["max"] < tol, (eu, ej)


def test_p2p_ranges_long_list_plan_chunks(vpm, orc):
    """A leaf list of 2.6M pairs (every 4-particle leaf against every leaf, 40
    times over), long enough that the host plan splits it over threads with
    runs of one target leaf crossing the chunk ends: results equal the
    oracle's, and a bad pair is named by its lowest index whichever chunk
    finds it (target ranges are checked before source ranges)."""
    n, leaf, rep = 1024, 4, 40
    P = vpm.workloads.random_cloud(n, seed=33)
    P[:, 6] *= 3.0
    starts = np.arange(0, n, leaf, dtype=np.int64)
    L = len(starts)
    tb = np.repeat(starts, rep * L)
    te = tb + leaf
    sb = np.tile(starts, rep * L)
    se = sb + leaf
    assert len(tb) >= (2 << 20)             # eight plan chunks
    ref = P.copy()
    orc.direct_ranges(ref, "gaussianerf", tb, te, sb, se)
    A = P.copy()
    vpm.p2p_ranges(A, tb, te, sb, se, kernel="gaussianerf", precision="f32c")
    eu, ej = _uj_err(vpm, A, ref)
    assert eu["max"] < F32_TOL and ej["max"] < F32_TOL, (eu, ej)
    cases = ((((3, 2_300_000, -1), (3, 1_100_000, n + 1)), 1_100_000),
             (((1, 2_000_000, n + 3), (0, 1_500_000, -2), (2, 500_000, -1)), 1_500_000))
    for bads, first in cases:
        lists = [tb.copy(), te.copy(), sb.copy(), se.copy()]
        for col, k, v in bads:
            lists[col][k] = v
        A = P.copy()
        with pytest.raises(vpm.VPMError) as ei:
            vpm.p2p_ranges(A, *lists, kernel="gaussianerf", precision="f32c")
        assert ei.value.status == -1 and ("range pair %d " % first) in str(ei.value), str(ei.value)
        assert np.array_equal(A, P)


@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_p2p_ranges_leaves_out_of_order(vpm, orc, precision):
    """Leaf lists whose target leaves come in shuffled order, with an empty
    target range at the start of some leaves: still one item per leaf, and
    the results equal the oracle's."""
    n = 900
    P = vpm.workloads.random_cloud(n, seed=34)
    P[:, 6] *= 3.0
    rng = np.random.default_rng(35)
    cuts = [0]
    while cuts[-1] < n:
        cuts.append(min(n, cuts[-1] + int(rng.integers(4, 30))))
    starts, ends = np.array(cuts[:-1]), np.array(cuts[1:])
    C = np.array([P[a:b, 0:3].mean(0) for a, b in zip(starts, ends)])
    tb, te, sb, se = [], [], [], []
    for i in rng.permutation(len(starts)):
        for j in range(len(starts)):
            if np.linalg.norm(C[i] - C[j]) < 0.3:
                tb.append(starts[i]); te.append(ends[i]); sb.append(starts[j]); se.append(ends[j])
        if i % 3 == 0:                                  # empty target range, same begin
            tb.append(starts[i]); te.append(starts[i]); sb.append(0); se.append(n)
    ref = P.copy()
    orc.direct_ranges(ref, "gaussianerf", tb, te, sb, se)
    A = P.copy()
    vpm.p2p_ranges(A, tb, te, sb, se, kernel="gaussianerf", precision=precision)
    eu, ej = _uj_err(vpm, A, ref)
    tol = F64_TOL if precision == "f64" else F32_TOL
    assert eu["max"] < tol and ej["max"] < tol, (eu, ej)
