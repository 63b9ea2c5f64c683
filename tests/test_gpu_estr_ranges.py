"""GPU parity of E_str over a direct list (vpm_estr_ranges, Estr_fmm! of
src/FLOWVPM_subfilterscale_models.jl:157-188) against the oracle's
restatement: leaf pairs with overlapping and repeated target ranges, static
particles (not skipped by Estr_fmm!) and SFS accumulated onto a nonzero
start.  Tolerances are those of the all-pairs E_str (test_gpu_uj.py): F64
to FP64 rounding; F32C within ESTR_SCALE_TOL of the summation scale
s_i = sum zeta |S_ij| over the list per particle (the budget is derived in
test_gpu_uj.py), with the RMS relative error within ESTR_RMS_TOL (E_str sums
cancel by factors of 10-100 on random J)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ESTR_SCALE_TOL = 5e-5
ESTR_RMS_TOL = 1e-5


def _estr_check(vpm, orc, A, ref, P, kernel, lists, what, transposed=True):
    # the scale takes the floor of the relative error, 1e-6 of its maximum
    # (particles below it have only pairs near or past the kernel's cut)
    scale = orc.estr_scale_ranges(P, kernel, *lists, transposed=transposed)
    e = np.linalg.norm(A - ref, axis=1)
    den = np.maximum(scale, 1e-6 * scale.max())
    ratio = np.where(den > 0, e / np.where(den > 0, den, 1.0), np.where(e > 0, np.inf, 0.0))
    st = vpm.metrics.err_stats(A, ref)
    print(what, "err/scale max %.2e" % ratio.max(), "relative", st, flush=True)
    assert ratio.max() < ESTR_SCALE_TOL and st["rms"] < ESTR_RMS_TOL, (ratio.max(), st)


def _leaf_lists(n, nleaves, seed, density=0.4):
    rng = np.random.default_rng(seed)
    cuts = np.sort(np.concatenate([[0, n], rng.choice(np.arange(1, n), nleaves - 1, replace=False)]))
    leaves = list(zip(cuts[:-1], cuts[1:]))
    tb, te, sb, se = [], [], [], []
    for (a, b) in leaves:
        for (c, d) in leaves:
            if rng.random() < density:
                tb.append(a); te.append(b); sb.append(c); se.append(d)
    return tb, te, sb, se


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "gaussian", "singular"])
@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_estr_ranges_vs_oracle(vpm, orc, kernel, precision):
    n = 900
    P = vpm.workloads.random_cloud(n, seed=16)
    P[:, 6] *= 3.0
    rng = np.random.default_rng(17)
    P[:, 15:24] = rng.normal(0, 1, (n, 9))
    P[:, 39:42] = rng.normal(0, 1e-3, (n, 3))
    P[[4, 400], 42] = 1.0
    tb, te, sb, se = _leaf_lists(n, 21, seed=18)
    tb += [5, 100]; te += [300, 120]; sb += [0, 40]; se += [50, 41]
    ref = P.copy()
    orc.estr_ranges(ref, kernel, tb, te, sb, se)
    A = P.copy()
    vpm.api.estr_ranges(A, tb, te, sb, se, kernel=kernel, precision=precision)
    assert np.array_equal(A[:, 0:39], P[:, 0:39])
    if precision == "f64" or kernel == "singular":
        st = vpm.metrics.err_stats(A[:, 39:42], ref[:, 39:42])
        print("estr_ranges", kernel, precision, st, flush=True)
        assert st["max"] < 1e-11, st
    else:
        _estr_check(vpm, orc, A[:, 39:42], ref[:, 39:42], P, kernel, (tb, te, sb, se),
                    "estr_ranges %s %s" % (kernel, precision))


def test_estr_ranges_whole_field_equals_estr_direct(vpm, orc):
    """With no static particle, the single pair (all, all) is Estr_direct!."""
    n = 700
    P = vpm.workloads.random_cloud(n, seed=19)
    P[:, 6] *= 3.0
    P[:, 15:24] = np.random.default_rng(20).normal(0, 1, (n, 9))
    ref = P.copy()
    orc.estr_direct(ref, "gaussianerf")
    A = P.copy()
    vpm.api.estr_ranges(A, [0], [n], [0], [n], kernel="gaussianerf", precision="f64")
    np.testing.assert_allclose(A[:, 39:42], ref[:, 39:42], rtol=1e-11, atol=1e-12 * np.abs(ref[:, 39:42]).max())


@pytest.mark.parametrize("leaf", [5, 12, 24, 48, "ragged"])
@pytest.mark.parametrize("kernel,transposed", [("gaussianerf", True), ("winckelmans", True), ("gaussian", True),
                                               ("gaussianerf", False)])
def test_estr_ranges_small_leaves(vpm, orc, leaf, kernel, transposed):
    """FMM-style lists over leaves of 5, 12, 24 and 48 particles and of uneven
    size (3 to 40), with an empty source range per leaf: the wave kernel, its
    lanes shared 8, 4, 2 or 1 ways, against the oracle; every kernel with a
    cut, and the classic (not transposed) form, whose packed source records
    hold J Gamma instead of J^T Gamma."""
    n = 1000
    P = vpm.workloads.random_cloud(n, seed=41)
    P[:, 6] *= 3.0
    rng = np.random.default_rng(42)
    P[:, 15:24] = rng.normal(0, 1, (n, 9))
    if leaf == "ragged":
        cuts = [0]
        while cuts[-1] < n:
            cuts.append(min(n, cuts[-1] + int(rng.integers(3, 41))))
        starts, ends = np.array(cuts[:-1]), np.array(cuts[1:])
    else:
        starts = np.arange(0, n, leaf)
        ends = np.minimum(starts + leaf, n)
    C = np.array([P[a:b, 0:3].mean(0) for a, b in zip(starts, ends)])
    tb, te, sb, se = [], [], [], []
    for i in range(len(starts)):
        for j in range(len(starts)):
            if np.linalg.norm(C[i] - C[j]) < 0.3:
                tb.append(starts[i]); te.append(ends[i]); sb.append(starts[j]); se.append(ends[j])
        tb.append(starts[i]); te.append(ends[i]); sb.append(starts[i]); se.append(starts[i])
    ref = P.copy()
    orc.estr_ranges(ref, kernel, tb, te, sb, se, transposed=transposed)
    A = P.copy()
    vpm.api.estr_ranges(A, tb, te, sb, se, kernel=kernel, precision="f32c", transposed=transposed)
    assert np.array_equal(A[:, 0:39], P[:, 0:39])
    _estr_check(vpm, orc, A[:, 39:42], ref[:, 39:42], P, kernel, (tb, te, sb, se),
                "estr_ranges leaves %s %s transposed=%s" % (leaf, kernel, transposed), transposed=transposed)
