"""GPU tests of planned direct lists (vpm_plan_ranges + vpm_p2p_plan /
vpm_estr_plan / vpm_zeta_plan): UJ_fmm plans one direct_list and hands it
to the near field and then to Estr_fmm! (src/FLOWVPM_UJ.jl:90-125).  A plan
made once must give the per-call list entry points' results bit for bit,
for U/J, then E_str on the updated J, and zeta, and those match the
oracle's restatement (direct_ranges, estr_ranges) within the tolerances of
the per-call tests."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

F32_TOL = 1e-5
F64_TOL = 1e-12
ESTR_SCALE_TOL = 5e-5          # of the E_str summation scale (test_gpu_uj.py)
ESTR_RMS_TOL = 1e-5


def _say(*a):
    print(*a, flush=True)


def _leaf_list(P, leaf, seed, radius):
    n = P.shape[0]
    if leaf == "ragged":
        rng = np.random.default_rng(seed)
        cuts = [0]
        while cuts[-1] < n:
            cuts.append(min(n, cuts[-1] + int(rng.integers(3, 41))))
        starts, ends = np.array(cuts[:-1]), np.array(cuts[1:])
    else:
        starts = np.arange(0, n, leaf)
        ends = np.minimum(starts + leaf, n)
    Cc = np.array([P[a:b, 0:3].mean(0) for a, b in zip(starts, ends)])
    tb, te, sb, se = [], [], [], []
    for i in range(len(starts)):
        for j in range(len(starts)):
            if np.linalg.norm(Cc[i] - Cc[j]) < radius:
                tb.append(starts[i]); te.append(ends[i]); sb.append(starts[j]); se.append(ends[j])
    return [np.array(x, dtype=np.int64) for x in (tb, te, sb, se)]


def _per_call(vpm, f, which, lists, kernel, precision, transposed=True):
    api = vpm.api
    cfg = api.make_config(kernel, precision, transposed)
    ptr = [a.ctypes.data_as(vpm._lib._i64p) for a in lists]
    fn = {"p2p": api.lib().vpm_p2p_ranges, "estr": api.lib().vpm_estr_ranges,
          "zeta": api.lib().vpm_zeta_ranges}[which]
    api.check(fn(f.handle, C.byref(cfg), ptr[0], ptr[1], ptr[2], ptr[3], len(lists[0])))


@pytest.mark.parametrize("leaf", [5, 16, 64, 200, "ragged"])
@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_plan_serves_uj_then_estr_like_per_call_lists(vpm, orc, leaf, precision):
    """One plan: U/J (onto a nonzero start, as after a far field), then E_str
    with the J just computed, then zeta; each equals the per-call entry
    point bit for bit, and U/J and E_str match the oracle."""
    api = vpm.api
    n = 1500
    P = vpm.workloads.random_cloud(n, seed=60)
    P[:, 6] *= 3.0
    rng = np.random.default_rng(61)
    P[:, 9:24] = rng.normal(0, 1e-3, (n, 15))       # far-field part already there
    P[:, 39:42] = rng.normal(0, 1e-4, (n, 3))
    lists = _leaf_list(P, leaf, 62, 0.35)
    _say("plan leaf=%s precision=%s pairs=%d" % (leaf, precision, len(lists[0])))
    kern = "gaussianerf"
    fa, fb = api.Field(n), api.Field(n)
    try:
        fa.set_particles(P)
        fb.set_particles(P)
        plan = fa.plan_ranges(*lists)
        plan.p2p(kern, precision)
        _per_call(vpm, fb, "p2p", lists, kern, precision)
        A, B = np.zeros_like(P), np.zeros_like(P)
        fa.get_particles(A)
        fb.get_particles(B)
        assert np.array_equal(A, B)
        ref = P.copy()
        orc.direct_ranges(ref, kern, *lists)
        eu = vpm.metrics.err_stats(A[:, 9:12], ref[:, 9:12])
        ej = vpm.metrics.err_stats(A[:, 15:24], ref[:, 15:24])
        tol = F64_TOL if precision == "f64" else F32_TOL
        assert eu["max"] < tol and ej["max"] < tol, (eu, ej)
        # E_str of the same list on the U/J just computed (UJ_fmm -> Estr_fmm!)
        plan.estr(kern, precision, True)
        _per_call(vpm, fb, "estr", lists, kern, precision, True)
        fa.get_particles(A)
        fb.get_particles(B)
        assert np.array_equal(A, B)
        ref2 = ref.copy()
        ref2[:, 9:24] = A[:, 9:24]                  # the oracle's E_str on the device's J
        orc.estr_ranges(ref2, kern, *lists)
        st = vpm.metrics.err_stats(A[:, 39:42], ref2[:, 39:42])
        if precision == "f64":
            assert st["max"] < 1e-11, st
        else:
            scale = orc.estr_scale_ranges(ref2, kern, *lists)
            ratio = np.linalg.norm(A[:, 39:42] - ref2[:, 39:42], axis=1) / np.maximum(scale, 1e-6 * scale.max())
            assert ratio.max() < ESTR_SCALE_TOL and st["rms"] < ESTR_RMS_TOL, (ratio.max(), st)
        plan.zeta(kern, precision)
        _per_call(vpm, fb, "zeta", lists, kern, precision)
        fa.get_particles(A)
        fb.get_particles(B)
        assert np.array_equal(A, B)
        plan.close()
    finally:
        fa.close()
        fb.close()


def test_plan_rejects_stale_and_foreign_fields(vpm):
    """A plan is bound to its field and to the particle count it was made
    for; bad pairs fail at planning time with the lowest bad index; the
    field may be destroyed before the plan."""
    api = vpm.api
    n = 300
    P = vpm.workloads.random_cloud(n, seed=63)
    fa, fb = api.Field(n + 10), api.Field(n)
    try:
        fa.set_particles(P)
        fb.set_particles(P)
        lists = [np.array([0, 100]), np.array([100, 300]), np.array([0, 0]), np.array([300, 300])]
        plan = fa.plan_ranges(*lists)
        cfg = api.make_config("gaussianerf", "f32c", True)
        rc = api.lib().vpm_p2p_plan(fb.handle, C.byref(cfg), plan.handle)
        assert rc == vpm._lib.VPM_EINVAL and b"another field" in api.lib().vpm_last_error()
        fa.add_particles(P[:3])
        with pytest.raises(vpm.VPMError) as ei:
            plan.estr()
        assert "np = 300" in str(ei.value)
        fa.remove_particle(0)
        fa.remove_particle(0)
        fa.remove_particle(0)
        plan.p2p()                                 # np is 300 again
        with pytest.raises(vpm.VPMError) as ei:
            fa.plan_ranges([0, 5, 0], [10, 400, 10], [0, 0, -1], [5, 5, 5])
        assert "range pair 1 " in str(ei.value)
        empty = fa.plan_ranges([], [], [], [])
        empty.p2p()
        empty.zeta()
        empty.close()
        fa.close()
        rc = api.lib().vpm_p2p_plan(fb.handle, C.byref(cfg), plan.handle)
        assert rc == vpm._lib.VPM_EINVAL
        plan.close()
    finally:
        fa.close()
        fb.close()


@pytest.mark.parametrize("tleaf", [5, 8, 13])
def test_plan_small_targets_long_and_empty_sources(vpm, orc, tleaf):
    """Target leaves of 5, 8 and 13 particles (lanes shared 8 and 4 ways)
    whose source streams mix ranges longer than a 64-source chunk, ranges of
    one particle and empty ranges, so chunks end inside a range, after an
    empty one and exactly at a range's end.  U/J, E_str on the J just
    computed and zeta each match the oracle; the planned call equals the
    per-call list bit for bit."""
    api = vpm.api
    n = 1200
    P = vpm.workloads.random_cloud(n, seed=64)
    P[:, 6] *= 3.0
    rng = np.random.default_rng(65)
    P[:, 39:42] = rng.normal(0, 1e-4, (n, 3))
    srcs = [(0, 150), (300, 301), (400, 400), (500, 564), (564, 700), (700, 701), (701, 701), (900, 1000)]
    tb, te, sb, se = [], [], [], []
    for a in range(0, n, tleaf):
        b = min(a + tleaf, n)
        k = (a // tleaf) % len(srcs)
        for (c, d) in srcs[k:] + srcs[:k]:          # a different start per leaf
            tb.append(a); te.append(b); sb.append(c); se.append(d)
        tb.append(a); te.append(b); sb.append(a); se.append(b)
    lists = [np.array(x, dtype=np.int64) for x in (tb, te, sb, se)]
    kern = "gaussianerf"
    _say("small targets leaf=%d pairs=%d" % (tleaf, len(tb)))
    fa, fb = api.Field(n), api.Field(n)
    try:
        fa.set_particles(P)
        fb.set_particles(P)
        plan = fa.plan_ranges(*lists)
        plan.p2p(kern, "f32c")
        _per_call(vpm, fb, "p2p", lists, kern, "f32c")
        A, B = np.zeros_like(P), np.zeros_like(P)
        fa.get_particles(A)
        fb.get_particles(B)
        assert np.array_equal(A, B)
        ref = P.copy()
        orc.direct_ranges(ref, kern, *lists)
        eu = vpm.metrics.err_stats(A[:, 9:12], ref[:, 9:12])
        ej = vpm.metrics.err_stats(A[:, 15:24], ref[:, 15:24])
        assert eu["max"] < F32_TOL and ej["max"] < F32_TOL, (eu, ej)
        plan.estr(kern, "f32c", True)
        fa.get_particles(A)
        ref2 = ref.copy()
        ref2[:, 9:24] = A[:, 9:24]
        orc.estr_ranges(ref2, kern, *lists)
        scale = orc.estr_scale_ranges(ref2, kern, *lists)
        ratio = np.linalg.norm(A[:, 39:42] - ref2[:, 39:42], axis=1) / np.maximum(scale, 1e-6 * scale.max())
        st = vpm.metrics.err_stats(A[:, 39:42], ref2[:, 39:42])
        assert ratio.max() < ESTR_SCALE_TOL and st["rms"] < ESTR_RMS_TOL, (ratio.max(), st)
        R = A.copy()
        orc.zeta_ranges(R, kern, *lists)
        S = A.copy()
        S[:, 3:6] = np.abs(S[:, 3:6])
        S[:, 15:18] = np.abs(S[:, 15:18])
        orc.zeta_ranges(S, kern, *lists)
        plan.zeta(kern, "f32c")
        fa.get_particles(A)
        scaled = (np.abs(A[:, 15:18] - R[:, 15:18]) / np.maximum(S[:, 15:18], 1e-300)).max()
        assert scaled < 1e-5, scaled
        plan.close()
    finally:
        fa.close()
        fb.close()
