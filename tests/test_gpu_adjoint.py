"""GPU parity of the fused FP32 adjoint (VPM_PREC_F32C of vpm_uj_adjoint).

Reference points: the oracle's FP64 adjoint (CPU restatement of the pullback
of src/FLOWVPM_rrules.jl:179-330, itself pinned to mpmath and to central
differences), the library's FP64 adjoint, and central differences of the
oracle's exact loss.  Errors are per-particle relative norms (metrics.rel_err,
floor 1e-6 of the largest norm).

Tolerances of the FP32 path: every pair is evaluated in FP32 (hi/lo
positions and cancellation-free reduced forms near the core), partial sums
over 64 or 256 pairs in FP32, everything above in FP64.  Each pair term is
then exact to a few FP32 roundings, so the error of an output is bounded by
a multiple of its summation scale sum_i |term_i| (oracle.uj_adjoint_scale),
however much the terms cancel.  Asserted per particle:
    |got - ref| <= ADJ_SCALED_TOL * |scale|     (vector norms),
(with the scale floored at SCALE_FLOOR of the field's largest), and over all
particles ||got - ref|| / ||ref|| <= ADJ_NORM_TOL.  The plain
per-particle relative error is printed too: for dL/dsigma, whose near-pair
terms can cancel almost completely (sum |term| / |sum| reaches 1e4 on the
random cloud of 1,500), it reaches ~1e-3 at such particles while the
scaled error stays at FP32 rounding.
"""
import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu

KERNELS = ["gaussianerf", "winckelmans", "gaussian", "singular"]
ADJ_SCALED_TOL = 1e-5
ADJ_NORM_TOL = 1e-5
# dL/dsigma terms are proportional to E = e^{-q^2/2} (Gaussian-erf) or
# e^{-q^3} (Gaussian), whose relative condition in q is q^2 <= 49 and
# 3 q^3 <= 81 inside the far radius: FP32-rounded q (and 1/sigma) give each
# term a relative error of up to ~1e2 roundings, hence the looser bound.
ADJ_SCALED_TOL_SIGMA = 1e-4
# A particle with no neighbour inside the far-field radius has dL/dsigma made
# only of e^{-q^2/2}-small terms (q >= 7, Gaussian: e^{-q^3}, q >= 3), which
# the FP32 path drops; its scale is floored at this fraction of the field's
# largest.
SCALE_FLOOR = 1e-6


def _say(*a):
    print(*a, flush=True)


def _check(vpm, tag, got, ref, scale):
    for name, a, r, m in zip(("X", "Gamma", "sigma"), got, ref, scale):
        a = np.asarray(a).reshape(len(a), -1)
        r = np.asarray(r).reshape(len(r), -1)
        m = np.asarray(m).reshape(len(m), -1)
        assert np.all(np.isfinite(a)), (tag, name)
        num = np.linalg.norm(a - r, axis=1)
        den = np.linalg.norm(m, axis=1)
        scaled = num / np.maximum(den, max(SCALE_FLOOR * den.max(), 1e-300))
        normwise = float(np.linalg.norm(a - r) / max(np.linalg.norm(r), 1e-300))
        st = vpm.metrics.err_stats(a, r)
        _say("adjoint f32c", tag, name, "rel", st, "scaled max %.2e" % scaled.max(), "normwise %.2e" % normwise)
        tol = ADJ_SCALED_TOL_SIGMA if name == "sigma" else ADJ_SCALED_TOL
        assert scaled.max() < tol and normwise < ADJ_NORM_TOL, (tag, name, scaled.max(), normwise)


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_adjoint_f32c_vs_reference_pullback(vpm, orc, kernel):
    """Against the reference's own CPU gradient: UJS_pullback
    (rrules.jl:183-326) restated literally in the oracle, on the cloud of
    test_adjoint_f32c_vs_oracle, to the same summation-scale bounds."""
    P = vpm.workloads.random_cloud(1500, seed=21)
    w, W = vpm.workloads.adjoint_weights(1500)
    ref = orc.uj_adjoint_rrules(P, kernel, w, W)
    got = vpm.uj_adjoint(P, w, W, kernel=kernel, precision="f32c")
    _check(vpm, "rrules " + kernel, got, ref, orc.uj_adjoint_scale(P, kernel, w, W))


@pytest.mark.parametrize("kernel", KERNELS)
def test_adjoint_f32c_vs_oracle(vpm, orc, kernel):
    P = vpm.workloads.random_cloud(1500, seed=21)
    w, W = vpm.workloads.adjoint_weights(1500)
    ref = orc.uj_adjoint(P, kernel, w, W)
    got = vpm.uj_adjoint(P, w, W, kernel=kernel, precision="f32c")
    _check(vpm, kernel, got, ref, orc.uj_adjoint_scale(P, kernel, w, W))


def test_adjoint_f32c_golden(vpm, orc):
    g = golden("adjoint.json")
    P = vpm.fields.make_particles(g["X"], g["Gamma"], g["sigma"])
    for kernel in KERNELS:
        got = vpm.uj_adjoint(P, np.array(g["Ubar"]), np.array(g["Jbar"]), kernel=kernel, precision="f32c")
        ref = g[kernel]
        scale = orc.uj_adjoint_scale(P, kernel, np.array(g["Ubar"]), np.array(g["Jbar"]))
        _check(vpm, "golden " + kernel, got, (ref["Xbar"], ref["Gbar"], ref["Sbar"]), scale)


@pytest.mark.parametrize("n", [1, 2, 63, 64, 65, 255, 256, 257, 700, 4099])
def test_adjoint_f32c_ragged_sizes(vpm, orc, n):
    """Partial waves, partial source groups and partial chunks (sigma in
    [0.1, 0.3] so that particles have neighbours inside their cores)."""
    P = vpm.workloads.random_cloud(n, seed=100 + n)
    P[:, 6] *= 7.5
    w, W = vpm.workloads.adjoint_weights(n, seed=n)
    for kernel in ("gaussianerf", "singular"):
        ref = orc.uj_adjoint(P, kernel, w, W)
        got = vpm.uj_adjoint(P, w, W, kernel=kernel, precision="f32c")
        if n == 1:
            for a in got:
                assert np.all(np.asarray(a) == 0.0)
            continue
        _check(vpm, "n=%d %s" % (n, kernel), got, ref, orc.uj_adjoint_scale(P, kernel, w, W))


def test_adjoint_f32c_coincident_and_deterministic(vpm, orc):
    """Coincident particles contribute nothing to each other; repeated calls
    give bit-identical results (fixed summation order, no atomics)."""
    P = vpm.workloads.random_cloud(600, seed=5)
    P[300:310, 0:3] = P[0:10, 0:3]        # ten exact duplicates
    w, W = vpm.workloads.adjoint_weights(600, seed=6)
    ref = orc.uj_adjoint(P, "gaussianerf", w, W)
    a = vpm.uj_adjoint(P, w, W, kernel="gaussianerf", precision="f32c")
    b = vpm.uj_adjoint(P, w, W, kernel="gaussianerf", precision="f32c")
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    _check(vpm, "coincident", a, ref, orc.uj_adjoint_scale(P, "gaussianerf", w, W))


@pytest.mark.slow
@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_adjoint_cfg3(vpm, orc, kernel):
    """Random cloud N=32768 (positions U[0,1]^3, Gamma ~ N(0,1e-3^2), sigma ~
    U[0.02,0.04]) with weights w ~ N(0,1)^3, W ~ N(0,1)^(3x3): the FP32
    adjoint against the FP64 adjoint on every particle, and against central
    differences (h = 1e-6) of the oracle's exact loss on 256 random
    coordinates of X, Gamma and sigma."""
    n = 32768
    P = vpm.workloads.random_cloud(n, seed=33)
    w, W = vpm.workloads.adjoint_weights(n, seed=34)
    got = vpm.uj_adjoint(P, w, W, kernel=kernel, precision="f32c")
    ref = vpm.uj_adjoint(P, w, W, kernel=kernel, precision="f64")
    scale = orc.uj_adjoint_scale(P, kernel, w, W)
    _check(vpm, "cfg3 " + kernel, got, ref, scale)
    rng = np.random.default_rng(35)
    h = 1e-6
    flat = [np.asarray(x).reshape(n, -1) for x in got]
    mags = [np.asarray(x).reshape(n, -1) for x in scale]
    worst = 0.0
    for k, col in zip(rng.choice(n, 256, replace=False), rng.integers(0, 7, 256)):
        Pp, Pm = P.copy(), P.copy()
        Pp[k, col] += h
        Pm[k, col] -= h
        fd = (orc.loss_local(Pp, kernel, w, W, k) - orc.loss_local(Pm, kernel, w, W, k)) / (2 * h)
        grp, c = (0, col) if col < 3 else ((1, col - 3) if col < 6 else (2, 0))
        an, mag = flat[grp][k, c], mags[grp][k, c]
        e = abs(fd - an) / mag
        worst = max(worst, e)
        assert e < (ADJ_SCALED_TOL_SIGMA if col == 6 else ADJ_SCALED_TOL), (k, col, fd, an, mag)
    _say("cfg3 finite differences, worst error / summation scale", kernel, worst)


@pytest.mark.parametrize("kernel", KERNELS)
def test_adjoint_f32c_loss_on_u_only(vpm, orc, kernel):
    """A loss on U alone (every Jbar zero) takes the kernel without the Jbar
    terms.  Against the oracle with the same cotangents, to the bounds of the
    general kernel; and against the general kernel run with Jbar = 1e-30 R,
    whose Jbar terms fall below FP32 rounding of every kept term: the same
    gradients bit for bit, since the terms dropped are exact zeros and the
    kept ones are rounded in the same order."""
    P = vpm.workloads.random_cloud(1500, seed=23)
    w, W = vpm.workloads.adjoint_weights(1500)
    Z = np.zeros_like(W)
    ref = orc.uj_adjoint(P, kernel, w, Z)
    got = vpm.uj_adjoint(P, w, Z, kernel=kernel, precision="f32c")
    _check(vpm, "u-only " + kernel, got, ref, orc.uj_adjoint_scale(P, kernel, w, Z))
    gen = vpm.uj_adjoint(P, w, 1e-30 * W, kernel=kernel, precision="f32c")
    for name, a, b in zip(("X", "Gamma", "sigma"), got, gen):
        d = np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300)
        _say("u-only vs general", kernel, name, "%.1e" % d, "identical entries %.4f"
             % np.mean(np.asarray(a) == np.asarray(b)))
        assert np.array_equal(np.asarray(a), np.asarray(b)), (kernel, name, d)
