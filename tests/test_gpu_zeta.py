"""GPU parity of zeta_direct (vpm_zeta_direct; src/FLOWVPM_viscous.jl:486-515).

F64 must match the oracle to FP64 rounding (same expression, different
summation order).  F32C: zeta >= 0, so the summation scale of each output is
the same sum with |Gamma| (the oracle run on |Gamma|); every component is
within ZETA_SCALED_TOL of that scale.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

KERNELS = ["gaussianerf", "winckelmans", "gaussian", "singular"]
ZETA_SCALED_TOL = 1e-5


def _say(*a):
    print(*a, flush=True)


def _cloud(vpm, n, seed):
    P = vpm.workloads.random_cloud(n, seed=seed)
    P[:, 6] *= 3.0                  # sigma in [0.06, 0.12]: each particle sees neighbours
    P[n // 2: n // 2 + 3, 0:3] = P[0:3, 0:3]    # coincident particles
    P[[1, 9], 42] = 1.0             # static particles are targets and sources too
    P[:, 18:24] = 5.0               # J rows beyond 1:3 must stay
    return P


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("n", [1, 255, 2000])
def test_zeta_direct_vs_oracle(vpm, orc, kernel, n):
    api = vpm.api
    P = _cloud(vpm, n, seed=60 + n) if n > 3 else vpm.workloads.random_cloud(n, seed=2)
    R = P.copy()
    orc.zeta_direct(R, kernel)
    S = P.copy()
    S[:, 3:6] = np.abs(S[:, 3:6])
    orc.zeta_direct(S, kernel)
    f = api.Field(n, 0)
    for prec in ("f64", "f32c"):
        f.set_particles(P)
        f.zeta_direct(kernel, prec)
        G = np.zeros_like(P)
        f.get_particles(G)
        assert np.array_equal(G[:, 18:24], P[:, 18:24])
        assert np.array_equal(G[:, 0:15], P[:, 0:15])
        err = np.abs(G[:, 15:18] - R[:, 15:18])
        if prec == "f64" or kernel == "singular":
            tol = 1e-12 * np.abs(S[:, 15:18]).max() + 1e-300
            assert err.max() <= tol, (prec, err.max())
        else:
            scaled = (err / np.maximum(S[:, 15:18], 1e-300)).max()
            _say("zeta", kernel, n, prec, "scaled max %.2e" % scaled)
            assert scaled < ZETA_SCALED_TOL, (kernel, n, scaled)
    f.close()


def test_zeta_direct_pfield_mirror(vpm, orc):
    api = vpm.api
    P = _cloud(vpm, 300, seed=8)
    pf = api.ParticleField(300, kernel="gaussianerf", precision="f64")
    api.add_particles(pf, P)
    api.zeta_direct(pf)
    R = P.copy()
    orc.zeta_direct(R, "gaussianerf")
    np.testing.assert_allclose(pf.particles[:300, 15:18], R[:, 15:18], rtol=1e-12, atol=1e-12)


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


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("nleaves,extra", [(21, False), (21, True), (4, False), (180, False), (75, False),
                                           (37, False), (75, True)])
def test_zeta_ranges_vs_oracle(vpm, orc, kernel, nleaves, extra):
    """zeta_fmm's pair loop over a direct list (vpm_zeta_ranges): leaf lists
    of ~5, ~12, ~24 (a wave shared by 8, 4, 2 leaves) or ~43 targets (one
    wave per leaf), or ~225 (256-lane blocks), with
    overlapping and repeated target ranges when `extra`; J[1:3] accumulates
    onto a nonzero start and static particles are not skipped.  F64 within
    FP64 rounding of the oracle; F32C within ZETA_SCALED_TOL of the
    summation scale (the same sums with |Gamma| and |J start|)."""
    n = 900
    P = _cloud(vpm, n, seed=70 + nleaves)
    P[:, 15:18] = np.random.default_rng(71).normal(0, 1e-2, (n, 3))
    tb, te, sb, se = _leaf_lists(n, nleaves, seed=72 + nleaves)
    if extra:
        tb += [5, 100, 100]; te += [300, 120, 120]; sb += [0, 40, 40]; se += [50, 41, 41]
    R = P.copy()
    orc.zeta_ranges(R, kernel, tb, te, sb, se)
    S = P.copy()
    S[:, 3:6] = np.abs(S[:, 3:6])
    S[:, 15:18] = np.abs(S[:, 15:18])
    orc.zeta_ranges(S, kernel, tb, te, sb, se)
    for prec in ("f64", "f32c"):
        A = P.copy()
        vpm.api.zeta_ranges(A, tb, te, sb, se, kernel=kernel, precision=prec)
        assert np.array_equal(A[:, 0:15], P[:, 0:15]) and np.array_equal(A[:, 18:], P[:, 18:])
        scaled = (np.abs(A[:, 15:18] - R[:, 15:18]) / np.maximum(S[:, 15:18], 1e-300)).max()
        _say("zeta_ranges", kernel, nleaves, extra, prec, "scaled max %.2e" % scaled)
        if prec == "f64" or kernel == "singular":
            assert scaled < 1e-12, scaled
        else:
            assert scaled < ZETA_SCALED_TOL, scaled


def test_zeta_ranges_whole_field_equals_zeta_direct(vpm, orc):
    """The single pair (all, all) onto zeroed J rows is zeta_direct."""
    n = 1000
    P = _cloud(vpm, n, seed=74)
    P[:, 15:18] = 0.0
    R = P.copy()
    orc.zeta_direct(R, "gaussianerf")
    A = P.copy()
    vpm.api.zeta_ranges(A, [0], [n], [0], [n], kernel="gaussianerf", precision="f64")
    np.testing.assert_allclose(A[:, 15:18], R[:, 15:18], rtol=1e-11, atol=1e-12 * np.abs(R[:, 15:18]).max())
