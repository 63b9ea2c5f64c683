"""Pins the FP64 oracle against golden vectors, closed forms and invariants.

Each check is chosen so that a plausible restatement mistake (a dropped term,
a wrong sign or index, a transposed J) fails at least one of them.
"""
import math

import numpy as np
import pytest

from conftest import golden

KERNELS = ["singular", "gaussian", "gaussianerf", "winckelmans"]


def test_erf_matches_mpmath(orc):
    g = golden("erf.json")
    for x, ref in zip(g["x"], g["erf"]):
        got = orc.erf64(x)
        # fdlibm's s_erf is accurate to < 1 ulp; allow 4 ulp
        assert abs(got - ref) <= 4 * np.spacing(abs(ref)) + 1e-300, (x, got, ref)


@pytest.mark.parametrize("kernel", KERNELS)
def test_kernel_functions(orc, kernel):
    g = golden("kernels.json")
    for q, gr, dgr, d2gr, zr in zip(g["q"], g[kernel]["g"], g[kernel]["dg"],
                                    g[kernel]["d2g"], g[kernel]["zeta"]):
        gg, dg = orc.g_dgdr(kernel, q)
        _, _, d2g = orc.g_derivs(kernel, q)
        z = orc.zeta(kernel, q)
        # Gaussian-type g cancels catastrophically for small q in FP64 (the
        # reference evaluates it the same way), so g is held to an absolute
        # bound near q=0 and to a relative bound elsewhere.
        assert abs(gg - gr) <= 1e-14 * max(abs(gr), 1.0) if q < 1 else abs(gg - gr) <= 4e-15 * abs(gr)
        assert abs(dg - dgr) <= 1e-14 * abs(dgr) + 1e-300
        assert abs(d2g - d2gr) <= 1e-13 * abs(d2gr) + 1e-300
        assert abs(z - zr) <= 1e-14 * abs(zr) + 1e-300
        # g' = 4 pi q^2 zeta for the three smooth kernels (kernel.jl:51-76)
        if kernel != "singular":
            assert abs(dg - 4 * math.pi * q * q * z) <= 1e-13 * abs(dg) + 1e-300


def _field_from(g, vpm):
    return vpm.fields.make_particles(g["X"], g["Gamma"], g["sigma"])


@pytest.mark.parametrize("kernel", KERNELS)
def test_uj_matches_mpmath(orc, vpm, kernel):
    g = golden("uj_%s.json" % kernel)
    P = _field_from(g, vpm)
    orc.uj_direct(P, kernel)
    U, J = np.array(g["U"]), np.array(g["J"])
    # Particles 0..15 form pairs at q = 1e-4 .. 20 (make_golden.particle_set).
    # Away from them the restatement is at FP64 rounding level.
    eu = vpm.metrics.err_stats(P[16:, 9:12], U[16:])
    ej = vpm.metrics.err_stats(P[16:, 15:24], J[16:])
    assert eu["max"] < 1e-12 and ej["max"] < 1e-12, (eu, ej)
    # At the near pairs the reference's own FP64 formula cancels: the
    # Gaussian kernel's 1 - exp(-q^3) (kernel.jl:64-65) keeps ~eps/q^3
    # relative accuracy, i.e. ~1e-4 at q = 1e-4, and the oracle reproduces
    # that on purpose; Gaussian-erf's erf(q/sqrt2) - sqrt(2/pi) q exp(-q^2/2)
    # (kernel.jl:55-56) keeps ~eps/q^2, i.e. ~1e-8 relative at q = 1e-4.
    loose = {"gaussian": 3e-5, "gaussianerf": 3e-7}.get(kernel, 1e-11)
    eu = vpm.metrics.err_stats(P[:, 9:12], U)
    ej = vpm.metrics.err_stats(P[:, 15:24], J)
    assert eu["max"] < loose and ej["max"] < loose, (eu, ej)


def test_two_body_check_fmm(orc, vpm):
    """scripts/check_fmm.jl:90-98 against its analytic singular formulas."""
    g = golden("check_fmm.json")
    P = _field_from(g, vpm)
    orc.uj_direct(P, "singular")
    np.testing.assert_allclose(P[:, 9:12], g["U"], rtol=1e-13, atol=0)
    np.testing.assert_allclose(P[:, 15:24], g["J"], rtol=1e-12, atol=1e-14)
    # Gaussian-erf at q = 6.69 matches the singular result to ~1e-9
    P2 = _field_from(g, vpm)
    orc.uj_direct(P2, "gaussianerf")
    np.testing.assert_allclose(P2[:, 9:12], g["U"], rtol=1e-8)


def _cloud(vpm, n=40, seed=0, sig=(0.05, 0.3)):
    rng = np.random.default_rng(seed)
    return vpm.fields.make_particles(rng.uniform(0, 1, (n, 3)), rng.normal(0, 1, (n, 3)),
                                     rng.uniform(*sig, n))


@pytest.mark.parametrize("kernel", KERNELS)
def test_uj_invariants(orc, vpm, kernel):
    P = _cloud(vpm)
    A = P.copy()
    orc.uj_direct(A, kernel)
    U, J = A[:, 9:12], vpm.fields.J_matrix(A)
    # tr J = 0 pairwise (d.(d x G) = 0, epsilon term traceless)
    assert np.abs(np.trace(J, axis1=1, axis2=2)).max() < 1e-12 * np.abs(J).max()
    # rotation equivariance: U -> R U, J -> R J R^T
    th = 0.7
    Rz = np.array([[math.cos(th), -math.sin(th), 0], [math.sin(th), math.cos(th), 0], [0, 0, 1]])
    Rx = np.array([[1, 0, 0], [0, math.cos(0.3), -math.sin(0.3)], [0, math.sin(0.3), math.cos(0.3)]])
    R = Rz @ Rx
    B = P.copy()
    B[:, 0:3] = P[:, 0:3] @ R.T
    B[:, 3:6] = P[:, 3:6] @ R.T
    orc.uj_direct(B, kernel)
    np.testing.assert_allclose(B[:, 9:12], U @ R.T, rtol=1e-9, atol=1e-11 * np.abs(U).max())
    JR = np.einsum("ab,nbc,dc->nad", R, J, R)
    np.testing.assert_allclose(vpm.fields.J_matrix(B), JR, rtol=1e-9, atol=1e-11 * np.abs(J).max())
    # scaling: X, sigma -> lam X, lam sigma  =>  U -> U/lam^2, J -> J/lam^3
    lam = 2.5
    C = P.copy()
    C[:, 0:3] *= lam
    C[:, 6] *= lam
    orc.uj_direct(C, kernel)
    np.testing.assert_allclose(C[:, 9:12], U / lam ** 2, rtol=1e-10, atol=1e-13 * np.abs(U).max())
    np.testing.assert_allclose(C[:, 15:24], A[:, 15:24] / lam ** 3, rtol=1e-10,
                               atol=1e-13 * np.abs(A[:, 15:24]).max())
    # translation and permutation invariance
    D = P.copy()
    D[:, 0:3] += np.array([0.25, -0.5, 0.125])
    perm = np.random.default_rng(1).permutation(len(P))
    D = D[perm].copy()
    orc.uj_direct(D, kernel)
    np.testing.assert_allclose(D[:, 9:24], A[perm, 9:24], rtol=1e-9, atol=1e-12 * np.abs(A[:, 9:24]).max())


def test_far_field_is_singular(orc, vpm):
    """For q >> 1 every smooth kernel reduces to singular Biot-Savart."""
    P = _cloud(vpm, n=30, sig=(1e-4, 2e-4))
    ref = P.copy()
    orc.uj_direct(ref, "singular")
    for k in ("gaussian", "gaussianerf", "winckelmans"):
        A = P.copy()
        orc.uj_direct(A, k)
        tol = 1e-6 if k == "winckelmans" else 1e-12   # Winckelmans: 1 - g ~ 1.9/q^4
        np.testing.assert_allclose(A[:, 9:24], ref[:, 9:24], rtol=tol, atol=tol * np.abs(ref[:, 9:24]).max())


def test_self_pair_and_accumulate(orc, vpm):
    """Skip only r2 == 0 (fmm.jl:118); UJ_direct(reset=false) accumulates and
    static particles are not reset (particlefield.jl:468)."""
    P = _cloud(vpm, n=10)
    P[3, 0:3] = P[2, 0:3]                # coincident pair: contributes nothing
    A = P.copy()
    orc.uj_direct(A, "gaussianerf")
    B = np.delete(P, 3, axis=0)
    orc.uj_direct(B, "gaussianerf")
    # particle 2 sees exactly the field it sees without particle 3: the r=0
    # pair is skipped and the additions happen in the same order
    assert np.array_equal(A[2, 9:24], B[2, 9:24])
    # accumulate semantics
    C = A.copy()
    orc.uj_direct(C, "gaussianerf", reset=False)
    np.testing.assert_allclose(C[:, 9:24], 2 * A[:, 9:24], rtol=1e-14)
    # static targets keep accumulating across UJ calls even with reset=true
    S = P.copy()
    S[5, 42] = 1.0
    orc.uj_direct(S, "gaussianerf")
    orc.uj_direct(S, "gaussianerf")
    np.testing.assert_allclose(S[5, 9:24], 2 * A[5, 9:24], rtol=1e-13)
    np.testing.assert_allclose(S[4, 9:24], A[4, 9:24], rtol=1e-13)


def test_coincident_pair_skipped(orc, vpm):
    P = _cloud(vpm, n=6)
    P[1, 0:3] = P[0, 0:3]
    A = P.copy()
    orc.uj_direct(A, "winckelmans")
    # removing particle 1's source influence on 0 changes nothing at 0 except
    # through r=0, which is skipped; compare with a field where 1 is moved far
    Bf = P.copy()
    Bf[1, 3:6] = 0.0                     # zero strength: no source influence
    orc.uj_direct(Bf, "winckelmans")
    np.testing.assert_allclose(A[0, 9:24], Bf[0, 9:24], rtol=1e-13, atol=1e-15)


def test_direct_ranges_equals_union(orc, vpm):
    P = _cloud(vpm, n=50)
    A = P.copy()
    orc.direct(A, "gaussianerf")
    B = P.copy()
    cuts = [0, 7, 19, 33, 50]
    tb, te, sb, se = [], [], [], []
    for i in range(4):
        for j in range(4):
            tb.append(cuts[i]); te.append(cuts[i + 1]); sb.append(cuts[j]); se.append(cuts[j + 1])
    orc.direct_ranges(B, "gaussianerf", tb, te, sb, se)
    np.testing.assert_allclose(B[:, 9:24], A[:, 9:24], rtol=1e-12, atol=1e-14 * np.abs(A[:, 9:24]).max())


@pytest.mark.parametrize("kernel", KERNELS)
@pytest.mark.parametrize("transposed", [True, False])
def test_estr_matches_mpmath(orc, vpm, kernel, transposed):
    g = golden("estr.json")
    P = vpm.fields.make_particles(g["X"], g["Gamma"], g["sigma"], static=g["static"])
    P[:, 15:24] = np.array(g["J"])
    orc.estr_direct(P, kernel, transposed=transposed)
    ref = np.array(g["%s_%s" % (kernel, "T" if transposed else "C")])
    st = vpm.metrics.err_stats(P[:, 39:42], ref)
    assert st["max"] < 1e-12, st
    # static targets untouched
    assert np.all(P[np.array(g["static"]) != 0, 39:42] == 0)


def test_estr_zero_for_uniform_J(orc, vpm):
    P = _cloud(vpm, n=20)
    P[:, 15:24] = np.arange(9.0)
    orc.estr_direct(P, "gaussianerf")
    assert np.abs(P[:, 39:42]).max() == 0.0


@pytest.mark.parametrize("kernel", KERNELS)
def test_adjoint_matches_mpmath(orc, vpm, kernel):
    g = golden("adjoint.json")
    P = vpm.fields.make_particles(g["X"], g["Gamma"], g["sigma"])
    Xb, Gb, Sb = orc.uj_adjoint(P, kernel, np.array(g["Ubar"]), np.array(g["Jbar"]))
    ref = g[kernel]
    np.testing.assert_allclose(Xb, ref["Xbar"], rtol=1e-9, atol=1e-12 * np.abs(ref["Xbar"]).max())
    np.testing.assert_allclose(Gb, ref["Gbar"], rtol=1e-9, atol=1e-12 * np.abs(ref["Gbar"]).max())
    np.testing.assert_allclose(Sb, ref["Sbar"], rtol=1e-9, atol=1e-12 * max(np.abs(ref["Sbar"]).max(), 1e-30))


def test_adjoint_matches_fd(orc, vpm):
    """Central differences of the exact loss (h = 1e-6) on a larger cloud."""
    n = 200
    P = _cloud(vpm, n=n, seed=4)
    rng = np.random.default_rng(9)
    w, W = rng.normal(size=(n, 3)), rng.normal(size=(n, 9))
    Xb, Gb, Sb = orc.uj_adjoint(P, "gaussianerf", w, W)
    h = 1e-6
    for k in rng.choice(n, 8, replace=False):
        for c in range(7):
            col = [0, 1, 2, 3, 4, 5, 6][c]
            Pp, Pm = P.copy(), P.copy()
            Pp[k, col] += h
            Pm[k, col] -= h
            fd = (orc.loss_local(Pp, "gaussianerf", w, W, k) - orc.loss_local(Pm, "gaussianerf", w, W, k)) / (2 * h)
            an = Xb[k, col] if col < 3 else (Gb[k, col - 3] if col < 6 else Sb[k])
            assert abs(fd - an) <= 2e-6 * max(abs(an), 1e-3 * np.abs(Xb).max()), (k, col, fd, an)


def _random_state(vpm, n, seed):
    """Particle records with random X, Gamma, sigma, U, J and M (the inputs of
    update_particle_states)."""
    rng = np.random.default_rng(seed)
    P = _cloud(vpm, n=n, seed=seed)
    P[:, 9:12] = rng.normal(0, 1, (n, 3))       # U
    P[:, 15:24] = rng.normal(0, 1, (n, 9))      # J
    P[:, 27:36] = rng.normal(0, 1, (n, 9))      # M
    return P


@pytest.mark.parametrize("transposed", [True, False])
def test_update_adjoint_matches_fd(orc, vpm, transposed):
    """Pullback of update_particle_states against central differences of the
    update itself, for every input row it reads (X, Gamma, sigma, M, U, J)."""
    n = 6
    P = _random_state(vpm, n, seed=11)
    P[5, 42] = 1.0                              # one static particle
    rng = np.random.default_rng(12)
    Wout = np.zeros_like(P)
    for rows in ((0, 3), (3, 6), (6, 7), (27, 30), (30, 33), (34, 35)):
        Wout[:, rows[0]:rows[1]] = rng.normal(0, 1, (n, rows[1] - rows[0]))
    a, b, dt, f, g = -5.0 / 9.0, 15.0 / 16.0, 0.01, 0.1, 0.2
    Uinf = np.array([0.3, -0.1, 0.2])
    zeta0 = orc.zeta("gaussianerf", 0.0) if hasattr(orc, "zeta") else 1.0

    def loss(Q):
        Q = Q.copy()
        orc.update_particle_states(Q, a, b, dt, Uinf, f, g, zeta0, transposed)
        return float(np.sum(Wout * Q))

    B = Wout.copy()
    orc.update_particle_states_adjoint(P, B, a, b, dt, f, g, transposed)
    h = 1e-6
    rows = list(range(0, 7)) + list(range(9, 12)) + list(range(15, 24)) + [27, 28, 29, 30, 31, 32, 34]
    for i in range(n):
        for r in rows:
            Pp, Pm = P.copy(), P.copy()
            Pp[i, r] += h
            Pm[i, r] -= h
            fd = (loss(Pp) - loss(Pm)) / (2 * h)
            an = B[i, r]
            assert abs(fd - an) <= 1e-7 * max(1.0, abs(an)), (i, r, fd, an)


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_rk3_step_adjoint_matches_fd(orc, vpm, kernel):
    """Pullback of one whole rk3 step (three UJ evaluations and updates)
    against central differences of a linear functional of the state after
    the step, on X, Gamma and sigma of the state before it."""
    n = 30
    P = _cloud(vpm, n=n, seed=21, sig=(0.15, 0.3))
    P[:, 3:6] *= 0.05
    rng = np.random.default_rng(22)
    Wout = np.zeros_like(P)
    Wout[:, 0:7] = rng.normal(0, 1, (n, 7))
    s = orc.make_settings(kernel=kernel, f=0.0, g=0.2, relax=0, Uinf=(0.1, 0.0, -0.05))
    dt = 0.05

    def loss(Q):
        Q = Q.copy()
        orc.rk3_step(Q, s, dt, relax_now=False)
        return float(np.sum(Wout[:, 0:7] * Q[:, 0:7]))

    B = Wout.copy()
    orc.rk3_step_adjoint(P, s, dt, B)
    assert np.all(B[:, 9:36] == 0.0)
    h = 1e-6
    scale = np.abs(B[:, 0:7]).max()
    for i in rng.choice(n, 8, replace=False):
        for r in range(7):
            Pp, Pm = P.copy(), P.copy()
            Pp[i, r] += h
            Pm[i, r] -= h
            fd = (loss(Pp) - loss(Pm)) / (2 * h)
            an = B[i, r]
            assert abs(fd - an) <= 1e-6 * max(abs(an), 1e-3 * scale), (i, r, fd, an)


def test_rk3_step_adjoint_rejects_sfs(orc, vpm):
    P = _cloud(vpm, n=5)
    s = orc.make_settings(sfs=1)
    with pytest.raises(ValueError):
        orc.rk3_step_adjoint(P, s, 0.01, np.zeros_like(P))


def test_adjoint_scale_bounds_adjoint(orc, vpm):
    """The summation scale dominates the adjoint componentwise and equals it
    when every contribution has one sign (a two-particle field)."""
    P = _cloud(vpm, n=50, seed=8)
    rng = np.random.default_rng(8)
    w, W = rng.normal(size=(50, 3)), rng.normal(size=(50, 9))
    for kernel in KERNELS:
        got = orc.uj_adjoint(P, kernel, w, W)
        mag = orc.uj_adjoint_scale(P, kernel, w, W)
        for a, m in zip(got, mag):
            assert np.all(np.abs(a) <= m * (1 + 1e-12) + 1e-300)
    P2 = P[:2].copy()
    w2, W2 = np.zeros((2, 3)), np.zeros((2, 9))
    w2[0] = (1.0, 0.0, 0.0)
    got = orc.uj_adjoint(P2, "gaussianerf", w2, W2)
    mag = orc.uj_adjoint_scale(P2, "gaussianerf", w2, W2)
    np.testing.assert_allclose(np.abs(got[1]), mag[1], rtol=1e-14)


@pytest.mark.parametrize("kernel", KERNELS)
def test_zeta_direct_matches_mpmath(orc, vpm, kernel):
    """zeta_direct against 30-digit brute force of the closed-form zeta
    (kernel.jl:45,51,60,69-74) on a small cloud with a duplicated particle;
    J[1:3] of static particles is written too and the other J rows untouched."""
    import mpmath as mp
    mp.mp.dps = 30
    P = _cloud(vpm, n=24, seed=31)
    P[5, 0:3] = P[4, 0:3]                   # coincident pair: the singular delta sees both
    P[7, 42] = 1.0
    P[:, 18:24] = 7.0
    Q = P.copy()
    orc.zeta_direct(Q, kernel)
    c1 = 1 / (2 * mp.pi) ** mp.mpf(1.5)
    def zeta(q):
        if kernel == "singular":
            return mp.mpf(1) if q == 0 else mp.mpf(0)
        if kernel == "gaussianerf":
            return c1 * mp.exp(-q * q / 2)
        if kernel == "gaussian":
            return 3 / (4 * mp.pi) * mp.exp(-q ** 3)
        return 7.5 / (4 * mp.pi) / (q * q + 1) ** mp.mpf(3.5)
    for i in range(24):
        acc = [mp.mpf(0)] * 3
        for j in range(24):
            d = [mp.mpf(P[i, k]) - mp.mpf(P[j, k]) for k in range(3)]
            r = mp.sqrt(sum(x * x for x in d))
            s = mp.mpf(P[j, 6])
            z = zeta(r / s) / s ** 3
            acc = [acc[k] + mp.mpf(P[j, 3 + k]) * z for k in range(3)]
        ref = np.array([float(a) for a in acc])
        np.testing.assert_allclose(Q[i, 15:18], ref, rtol=1e-12, atol=1e-14 * max(1.0, np.abs(ref).max()))
    assert np.all(Q[:, 18:24] == 7.0)


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_estr_ranges_matches_estr_direct(orc, vpm, kernel):
    """Estr_fmm over the single pair (all, all) equals Estr_direct! when no
    particle is static; split into sub-ranges it sums to the same."""
    P = _cloud(vpm, n=40, seed=12)
    P[:, 15:24] = np.random.default_rng(13).normal(size=(40, 9))
    A = P.copy()
    orc.estr_direct(A, kernel)
    B = P.copy()
    orc.estr_ranges(B, kernel, [0], [40], [0], [40])
    np.testing.assert_allclose(B[:, 39:42], A[:, 39:42], rtol=1e-13, atol=1e-13 * np.abs(A[:, 39:42]).max())
    C = P.copy()
    orc.estr_ranges(C, kernel, [0, 0, 25, 25], [25, 25, 40, 40], [0, 17, 0, 17], [17, 40, 17, 40])
    np.testing.assert_allclose(C[:, 39:42], A[:, 39:42], rtol=1e-12, atol=1e-12 * np.abs(A[:, 39:42]).max())
    # static particles are not skipped: a static target still receives
    D = P.copy()
    D[3, 42] = 1.0
    orc.estr_ranges(D, kernel, [0], [40], [0], [40])
    np.testing.assert_allclose(D[:, 39:42], B[:, 39:42], rtol=0, atol=0)


@pytest.mark.parametrize("kernel", ["gaussianerf", "gaussian", "winckelmans", "singular"])
def test_zeta_ranges_matches_zeta_direct(orc, vpm, kernel):
    """zeta_fmm's pair loop over the single pair (all, all) onto zeroed J
    rows is zeta_direct (pinned to mpmath), bit for bit (same order); split
    into sub-ranges it sums to the same; it accumulates (no reset) and skips
    no static particle."""
    P = _cloud(vpm, n=40, seed=14)
    P[7] = P[8]
    P[8, 3:6] *= -0.5                          # a coincident pair with its own strength
    A = P.copy()
    orc.zeta_direct(A, kernel)
    B = P.copy()
    B[:, 15:18] = 0.0
    orc.zeta_ranges(B, kernel, [0], [40], [0], [40])
    assert np.array_equal(B[:, 15:18], A[:, 15:18])
    assert np.array_equal(B[:, 18:24], P[:, 18:24])
    C = P.copy()
    C[:, 15:18] = 0.0
    orc.zeta_ranges(C, kernel, [0, 0, 25, 25], [25, 25, 40, 40], [0, 17, 0, 17], [17, 40, 17, 40])
    np.testing.assert_allclose(C[:, 15:18], A[:, 15:18], rtol=1e-12, atol=1e-12 * np.abs(A[:, 15:18]).max())
    D = P.copy()
    D[3, 42] = 1.0
    orc.zeta_ranges(D, kernel, [0], [40], [0], [40])
    np.testing.assert_allclose(D[:, 15:18], P[:, 15:18] + A[:, 15:18], rtol=1e-14,
                               atol=1e-14 * np.abs(A[:, 15:18]).max())


def _lattice(vpm, m=4, h=0.1, sig=1.3, seed=3):
    rng = np.random.default_rng(seed)
    g = np.arange(m) * h
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + rng.uniform(-0.1, 0.1, (m ** 3, 3)) * h
    G = rng.normal(0, 1e-3, (m ** 3, 3))
    P = vpm.fields.make_particles(X, G, np.full(m ** 3, sig * h))
    P[:, 7] = h ** 3                           # vol
    return P


@pytest.mark.parametrize("kernel", ["gaussianerf", "gaussian", "winckelmans"])
def test_core_spreading_reset_equals_dense_solve(orc, vpm, kernel):
    """Iterated to a tight tolerance, the RBF conjugate gradient gives the
    Gamma that solves the zeta system A Gamma = omega_target (dense solve),
    with A_ij = zeta(r_ij/sgm0)/sgm0^3 (self included).  The tolerance stays
    at 1e-7: below an absolute r.r of 2 eps the reference's CG sets beta = 1
    (viscous.jl:411-416) and stalls, which the oracle restates."""
    P = _lattice(vpm)
    n = P.shape[0]
    Q = P.copy()
    orc.zeta_direct(Q, kernel)
    omega = Q[:, 15:18].copy()
    sg0 = 0.07          # 0.7 h: a well-conditioned zeta matrix, so CG reaches 1e-12
    R = P.copy()
    rc, its, res = orc.core_spreading_reset(R, kernel, sg0, itmax=500, tol=1e-7, iterror=True)
    assert rc == 0 and its < 500, (rc, its, res)
    assert np.all(R[:, 6] == sg0)
    X = P[:, 0:3]
    r = np.linalg.norm(X[:, None, :] - X[None, :, :], axis=2) / sg0
    if kernel == "gaussianerf":
        z = np.exp(-r * r / 2) / (2 * np.pi) ** 1.5
    elif kernel == "gaussian":
        z = 3 / (4 * np.pi) * np.exp(-r ** 3)
    else:
        z = 7.5 / (4 * np.pi) / (r * r + 1) ** 3.5
    A = z / sg0 ** 3
    x = np.linalg.solve(A, omega)
    np.testing.assert_allclose(R[:, 3:6], x, rtol=1e-5, atol=1e-5 * np.abs(x).max())


def test_core_spreading_reset_residual_and_static(orc, vpm):
    """At the default tolerance (1e-3) the new strengths reproduce the target
    vorticity to that relative residual per component.  A static particle
    keeps Gamma and sigma (its zeta contribution enters every A p of the
    reference's CG as a constant, so with statics only the recursive residual
    is controlled)."""
    P = _lattice(vpm)
    Q = P.copy()
    orc.zeta_direct(Q, "gaussianerf")
    omega = Q[:, 15:18].copy()
    R = P.copy()
    rc, its, res = orc.core_spreading_reset(R, "gaussianerf", 0.07, itmax=50)
    assert rc == 0 and np.all(res <= 1e-3), (rc, its, res)
    S = R.copy()
    orc.zeta_direct(S, "gaussianerf")
    for c in range(3):
        e = np.linalg.norm(S[:, 15 + c] - omega[:, c]) / np.linalg.norm(omega[:, c])
        assert e <= 1.5e-3, (c, e)
    P2 = P.copy()
    P2[10, 42] = 1.0
    R2 = P2.copy()
    rc, its, res = orc.core_spreading_reset(R2, "gaussianerf", 0.07, itmax=50)
    assert rc == 0 and np.all(res <= 1e-3)
    assert np.array_equal(R2[10, 3:7], P2[10, 3:7])
    # with iterror, an unconverged solve reports status 1
    R3 = P.copy()
    rc, its, res = orc.core_spreading_reset(R3, "gaussianerf", 0.07, itmax=2)
    assert rc == 1 and its == 2


def test_rk3_step_runs_core_spreading_reset(orc, vpm):
    """When the core growth reaches beta in the last stage, the step resets
    sigma to sgm0 and t_sgm to 0 (viscous.jl:202-220), and the new strengths
    reproduce the zeta field of the grown cores to the CG tolerance; with
    rbf_itmax = 0 the reset is reported (status 1) and not performed."""
    P = _lattice(vpm, sig=0.7)
    sg0 = float(P[0, 6])
    s = orc.make_settings(kernel="gaussianerf", relax=0, viscous=1, nu=1e-3, sgm0=sg0, beta=1.001,
                          rbf_itmax=50)
    R = P.copy()
    assert orc.rk3_step(R, s, 0.01, relax_now=False) == 0
    assert s.t_sgm == 0.0 and np.all(R[:, 6] == sg0)
    omega = R[:, 33:36].copy()                 # target: zeta field of the grown cores (M[7:9])
    S = R.copy()
    orc.zeta_direct(S, "gaussianerf")
    for c in range(3):
        assert np.linalg.norm(S[:, 15 + c] - omega[:, c]) <= 1.5e-3 * np.linalg.norm(omega[:, c])
    s0 = orc.make_settings(kernel="gaussianerf", relax=0, viscous=1, nu=1e-3, sgm0=sg0, beta=1.001,
                           rbf_itmax=0)
    R0 = P.copy()
    assert orc.rk3_step(R0, s0, 0.01, relax_now=False) == 1
    assert s0.t_sgm == pytest.approx(0.01) and np.all(R0[:, 6] > sg0)


def _adjoint_identities(P, U, J, Ubar, Jbar, Xb, Gb, Sb):
    tu, tj = np.sum(Ubar * U, axis=1), np.sum(Jbar * J, axis=1)
    L = np.sum(tu) + np.sum(tj)
    lhs = np.sum(Gb * P[:, 3:6], axis=1)
    e_lin = abs(np.sum(lhs) - L) / (np.sum(np.abs(lhs)) + np.sum(np.abs(tu)) + np.sum(np.abs(tj)))
    e_tr = (np.abs(np.sum(Xb, axis=0)) / np.sum(np.abs(Xb), axis=0)).max()
    Xc = P[:, 0:3] - P[:, 0:3].mean(axis=0)
    hx, hs, rhs = np.sum(Xb * Xc, axis=1), Sb * P[:, 6], -2.0 * tu - 3.0 * tj
    e_hom = abs(np.sum(hx) + np.sum(hs) - np.sum(rhs)) / (np.sum(np.abs(hx)) + np.sum(np.abs(hs)) + np.sum(np.abs(rhs)))
    return e_lin, e_tr, e_hom


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "gaussian", "singular"])
def test_adjoint_identities_hold_and_catch_mistakes(orc, vpm, kernel):
    """The size-independent identities the full-size GPU test applies to the
    adjoint (linearity in Gamma, translation invariance, homogeneity of
    degree -2 / -3 under (X, sigma) scaling) hold for the oracle to rounding,
    and each fails when one part of the gradient is planted wrong.  The
    weights are the outputs themselves (Ubar = U, Jbar = J, held fixed), so
    that L = sum |U|^2 + sum |J|^2 has no cancellation and a planted error
    shows at its own size."""
    n = 300
    P = vpm.workloads.random_cloud(n, seed=9)
    P[:, 6] *= 5.0
    Q = P.copy()
    orc.uj_direct(Q, kernel)
    U, J = Q[:, 9:12], Q[:, 15:24]
    Ubar, Jbar = U.copy(), J.copy()
    Xb, Gb, Sb = orc.uj_adjoint(P, kernel, Ubar, Jbar)
    ok = _adjoint_identities(P, U, J, Ubar, Jbar, Xb, Gb, Sb)
    assert max(ok) < 1e-12, ok
    bad = []
    if kernel != "singular":                                       # no sigma dependence there
        bad.append(_adjoint_identities(P, U, J, Ubar, Jbar, Xb, Gb, 0 * Sb)[2])       # sigma term dropped
    bad.append(_adjoint_identities(P, U, J, Ubar, Jbar, Xb, 0.99 * Gb, Sb)[0])        # Gamma gradient scaled
    Xd = Xb.copy()
    Xd[np.argmax(np.linalg.norm(Xb, axis=1))] = 0.0
    bad.append(_adjoint_identities(P, U, J, Ubar, Jbar, Xd, Gb, Sb)[1])               # one particle's dX dropped
    assert min(bad) > 1e-4, bad


# ---- per-particle schemes: properties fixed by the mathematics -----------
# (timeintegration.jl:388-532, relaxation.jl:62-142, viscous.jl:167-180,
# subfilterscale.jl:541-673 and 204-250)

def test_rk3_uniform_advection_is_exact(orc, vpm):
    """A lone particle (no self-interaction: U = 0, J = 0) in a free stream
    moves by exactly dt U_inf per step: the low-storage coefficients sum to
    one (1/3 + 15/16 4/9 + 8/15 15/32 = 1); Gamma and sigma do not change."""
    P = vpm.fields.make_particles([[0.1, -0.2, 0.3]], [[0.4, 0.5, -0.6]], 0.05)
    Uinf = (0.3, -0.2, 0.1)
    s = orc.make_settings(kernel="gaussianerf", f=0.0, g=0.2, relax=0, Uinf=Uinf)
    R = P.copy()
    for _ in range(4):
        orc.rk3_step(R, s, 0.05, relax_now=False)
    np.testing.assert_allclose(R[0, 0:3], P[0, 0:3] + 4 * 0.05 * np.array(Uinf), rtol=0, atol=1e-15)
    assert np.array_equal(R[0, 3:7], P[0, 3:7])


def test_rk3_third_order_convergence(orc, vpm):
    """On 60 strongly interacting particles (rVPM, stretching on, no
    relaxation; they move about a tenth of the box by T = 1), the RK3 error
    of X and Gamma at fixed time falls by 2^3 each time dt halves (measured
    8.2, 8.1, 8.1); a wrong stage coefficient or a missed register reset
    loses an order."""
    rng = np.random.default_rng(15)
    n = 60
    P = vpm.fields.make_particles(rng.uniform(0, 1, (n, 3)), rng.normal(0, 2e-2, (n, 3)), 0.15)
    s = orc.make_settings(kernel="gaussianerf", f=0.0, g=0.2, relax=0)
    T = 1.0

    def run(nsteps):
        R = P.copy()
        for _ in range(nsteps):
            orc.rk3_step(R, s, T / nsteps, relax_now=False)
        return R

    ref = run(512)
    err = [np.abs(run(k)[:, 0:6] - ref[:, 0:6]).max() for k in (4, 8, 16, 32)]
    ratios = [err[i] / err[i + 1] for i in range(3)]
    assert all(7.0 < r < 9.0 for r in ratios), (err, ratios)


def test_relaxation_properties(orc, vpm):
    """Pedrizzetti (relaxation.jl:62-81): rlxf = 0 changes nothing and
    rlxf = 1 aligns Gamma with the vorticity omega = (J32 - J23, J13 - J31,
    J21 - J12) keeping |Gamma|; the corrected scheme (:117-142) keeps |Gamma|
    for any rlxf; omega = 0 (symmetric J) leaves Gamma as it is."""
    rng = np.random.default_rng(12)
    n = 50
    P = vpm.workloads.random_cloud(n, seed=13)
    P[:, 3:6] = rng.normal(size=(n, 3))
    P[:, 15:24] = rng.normal(size=(n, 9))
    S = P[0, 15:24].reshape(3, 3)
    P[0, 15:24] = (S + S.T).reshape(-1)                   # omega = 0 at particle 0
    J = P[:, 15:24]
    w = np.stack([J[:, 5] - J[:, 7], J[:, 6] - J[:, 2], J[:, 1] - J[:, 3]], axis=1)
    g0 = np.linalg.norm(P[:, 3:6], axis=1)
    R = P.copy()
    orc.relax(R, 0.0)
    np.testing.assert_allclose(R[:, 3:6], P[:, 3:6], rtol=1e-15, atol=0)
    R = P.copy()
    orc.relax(R, 1.0)
    wn = np.linalg.norm(w[1:], axis=1)
    np.testing.assert_allclose(R[1:, 3:6], g0[1:, None] * w[1:] / wn[:, None], rtol=1e-13, atol=1e-15)
    assert np.array_equal(R[0, 3:6], P[0, 3:6])
    for rlxf in (0.05, 0.3, 0.7):
        R = P.copy()
        orc.relax(R, rlxf, corrected=True)
        np.testing.assert_allclose(np.linalg.norm(R[:, 3:6], axis=1), g0, rtol=1e-13)


def test_core_spreading_advances_sigma_squared_linearly(orc, vpm):
    """With f = g = 0 (no stretching term in the sigma update), a
    CoreSpreading rk3 step advances sigma^2 by exactly 2 nu dt (the RK3
    registers integrate the constant rate d sigma^2/dt = 2 nu); static
    particles keep their sigma; t_sgm advances by dt."""
    P = vpm.workloads.random_cloud(120, seed=14)
    P[:, 6] *= 3.0
    P[[5, 50], 42] = 1.0
    nu, dt = 1e-3, 0.01
    s = orc.make_settings(kernel="gaussianerf", f=0.0, g=0.0, relax=0, viscous=1, nu=nu, sgm0=0.05,
                          beta=100.0)
    R = P.copy()
    assert orc.rk3_step(R, s, dt, relax_now=False) == 0
    live = P[:, 42] == 0
    np.testing.assert_allclose(R[live, 6] ** 2 - P[live, 6] ** 2, 2 * nu * dt, rtol=1e-10)
    assert np.array_equal(R[~live, 6], P[~live, 6])
    assert s.t_sgm == pytest.approx(dt)


def test_dynamic_sfs_coefficient_rules(orc, vpm):
    """dynamicprocedure_pseudo3level_afterUJ (subfilterscale.jl:541-673) and
    clipping_backscatter (:225-247, 287-296) on a two-ring cloud: every
    non-static coefficient satisfies |C| <= maxC, |C| >= minC unless its
    numerator vanished, C >= 0 with force_positive, C = C2/C3 up to the sign
    taken by force_positive (or 0 when clipped), C (Gamma . SFS) >= 0 after
    clipping, and M is flushed."""
    P = vpm.workloads.leapfrog_cfg2(Nphi=24)
    n = P.shape[0]
    P[[3, 40], 42] = 1.0
    minC, maxC = 0.01, 0.6
    R = P.copy()
    assert orc.sfs_before(R, "gaussianerf", True, 0.999, minC, maxC) == 0
    orc.uj_direct(R, "gaussianerf", reset=True, reset_sfs=True, sfs=True)
    assert orc.sfs_after(R, "gaussianerf", True, 0.999, 0.005, minC, maxC, force_positive=True, clip=True) == 0
    live = R[:, 42] == 0
    C = R[live, 36]
    ratio = R[live, 37] / R[live, 38]
    gs = np.sum(R[live, 3:6] * R[live, 39:42], axis=1)
    clipped = C == 0
    assert np.all(C >= 0) and np.all(C <= maxC * (1 + 1e-14))
    assert np.all(C[~clipped] >= minC * (1 - 1e-14))
    np.testing.assert_allclose(C[~clipped], np.abs(ratio[~clipped]), rtol=1e-15)
    assert np.all(C * gs >= 0)
    assert np.all(R[live, 27:36] == 0)
    assert 0 < np.count_nonzero(~clipped) < n


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "gaussian", "singular"])
def test_uj_adjoint_equals_reference_pullback(orc, vpm, kernel):
    """The closed-form adjoint (pinned by central differences and mpmath)
    equals the reference's own CPU gradient, UJS_pullback of rrules.jl:183-326
    restated literally, to FP64 rounding for every output, although the
    reference's comments doubt parts of it (rrules.jl:233, 237)."""
    n = 40
    P = vpm.workloads.random_cloud(n, seed=31)
    P[:, 6] *= 8.0
    P[7, 0:3] = P[8, 0:3]                     # a coincident pair (skipped by both)
    rng = np.random.default_rng(32)
    Ub, Jb = rng.normal(size=(n, 3)), rng.normal(size=(n, 9))
    a = orc.uj_adjoint(P, kernel, Ub, Jb)
    b = orc.uj_adjoint_rrules(P, kernel, Ub, Jb)
    for name, x, y in zip(("X", "Gamma", "sigma"), a, b):
        x, y = np.asarray(x).reshape(n, -1), np.asarray(y).reshape(n, -1)
        if kernel == "singular" and name == "sigma":
            assert np.all(x == 0) and np.all(y == 0)
            continue
        e = np.linalg.norm(x - y) / np.linalg.norm(x)
        assert e < 1e-13, (name, e)


def test_update_pullback_vs_reference_state_pullback(orc, vpm):
    """The update pullback (derived from the rVPM update, pinned by central
    differences) against the reference's own state_pullback
    (rrules.jl:436-547) restated literally, particle by particle (its
    Gammabar reads the first particle's register cotangent, :526, so each
    call holds one particle): X, U, the registers and sigma agree to
    rounding; Gamma and J differ, because the reference differentiates the
    -3 MM4 Gamma_zeta term only at zeta = eta (:521, :524); with the whole
    sum there the two agree to rounding on every output."""
    n = 12
    rng = np.random.default_rng(41)
    P = vpm.workloads.random_cloud(n, seed=42)
    P[:, 3:6] = rng.normal(size=(n, 3))
    P[:, 15:24] = rng.normal(size=(n, 9))
    B0 = np.zeros_like(P)
    for rows in (slice(0, 7), slice(27, 33), slice(34, 35)):
        B0[:, rows] = rng.normal(size=(n, rows.stop - rows.start))
    a, b, dt, f, g = -5.0 / 9.0, 15.0 / 16.0, 0.01, 0.0, 0.2
    zeta0 = orc.zeta("gaussianerf", 0.0)   # unused without SFS (C = 0)
    ours = B0.copy()
    orc.update_particle_states_adjoint(P, ours, a, b, dt, f, g, transposed=True)
    lit, fix = B0.copy(), B0.copy()
    for i in range(n):
        for R, full in ((lit, False), (fix, True)):
            Pi, Bi = P[i:i + 1].copy(), R[i:i + 1].copy()
            orc.update_particle_states_pullback_rrules(Pi, Bi, a, b, dt, f, g, zeta0, full=full)
            R[i] = Bi[0]
    same = np.r_[0:3, 6, 9:12, 27:33, 34]
    np.testing.assert_allclose(lit[:, same], ours[:, same], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(fix[:, np.r_[0:12, 15:24, 27:33, 34]], ours[:, np.r_[0:12, 15:24, 27:33, 34]],
                               rtol=1e-12, atol=1e-14)
    dG = np.linalg.norm(lit[:, 3:6] - ours[:, 3:6]) / np.linalg.norm(ours[:, 3:6])
    dJ = np.linalg.norm(lit[:, 15:24] - ours[:, 15:24]) / np.linalg.norm(ours[:, 15:24])
    assert dG > 1e-4 and dJ > 1e-4, (dG, dJ)


@pytest.mark.parametrize("transposed", [True, False])
@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
def test_dynamic_sfs_coefficient_from_independent_filters(vpm, orc, transposed, kernel):
    """The oracle's dynamic procedure (sfs_before, the domain-filter UJ with
    E_str, sfs_after; subfilterscale.jl:447-673) against the coefficient
    rebuilt from independent evaluations of the two filter widths: U/J and
    E_str of the field with every non-static sigma scaled by alpha (test
    filter) and unscaled (domain filter), stretching from metrics.stretching,
    then
        nume = (3 alpha - 2) (S_test - S_dom) . Gamma,
        deno = (E_test - E_dom) . Gamma / (zeta0 / sigma^3),   C = nume / deno
    with rlxf = 1 (no Lagrangian memory), no clamping, no clipping.  A wrong
    filter direction (sigma / alpha), a missing factor, a row or scheme mix-up
    in the oracle shows here."""
    rng = np.random.default_rng(17)
    n = 60
    P = vpm.fields.make_particles(rng.uniform(0, 1, (n, 3)), rng.normal(0, 1e-2, (n, 3)),
                                  rng.uniform(0.12, 0.2, n))
    P[7, 42] = 1.0                      # one static particle: skipped throughout
    alpha = 0.9
    zeta0 = float(orc.zeta(kernel, 0.0))
    # the oracle's procedure
    Q = P.copy()
    assert orc.sfs_before(Q, kernel, transposed, alpha, 0.0, 1e30) == 0
    orc.uj_direct(Q, kernel, reset=True, reset_sfs=True, sfs=True, transposed=transposed)
    assert orc.sfs_after(Q, kernel, transposed, alpha, 1.0, 0.0, 1e30, force_positive=False, clip=False) == 0
    # the two filter widths, evaluated independently
    live = P[:, 42] == 0
    T = P.copy()
    T[live, 6] *= alpha
    orc.uj_direct(T, kernel, reset=True, reset_sfs=True, sfs=True, transposed=transposed)
    D = P.copy()
    orc.uj_direct(D, kernel, reset=True, reset_sfs=True, sfs=True, transposed=transposed)
    G = P[:, 3:6]
    St = vpm.metrics.stretching(T[:, 15:24], G, transposed)
    Sd = vpm.metrics.stretching(D[:, 15:24], G, transposed)
    nume = (3 * alpha - 2) * np.sum((St - Sd) * G, axis=1)
    deno = np.sum((T[:, 39:42] - D[:, 39:42]) * G, axis=1) / (zeta0 / P[:, 6] ** 3)
    np.testing.assert_allclose(Q[live, 37], nume[live], rtol=1e-9, atol=1e-14 * np.abs(nume).max())
    np.testing.assert_allclose(Q[live, 38], deno[live], rtol=1e-9, atol=1e-14 * np.abs(deno).max())
    np.testing.assert_allclose(Q[live, 36], nume[live] / deno[live], rtol=1e-8)
    assert np.all(Q[live, 27:36] == 0.0)                  # M flushed
    assert np.array_equal(Q[~live][:, 36:39], P[~live][:, 36:39])     # statics keep their C
    # the check has teeth: the filter taken the other way round (sigma / alpha) differs
    W = P.copy()
    W[live, 6] /= alpha
    orc.uj_direct(W, kernel, reset=True, reset_sfs=True, sfs=True, transposed=transposed)
    Sw = vpm.metrics.stretching(W[:, 15:24], G, transposed)
    wrong = (3 * alpha - 2) * np.sum((Sw - Sd) * G, axis=1)
    assert np.max(np.abs(wrong[live] - Q[live, 37])) > 1e-3 * np.abs(Q[live, 37]).max()


# ---- the stretching term and the rVPM update: invariants of the ODE ------
# (timeintegration.jl:479-526).  Per particle the update integrates
#   dX/dt = U + Uinf,
#   dGamma/dt = S - 3 Z Gamma - E,   dsigma/dt = -sigma Z,
#   Z = [(f+g)/(1+3f) S.Gamma - f/(1+3f) E.Gamma] / |Gamma|^2,
#   E = C SFS sigma^3 / zeta0,
# with S = J^T Gamma (transposed) or J Gamma (classic).  From these:
#  * d ln|Gamma| / dt = [(1-3g) S.Gamma - E.Gamma] / ((1+3f) |Gamma|^2) and
#    d ln sigma / dt = -Z, so without SFS |Gamma| sigma^((1-3g)/(f+g)) is
#    constant for any S, and with g = 0 |Gamma|^f sigma is constant for any
#    S and E;
#  * with uniform sigma the transposed stretching of a pair (p, q) is
#    H(d).(Gamma_q x Gamma_p) d + F (Gamma_q x Gamma_p) on p and the negative
#    of it on q (H, F even in d), so sum_p S_p = 0 and, with f = g = 0
#    (Z = 0) and no SFS, sum_p Gamma_p is conserved by any Runge-Kutta
#    scheme (a linear invariant); the classic J Gamma has no such symmetry.
# The mutants below restate the update with one plausible mistake each and
# show that the checks catch it.

def _update_np(P, a, b, dt, Uinf, f, g, zeta0, transposed, mutant=None):
    """The rVPM update restated in numpy (timeintegration.jl:479-526), with
    an optional planted mistake: "2Z" (2 Z Gamma for 3 Z Gamma), "+sZ" (the
    sigma register gains +dt sigma Z), "coefZ" ((f+g)/(1+3g) in Z), "sfsZ"
    (the SFS part of Z with the sign flipped), "classic" (J Gamma under the
    transposed flag)."""
    live = P[:, 42] == 0
    Q = P[live]
    M, G, J, U = Q[:, 27:36], Q[:, 3:6], Q[:, 15:24], Q[:, 9:12]
    C, sfs, sg = Q[:, 36], Q[:, 39:42], Q[:, 6].copy()
    M[:, 0:3] = a * M[:, 0:3] + dt * (U + np.asarray(Uinf))
    Q[:, 0:3] += b * M[:, 0:3]
    if transposed and mutant != "classic":
        S = np.stack([np.sum(J[:, 3 * k:3 * k + 3] * G, axis=1) for k in range(3)], 1)
    else:
        S = np.stack([np.sum(J[:, k::3] * G, axis=1) for k in range(3)], 1)
    Gn2 = np.sum(G * G, axis=1)
    cz = (f + g) / (1 + 3 * g) if mutant == "coefZ" else (f + g) / (1 + 3 * f)
    cs = -f / (1 + 3 * f) if mutant == "sfsZ" else f / (1 + 3 * f)
    E = C[:, None] * sfs * (sg ** 3 / zeta0)[:, None]
    Z = np.where(Gn2 > 0, (cz * np.sum(S * G, axis=1) - cs * np.sum(E * G, axis=1)) / np.where(Gn2 > 0, Gn2, 1), 0)
    k3 = 2.0 if mutant == "2Z" else 3.0
    M[:, 3:6] = a * M[:, 3:6] + dt * (S - k3 * Z[:, None] * G - E)
    M[:, 7] = a * M[:, 7] + (dt * sg * Z if mutant == "+sZ" else -dt * sg * Z)
    Q[:, 27:36] = M
    Q[:, 3:6] = G + b * M[:, 3:6]
    Q[:, 6] = sg + b * M[:, 7]
    P[live] = Q


_RK3 = ((0.0, 1.0 / 3.0), (-5.0 / 9.0, 15.0 / 16.0), (-153.0 / 128.0, 8.0 / 15.0))


def _rk3_np(orc, P, kernel, dt, f, g, transposed, mutant):
    """rungekutta3 without SFS, viscosity or relaxation
    (timeintegration.jl:410-434) over the oracle's U/J and _update_np."""
    P[:, 27:36] = 0.0
    for a, b in _RK3:
        orc.uj_direct(P, kernel, reset=True, transposed=transposed)
        _update_np(P, a, b, dt, (0.0, 0.0, 0.0), f, g, 1.0, transposed, mutant)


def test_update_np_restatement_equals_oracle(orc, vpm):
    """The numpy restatement used for the mutants is the oracle's update
    (both schemes, SFS term on, a static particle), so a mutant differs from
    the oracle in its planted mistake only."""
    rng = np.random.default_rng(60)
    n = 40
    P = vpm.workloads.random_cloud(n, seed=61)
    P[:, 9:12], P[:, 15:24] = rng.normal(size=(n, 3)), rng.normal(size=(n, 9))
    P[:, 27:36], P[:, 39:42] = rng.normal(size=(n, 9)), rng.normal(size=(n, 3))
    P[:, 36] = rng.uniform(0, 1, n)
    P[3, 42] = 1.0
    P[4, 3:6] = 0.0
    for transposed in (True, False):
        A, B = P.copy(), P.copy()
        orc.update_particle_states(A, -5 / 9, 15 / 16, 0.01, [0.1, 0.2, 0.3], 0.1, 0.2, 0.06, transposed)
        _update_np(B, -5 / 9, 15 / 16, 0.01, [0.1, 0.2, 0.3], 0.1, 0.2, 0.06, transposed)
        np.testing.assert_allclose(B, A, rtol=1e-13, atol=1e-16)
    # and the RK3 driver over it is the oracle's rk3_step
    A = vpm.fields.make_particles(P[:, 0:3], P[:, 3:6] * 1e-2, 0.15)
    B = A.copy()
    orc.rk3_step(A, orc.make_settings(kernel="gaussianerf", f=0.0, g=0.2, relax=0), 0.05, relax_now=False)
    _rk3_np(orc, B, "gaussianerf", 0.05, 0.0, 0.2, True, None)
    np.testing.assert_allclose(B[:, 0:7], A[:, 0:7], rtol=1e-13, atol=1e-16)


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "gaussian", "singular"])
def test_transposed_stretching_conserves_total_circulation(orc, vpm, kernel):
    """cVPM (f = g = 0), uniform sigma, inviscid, no SFS or relaxation: 10 RK3
    steps of 60 interacting particles keep sum Gamma to FP64 rounding of the
    step sums (below 1e-13 of sum_p |Gamma_p(T) - Gamma_p(0)|; measured 4e-17
    to 4e-14 against per-particle changes of 9e-3 to 50), while the classic
    scheme (J Gamma) moves it by 1e-2 or more."""
    rng = np.random.default_rng(21)
    n = 60
    P = vpm.fields.make_particles(rng.uniform(0, 1, (n, 3)), rng.normal(0, 2e-2, (n, 3)), 0.15)
    drift, change, scale = {}, {}, {}
    for transposed in (True, False):
        s = orc.make_settings(kernel=kernel, transposed=transposed, f=0.0, g=0.0, relax=0)
        R = P.copy()
        for _ in range(10):
            assert orc.rk3_step(R, s, 0.05, relax_now=False) == 0
        drift[transposed] = np.abs(R[:, 3:6].sum(0) - P[:, 3:6].sum(0)).max()
        change[transposed] = np.abs(R[:, 3:6] - P[:, 3:6]).max()
        scale[transposed] = np.abs(R[:, 3:6] - P[:, 3:6]).sum()
        assert np.all(R[:, 6] == P[:, 6])                 # Z = 0: sigma untouched
    assert change[True] > 5e-3 and change[False] > 5e-3, change
    assert drift[True] < 1e-13 * scale[True], (drift, scale)
    assert drift[False] > 1e-2, drift
    assert drift[False] > 1e10 * drift[True], drift


def _rvpm_invariant(R, f, g):
    return np.linalg.norm(R[:, 3:6], axis=1) * R[:, 6] ** ((1 - 3 * g) / (f + g))


@pytest.mark.parametrize("fg", [(0.0, 0.2), (0.25, 0.25), (0.5, 0.0)])
def test_rvpm_invariant_third_order(orc, vpm, fg):
    """rVPM without SFS keeps |Gamma| sigma^((1-3g)/(f+g)) per particle for any
    stretching, so over a whole rk3_step run its drift is the scheme's
    truncation error: it falls 8x per halving of dt (measured 7.9-8.1) while
    |Gamma| itself changes by 4-12 %.  Planted mistakes in the update (2 Z Gamma,
    +sigma Z, (f+g)/(1+3g)) break the invariant: their drift is 7e-2 to 0.74
    and does not fall with dt.  ((f+g)/(1+3g) equals the right coefficient
    when f = g, so that mutant is checked at the other two (f, g).)"""
    f, g = fg
    rng = np.random.default_rng(21)
    n = 60
    P = vpm.fields.make_particles(rng.uniform(0, 1, (n, 3)), rng.normal(0, 2e-2, (n, 3)),
                                  rng.uniform(0.12, 0.2, n))
    I0 = _rvpm_invariant(P, f, g)
    s = orc.make_settings(kernel="gaussianerf", transposed=True, f=f, g=g, relax=0)
    errs = []
    for nst in (4, 8, 16, 32):
        R = P.copy()
        for _ in range(nst):
            assert orc.rk3_step(R, s, 1.0 / nst, relax_now=False) == 0
        errs.append(np.abs(_rvpm_invariant(R, f, g) / I0 - 1).max())
    gchange = np.abs(np.linalg.norm(R[:, 3:6], axis=1) / np.linalg.norm(P[:, 3:6], axis=1) - 1).max()
    ratios = [errs[i] / errs[i + 1] for i in range(3)]
    assert gchange > 0.03, gchange
    assert all(7.0 < r < 9.0 for r in ratios), (errs, ratios)
    assert errs[-1] < 1e-6, errs
    mutants = ["2Z", "+sZ"] + (["coefZ"] if f != g else [])
    for mutant in mutants:
        me = []
        for nst in (8, 16):
            R = P.copy()
            for _ in range(nst):
                _rk3_np(orc, R, "gaussianerf", 1.0 / nst, f, g, True, mutant)
            me.append(np.abs(_rvpm_invariant(R, f, g) / I0 - 1).max())
        assert me[1] > 1e3 * errs[2] and me[0] / me[1] < 1.5, (mutant, me, errs)


@pytest.mark.parametrize("case", ["nosfs_rvpm", "nosfs_tube", "sfs_g0"])
def test_update_local_invariant_second_order(orc, vpm, case):
    """One update call (a = 0, b = 1: an Euler step of the ODE) on random
    records with random J, SFS and C changes the invariant by O(dt^2) only:
    4x less per halving of dt.  Without SFS the invariant is
    |Gamma| sigma^((1-3g)/(f+g)); with SFS and g = 0 it is |Gamma|^f sigma,
    which also fixes the SFS part of Z against the SFS part of dGamma/dt.
    Every planted mistake changes it at O(dt) (2x per halving)."""
    rng = np.random.default_rng(62)
    n = 200
    P = vpm.workloads.random_cloud(n, seed=63)
    P[:, 3:6] = rng.normal(size=(n, 3))
    P[:, 15:24] = rng.normal(size=(n, 9))
    P[:, 6] = rng.uniform(0.5, 1.5, n)
    if case == "sfs_g0":
        f, g = 0.5, 0.0
        P[:, 39:42] = rng.normal(size=(n, 3))
        P[:, 36] = rng.uniform(0.1, 1.0, n)
        inv = lambda R: np.linalg.norm(R[:, 3:6], axis=1) ** f * R[:, 6]
    else:
        f, g = (0.0, 0.2) if case == "nosfs_rvpm" else (0.5, 0.0)
        inv = lambda R: _rvpm_invariant(R, f, g)
    zeta0 = 0.7
    I0 = inv(P)

    def change(dt, mutant):
        R = P.copy()
        if mutant is None:
            orc.update_particle_states(R, 0.0, 1.0, dt, [0.0, 0.0, 0.0], f, g, zeta0, True)
        else:
            _update_np(R, 0.0, 1.0, dt, (0.0, 0.0, 0.0), f, g, zeta0, True, mutant)
        return np.abs(np.log(inv(R) / I0)).max()

    e = [change(dt, None) for dt in (2e-3, 1e-3, 5e-4)]
    assert all(3.6 < e[i] / e[i + 1] < 4.4 for i in range(2)), e
    mutants = ["2Z", "+sZ", "coefZ"] + (["sfsZ"] if case == "sfs_g0" else [])
    for mutant in mutants:
        m = [change(dt, mutant) for dt in (1e-3, 5e-4)]
        assert 1.8 < m[0] / m[1] < 2.2 and m[1] > 10 * e[2], (mutant, m, e)


def test_transposed_circulation_mutant_caught(orc, vpm):
    """The classic product J Gamma put where the transposed J^T Gamma belongs
    (the one mistake the rVPM invariant cannot see: S.Gamma is the same for
    both) moves sum Gamma at once."""
    rng = np.random.default_rng(21)
    n = 60
    P = vpm.fields.make_particles(rng.uniform(0, 1, (n, 3)), rng.normal(0, 2e-2, (n, 3)), 0.15)
    R, M = P.copy(), P.copy()
    for _ in range(3):
        _rk3_np(orc, R, "gaussianerf", 0.05, 0.0, 0.0, True, None)
        _rk3_np(orc, M, "gaussianerf", 0.05, 0.0, 0.0, True, "classic")
    good = np.abs(R[:, 3:6].sum(0) - P[:, 3:6].sum(0)).max()
    bad = np.abs(M[:, 3:6].sum(0) - P[:, 3:6].sum(0)).max()
    assert good < 1e-15 and bad > 1e-3, (good, bad)


def test_sfs_lagrangian_memory(orc, vpm):
    """dynamicprocedure_pseudo3level_afterUJ (subfilterscale.jl:571-617) with
    the same M before each of three calls, no clamping (minC = 0, maxC huge)
    and no clipping: the denominator memory starts at the first deno and
    stays there, and the numerator follows the Lagrangian average
    C2_k = nume (1 - (1 - rlxf)^k) with nume = (3 alpha - 2) M[1:3].Gamma and
    deno = M[4:6].Gamma sigma^3 / zeta0; C = C2 / C3 (|C2 / C3| with
    force_positive); M is flushed after every call."""
    rng = np.random.default_rng(64)
    n = 50
    P = vpm.workloads.random_cloud(n, seed=65)
    P[:, 3:6] = rng.normal(size=(n, 3))
    M = rng.normal(size=(n, 9))
    P[:, 36:39] = 0.0
    P[4, 42] = 1.0
    alpha, rlxf, kernel = 0.999, 0.3, "gaussianerf"
    zeta0 = float(orc.zeta(kernel, 0.0))
    G, sg = P[:, 3:6], P[:, 6]
    nume = (3 * alpha - 2) * np.sum(M[:, 0:3] * G, axis=1)
    deno = np.sum(M[:, 3:6] * G, axis=1) * sg ** 3 / zeta0
    live = P[:, 42] == 0
    for fp in (False, True):
        R = P.copy()
        for k in (1, 2, 3):
            R[live, 27:36] = M[live]
            assert orc.sfs_after(R, kernel, True, alpha, rlxf, 0.0, 1e300, force_positive=fp, clip=False) == 0
            c2 = nume * (1 - (1 - rlxf) ** k)
            np.testing.assert_allclose(R[live, 37], c2[live], rtol=1e-13)
            np.testing.assert_allclose(R[live, 38], deno[live], rtol=1e-13)
            c1 = c2 / deno
            np.testing.assert_allclose(R[live, 36], (np.abs(c1) if fp else c1)[live], rtol=1e-12)
            assert np.all(R[live, 27:36] == 0.0)
        assert np.array_equal(R[~live], P[~live])


@pytest.mark.parametrize("transposed", [True, False])
def test_estr_scale_bounds_estr(orc, vpm, transposed):
    """The E_str summation scale s_i = sum zeta |S_ij| (the bound the F32C
    E_str tests are held to) dominates |E_str_i| for every particle, equals
    it when each target has a single source term, skips static particles as
    Estr_direct! does, and over a list of sub-ranges sums to the all-pairs
    value."""
    n = 400
    P = vpm.workloads.random_cloud(n, seed=51)
    P[:, 6] *= 3.0
    P[:, 15:24] = np.random.default_rng(52).normal(size=(n, 9))
    P[[7, 90], 42] = 1.0
    ref = P.copy()
    orc.estr_direct(ref, "gaussianerf", transposed=transposed)
    s = orc.estr_scale(P, "gaussianerf", transposed=transposed)
    e = np.linalg.norm(ref[:, 39:42], axis=1)
    assert np.all(e <= s * (1 + 1e-12))
    assert np.all(s[[7, 90]] == 0.0) and np.all(s[np.arange(n) != 7] >= 0)
    assert np.median(s / np.maximum(e, 1e-300)) > 1.5          # the sums do cancel
    Q = P[:2].copy()
    Q[:, 42] = 0.0
    r2 = Q.copy()
    orc.estr_direct(r2, "gaussianerf", transposed=transposed)
    np.testing.assert_allclose(orc.estr_scale(Q, "gaussianerf", transposed=transposed),
                               np.linalg.norm(r2[:, 39:42], axis=1), rtol=1e-14)
    P[:, 42] = 0.0
    cuts = [0, 37, 150, 151, 400]
    tb, te, sb, se = [], [], [], []
    for a, b in zip(cuts[:-1], cuts[1:]):
        for c, d in zip(cuts[:-1], cuts[1:]):
            tb.append(a); te.append(b); sb.append(c); se.append(d)
    np.testing.assert_allclose(orc.estr_scale_ranges(P, "gaussianerf", tb, te, sb, se, transposed=transposed),
                               orc.estr_scale(P, "gaussianerf", transposed=transposed), rtol=1e-12)


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "gaussian", "singular"])
def test_u_scale_bounds_u(orc, vpm, kernel):
    """The U summation scale s_i = sum_j |U term_ij| over a direct list (the
    bound the full-size F32C list test holds U to) dominates |U_i|, equals it
    when a target has one source, skips coincident pairs as fmm.jl:118 does,
    is additive over sub-lists, and is independent of the sign of Gamma (so
    a symmetric neighbourhood, whose U cancels, keeps its full scale)."""
    n = 300
    P = vpm.workloads.random_cloud(n, seed=53)
    P[:, 6] *= 3.0
    P[:, 9:24] = 0.0
    full = ([0], [n], [0], [n])
    ref = P.copy()
    orc.direct_ranges(ref, kernel, *full)
    s = orc.u_scale_ranges(P, kernel, *full)
    u = np.linalg.norm(ref[:, 9:12], axis=1)
    assert np.all(u <= s * (1 + 1e-12))
    assert np.median(s / np.maximum(u, 1e-300)) > 1.5          # the sums do cancel
    # one source per target: the scale is |U|
    one = ([0, 1], [1, 2], [5, 9], [6, 10])
    r1 = P.copy()
    orc.direct_ranges(r1, kernel, *one)
    np.testing.assert_allclose(orc.u_scale_ranges(P, kernel, *one)[:2], np.linalg.norm(r1[:2, 9:12], axis=1),
                               rtol=1e-14)
    # a coincident source adds nothing
    Q = P.copy()
    Q[1, 0:3] = Q[0, 0:3]
    assert orc.u_scale_ranges(Q, kernel, [0], [1], [1], [2])[0] == 0.0
    # additive over sub-lists, and blind to the sign of Gamma
    cuts = [0, 37, 150, 151, 300]
    tb, te, sb, se = [], [], [], []
    for a, b in zip(cuts[:-1], cuts[1:]):
        for c, d in zip(cuts[:-1], cuts[1:]):
            tb.append(a); te.append(b); sb.append(c); se.append(d)
    np.testing.assert_allclose(orc.u_scale_ranges(P, kernel, tb, te, sb, se), s, rtol=1e-12)
    M = P.copy()
    M[:, 3:6] *= -1.0
    np.testing.assert_allclose(orc.u_scale_ranges(M, kernel, *full), s, rtol=1e-14)
