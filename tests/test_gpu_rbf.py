"""GPU parity of the CoreSpreading reset (vpm_core_spreading_reset;
src/FLOWVPM_viscous.jl:202-220 and rbf_conjugategradient, :314-478).

F64: the device CG runs the oracle's iteration with the same FP64 pair
arithmetic; only the summation order of zeta and of the dot products
differs, so the iteration count matches and the state agrees to
RBF_F64_TOL relative to max |Gamma|.
F32C: the zeta field has FP32 pair arithmetic (scaled error < 1e-5), so the
CG takes its own path to the same tolerance.  Its strengths are within
RBF_F32C_SOL of the oracle's (relative norm per component), and without
static particles they reproduce the oracle's target vorticity, through the
oracle's FP64 zeta, to the CG tolerance (RBF_F32C_RESID).  With static
particles the reference's CG controls only its recursive residual (the
statics add a constant to every A p), so the true residual is not bounded
there and is not checked.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

RBF_F64_TOL = 1e-10
RBF_F32C_RESID = 1.5e-3
RBF_F32C_SOL = 2e-3


def _say(*a):
    print(*a, flush=True)


def _lattice(vpm, m, h=0.1, sig=1.3, seed=3, statics=()):
    rng = np.random.default_rng(seed)
    g = np.arange(m) * h
    X = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) + rng.uniform(-0.1, 0.1, (m ** 3, 3)) * h
    G = rng.normal(0, 1e-3, (m ** 3, 3))
    P = vpm.fields.make_particles(X, G, np.full(m ** 3, sig * h))
    P[:, 7] = h ** 3
    for i in statics:
        P[i, 42] = 1.0
    return P


def _target(orc, P, kernel):
    Q = P.copy()
    orc.zeta_direct(Q, kernel)
    return Q[:, 15:18].copy()


def _resid(orc, R, omega, kernel, static):
    S = R.copy()
    orc.zeta_direct(S, kernel)
    live = ~static
    return np.array([np.linalg.norm(S[live, 15 + c] - omega[live, c]) / np.linalg.norm(omega[live, c])
                     for c in range(3)])


@pytest.mark.parametrize("kernel", ["gaussianerf", "gaussian", "winckelmans"])
@pytest.mark.parametrize("m,statics", [(4, True), (11, False), (11, True)])
def test_core_spreading_reset_vs_oracle(vpm, orc, kernel, m, statics):
    api = vpm.api
    P = _lattice(vpm, m, statics=(1, m ** 3 // 2) if statics else ())
    n = P.shape[0]
    sg0 = 0.07
    R = P.copy()
    rc, its, res = orc.core_spreading_reset(R, kernel, sg0, itmax=50, tol=1e-3, iterror=True)
    assert rc == 0
    omega = _target(orc, P, kernel)
    static = P[:, 42] != 0
    f = api.Field(n, 0)
    for prec in ("f64", "f32c"):
        f.set_particles(P)
        git, gres = f.core_spreading_reset(sg0, kernel, prec, itmax=50, tol=1e-3, iterror=True)
        G = np.zeros_like(P)
        f.get_particles(G)
        _say("rbf", kernel, n, prec, "iters", git, "oracle", its, "res", gres, res)
        keep = np.r_[0:15, 18:46]          # the zeta field covers statics too (J[1:3] only)
        assert np.array_equal(G[static][:, keep], P[static][:, keep])
        assert np.all(G[~static, 6] == sg0)
        assert np.all(gres <= 1e-3)
        if prec == "f64":
            assert git == its
            scale = np.abs(R[:, 3:6]).max()
            for rows in ((3, 6), (27, 36), (15, 18)):
                e = np.abs(G[:, rows[0]:rows[1]] - R[:, rows[0]:rows[1]]).max()
                ref = np.abs(R[:, rows[0]:rows[1]]).max()
                assert e <= RBF_F64_TOL * ref, (rows, e, ref)
            np.testing.assert_allclose(gres, res, rtol=1e-6)
            assert np.abs(G[:, 3:6] - R[:, 3:6]).max() <= RBF_F64_TOL * scale
        else:
            live = ~static
            d = np.array([np.linalg.norm(G[live, 3 + c] - R[live, 3 + c]) / np.linalg.norm(R[live, 3 + c])
                          for c in range(3)])
            _say("rbf f32c strengths vs oracle", d)
            assert np.all(d <= RBF_F32C_SOL), d
            if not statics:
                r = _resid(orc, G, omega, kernel, static)
                _say("rbf f32c residual vs oracle target", r)
                assert np.all(r <= RBF_F32C_RESID), r
    f.close()


def test_core_spreading_reset_unconverged(vpm, orc):
    """itmax reached with iterror: VPM_ECONVERGE; Gamma holds the last
    iterate, equal to the oracle's after the same two iterations."""
    api = vpm.api
    P = _lattice(vpm, 6)
    R = P.copy()
    rc, its, res = orc.core_spreading_reset(R, "gaussianerf", 0.07, itmax=2, tol=1e-9, iterror=True)
    assert rc == 1 and its == 2
    f = api.Field(P.shape[0], 0)
    f.set_particles(P)
    with pytest.raises(api.VPMError) as ei:
        f.core_spreading_reset(0.07, "gaussianerf", "f64", itmax=2, tol=1e-9, iterror=True)
    assert ei.value.status == -7
    G = np.zeros_like(P)
    f.get_particles(G)
    assert np.abs(G[:, 3:6] - R[:, 3:6]).max() <= RBF_F64_TOL * np.abs(R[:, 3:6]).max()
    # without iterror the same call succeeds
    f.set_particles(P)
    it, r = f.core_spreading_reset(0.07, "gaussianerf", "f64", itmax=2, tol=1e-9, iterror=False)
    assert it == 2 and np.all(r > 1e-9)
    f.close()


def _ring_pf(vpm, P, sg0, precision, beta, itmax):
    api = vpm.api
    pf = api.ParticleField(P.shape[0] + 8, kernel="gaussianerf", formulation=api.rVPM,
                           relaxation=api.pedrizzetti, viscous=api.CoreSpreading(1e-3, sg0, beta=beta, itmax=itmax),
                           precision=precision)
    api.add_particles(pf, P)
    return pf


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_rk3_step_with_reset_vs_oracle(vpm, orc, precision):
    """A rungekutta3 step whose last stage reaches the core-growth limit runs
    the reset there (viscous.jl:189-220): sigma returns to sgm0, t_sgm to 0,
    and the rest of the step (relaxation, U_prev) follows."""
    api = vpm.api
    P = _lattice(vpm, 9, sig=0.7)
    sg0 = float(P[0, 6])
    dt = 0.01
    beta = 1.003         # growth sqrt(2 nu t / sg0^2 + 1): 1.0020 after one step, 1.0041 after two
    pf = _ring_pf(vpm, P, sg0, precision, beta, 50)
    s = orc.make_settings(kernel="gaussianerf", f=0.0, g=0.2, relax=1, rlxf=0.3, viscous=1, nu=1e-3,
                          sgm0=sg0, beta=beta, rbf_itmax=50)
    R = P.copy()
    for step in range(2):
        rc = orc.rk3_step(R, s, dt, relax_now=True)
        assert rc == 0
        api.nextstep(pf, dt, relax=True)
    assert s.t_sgm == 0.0 and pf.viscous.t_sgm == 0.0
    G = pf.particles[:pf.np]
    assert np.all(G[:, 6] == sg0) and np.all(R[:, 6] == sg0)
    if precision == "f64":
        for name, (a, b) in {"X": (0, 3), "Gamma": (3, 6), "M": (27, 33), "U": (9, 12)}.items():
            e = np.abs(G[:, a:b] - R[:, a:b]).max() / np.abs(R[:, a:b]).max()
            _say("rk3+reset f64", name, "%.2e" % e)
            assert e < 1e-9, (name, e)
    else:
        e = np.abs(G[:, 0:3] - R[:, 0:3]).max()
        assert e < 1e-7, e
        eg = np.linalg.norm(G[:, 3:6] - R[:, 3:6]) / np.linalg.norm(R[:, 3:6])
        _say("rk3+reset f32c Gamma rel-norm %.2e" % eg)
        assert eg < 2e-3, eg


def test_rk3_step_reset_deferred(vpm):
    """itmax = 0 keeps the step's VPM_ERESET report (the reset is left to
    the caller) and leaves the grown core sizes in place."""
    api = vpm.api
    P = _lattice(vpm, 4, sig=0.7)
    sg0 = float(P[0, 6])
    pf = _ring_pf(vpm, P, sg0, "f64", 1.0001, 0)
    with pytest.raises(api.VPMError) as ei:
        api.nextstep(pf, 0.01, relax=False)
    assert ei.value.status == -6
    assert np.all(pf.particles[:pf.np, 6] > sg0)


def test_rbf_conjugategradient_alone_matches_reset(vpm, orc):
    """vpm_rbf_conjugategradient on a target set by hand (zeta field in
    M[7:9], sigma = sgm0) is the CG part of vpm_core_spreading_reset: the
    same records, bit for bit (F64)."""
    api = vpm.api
    P = _lattice(vpm, 7, statics=(3,))
    sg0 = 0.07
    f = api.Field(P.shape[0], 0)
    f.set_particles(P)
    it1, r1 = f.core_spreading_reset(sg0, "gaussianerf", "f64", itmax=50)
    A = np.zeros_like(P)
    f.get_particles(A)
    live = P[:, 42] == 0
    f.set_particles(P)
    f.zeta_direct("gaussianerf", "f64")
    B = np.zeros_like(P)
    f.get_particles(B)
    B[live, 33:36] = B[live, 15:18]
    B[live, 6] = sg0
    f.set_particles(B)
    it2, r2 = f.rbf_conjugategradient("gaussianerf", "f64", itmax=50)
    C = np.zeros_like(P)
    f.get_particles(C)
    assert it1 == it2 and np.array_equal(r1, r2)
    assert np.array_equal(A, C)
    f.close()
