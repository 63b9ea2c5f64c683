"""GPU parity of the per-particle FP64 schemes, the RK3 step and the adjoint.

The per-particle updates run in FP64 on the device, so with the F64 pair
path a whole rungekutta3 step must agree with the oracle to FP64 rounding
(1e-10 on the state after one step); with F32C pairs the state carries the
FP32c U/J error (bounded here by 1e-5 relative per particle, vector norms).
"""
import math
import os

import numpy as np
import pytest

from conftest import golden

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _state_err(vpm, A, R):
    out = {}
    for name, sl in (("X", slice(0, 3)), ("Gamma", slice(3, 6)), ("sigma", slice(6, 7))):
        out[name] = vpm.metrics.err_stats(A[:, sl], R[:, sl])["max"]
    return out


def _ring_pair(vpm, Nphi=128):
    P = vpm.workloads.leapfrog_cfg2(Nphi=Nphi)
    return P


def _pfield(vpm, P, **kw):
    api = vpm.api
    pf = api.ParticleField(P.shape[0], **kw)
    api.add_particles(pf, P)
    return pf


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_rk3_nosfs_relax_vs_oracle(vpm, orc, precision):
    api = vpm.api
    P = vpm.workloads.ring_cfg0()[::4].copy()          # 2,000 particles
    P[[3, 50], 42] = 1.0                               # two static particles
    pf = _pfield(vpm, P, kernel="gaussianerf", formulation=api.rVPM, relaxation=api.pedrizzetti,
                 precision=precision)
    s = orc.make_settings(kernel="gaussianerf", f=0.0, g=0.2, relax=1, rlxf=0.3)
    R = P.copy()
    dt = 0.01
    for step in range(2):
        orc.rk3_step(R, s, dt, relax_now=True)
        api.nextstep(pf, dt, relax=True)
    e = _state_err(vpm, pf.particles[:pf.np], R)
    _say("rk3 nosfs", precision, e)
    tol = 1e-10 if precision == "f64" else 1e-5
    assert max(e.values()) < tol, e
    # nextstep's U_prev = |U|
    np.testing.assert_allclose(pf.particles[:pf.np, 43], np.linalg.norm(pf.particles[:pf.np, 9:12], axis=1))


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_rk3_dynamic_sfs_core_spreading_vs_oracle(vpm, orc, precision):
    """Config-2 physics (rVPM, SFS_Cd_twolevel_nobackscatter, CoreSpreading,
    pedrizzetti) on a reduced leapfrog pair."""
    api = vpm.api
    P = _ring_pair(vpm, Nphi=64)                        # 8,192 particles
    sg0 = float(P[0, 6])
    pf = _pfield(vpm, P, kernel="gaussianerf", formulation=api.rVPM, relaxation=api.pedrizzetti,
                 SFS=api.SFS_Cd_twolevel_nobackscatter, viscous=api.CoreSpreading(1e-4, sg0),
                 precision=precision)
    s = orc.make_settings(kernel="gaussianerf", f=0.0, g=0.2, relax=1, rlxf=0.3, sfs=1, sfs_alpha=0.999,
                          sfs_rlxf=0.005, sfs_minC=0.0, sfs_maxC=1.0, sfs_force_positive=True,
                          sfs_clip=True, viscous=1, nu=1e-4, sgm0=sg0, beta=1.5)
    R = P.copy()
    dt = 0.01
    for step in range(2):
        rc = orc.rk3_step(R, s, dt, relax_now=True)
        assert rc == 0
        api.nextstep(pf, dt, relax=True)
    e = _state_err(vpm, pf.particles[:pf.np], R)
    C = vpm.metrics.err_stats(pf.particles[:pf.np, 36:39], R[:, 36:39])["max"]
    _say("rk3 dynsfs+cs", precision, e, "C", C)
    assert pf.viscous.t_sgm == pytest.approx(s.t_sgm)
    if precision == "f64":
        assert max(e.values()) < 1e-10, e
        assert C < 1e-8
    else:
        # The pseudo-3-level procedure takes M = S(alpha sigma) - S(sigma)
        # with alpha = 0.999 (subfilterscale.jl:571-585).  The F32C step
        # evaluates these differences directly (sfs_delta.hip) instead of
        # subtracting two U/J passes 1e-3 apart, which would amplify the
        # FP32 U/J error by ~1/(1 - alpha) (measured 2e-4 in Gamma that way).
        assert e["X"] < 1e-7 and e["sigma"] < 1e-7, e
        assert e["Gamma"] < 2e-6, e


def test_per_particle_kernels_vs_oracle(vpm, orc):
    """update_particle_states, both relaxations and core spreading, one call each."""
    api = vpm.api
    rng = np.random.default_rng(0)
    n = 777
    P = vpm.workloads.random_cloud(n, seed=3)
    P[:, 9:12] = rng.normal(size=(n, 3))
    P[:, 15:24] = rng.normal(size=(n, 9))
    P[:, 27:36] = rng.normal(size=(n, 9))
    P[:, 36] = rng.uniform(0, 1, n)
    P[:, 39:42] = rng.normal(size=(n, 3))
    P[5, 3:6] = 0.0                                    # |Gamma| = 0 branch
    P[9, 42] = 1.0
    f = api.Field(n, 0)
    for transposed in (True, False):
        for (a, b) in ((0.0, 1 / 3), (-5 / 9, 15 / 16)):
            R = P.copy()
            orc.update_particle_states(R, a, b, 0.01, [0.1, 0.2, 0.3], 0.1, 0.2, 0.06, transposed)
            f.set_particles(P)
            api.check(api.lib().vpm_update_particle_states(f.handle, a, b, 0.01,
                                                           np.array([0.1, 0.2, 0.3]).ctypes.data_as(api._lib._dptr),
                                                           0.1, 0.2, 0.06, int(transposed)))
            A = np.zeros_like(P)
            f.get_particles(A)
            np.testing.assert_array_equal(A, R)
    for corrected in (0, 1):
        R = P.copy()
        orc.relax(R, 0.3, corrected)
        f.set_particles(P)
        api.check(api.lib().vpm_relax(f.handle, 0.3, corrected))
        A = np.zeros_like(P)
        f.get_particles(A)
        np.testing.assert_allclose(A, R, rtol=1e-14, atol=0)
    R = P.copy()
    orc.core_spreading_rk3(R, 1e-4, 0.03, 1.5, 0.0, -5 / 9, 15 / 16, 0.01)
    f.set_particles(P)
    api.check(api.lib().vpm_core_spreading(f.handle, 1e-4, -5 / 9, 15 / 16, 0.01))
    A = np.zeros_like(P)
    f.get_particles(A)
    np.testing.assert_array_equal(A, R)
    f.close()


def test_single_ring_translation(vpm, orc):
    """test/runtests_singlevortexring.jl:22-23 with UJ_direct and RK3: nc=0
    ring of 100 particles, Winckelmans, cVPM, pedrizzetti every step, 50
    steps over two ring radii; the |Gamma|-weighted centroid speed must be
    within 2% of Uring = Gamma/(4 pi R) (ln(8R/Rcross) - 1/2) (:143)."""
    api = vpm.api
    P = vpm.workloads.ring_test()
    Uana = vpm.workloads.Uring(1.0, 1.0, 0.15, 0.5)
    nsteps = 50
    dt = (2.0 / Uana) / nsteps
    for precision in ("f64", "f32c"):
        pf = _pfield(vpm, P, kernel="winckelmans", formulation=api.cVPM, relaxation=api.pedrizzetti,
                     precision=precision)
        api.run_vpm(pf, dt, nsteps)
        Z = vpm.workloads.ring_centroid(pf.particles[:pf.np])
        U = np.linalg.norm(Z) / pf.t
        err = (U - Uana) / Uana
        _say("ring", precision, "U_vpm %.5f U_ana %.5f err %.4f" % (U, Uana, err))
        assert abs(err) < 0.02
    # and the GPU rollout matches the oracle's FP64 rollout
    R = P.copy()
    s = orc.make_settings(kernel="winckelmans", f=0.0, g=0.0, relax=1, rlxf=0.3)
    for _ in range(nsteps):
        orc.rk3_step(R, s, dt, relax_now=True)
    pf = _pfield(vpm, P, kernel="winckelmans", formulation=api.cVPM, relaxation=api.pedrizzetti,
                 precision="f64")
    api.run_vpm(pf, dt, nsteps)
    e = _state_err(vpm, pf.particles[:pf.np], R)
    assert max(e.values()) < 1e-9, e


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans", "gaussian", "singular"])
def test_adjoint_vs_oracle(vpm, orc, kernel):
    P = vpm.workloads.random_cloud(1500, seed=21)
    w, W = vpm.workloads.adjoint_weights(1500)
    Xr, Gr, Sr = orc.uj_adjoint(P, kernel, w, W)
    Xb, Gb, Sb = vpm.uj_adjoint(P, w, W, kernel=kernel, precision="f64")
    for name, a, r in (("X", Xb, Xr), ("Gamma", Gb, Gr), ("sigma", Sb[:, None], Sr[:, None])):
        st = vpm.metrics.err_stats(a, r)
        _say("adjoint", kernel, name, st)
        assert st["max"] < 1e-9, (name, st)


def test_adjoint_golden(vpm):
    g = golden("adjoint.json")
    P = vpm.fields.make_particles(g["X"], g["Gamma"], g["sigma"])
    for kernel in ("gaussianerf", "winckelmans", "gaussian", "singular"):
        Xb, Gb, Sb = vpm.uj_adjoint(P, np.array(g["Ubar"]), np.array(g["Jbar"]), kernel=kernel,
                                    precision="f64")
        ref = g[kernel]
        np.testing.assert_allclose(Xb, ref["Xbar"], rtol=1e-9, atol=1e-12 * np.abs(ref["Xbar"]).max())
        np.testing.assert_allclose(Gb, ref["Gbar"], rtol=1e-9, atol=1e-12 * np.abs(ref["Gbar"]).max())
        np.testing.assert_allclose(Sb, ref["Sbar"], rtol=1e-9, atol=1e-12 * max(np.abs(ref["Sbar"]).max(), 1e-30))


@pytest.mark.parametrize("resident", [True, False])
def test_run_vpm_static_particles_and_relax_cadence_vs_oracle(vpm, orc, resident):
    """run_vpm! (utils.jl:86-119): static particles appended before each
    step and removed after it (remove_particle from the end, :106-111),
    relaxation every nsteps_relax = 2 steps from step 1, and a runtime
    function that sees the field after each step and ends the run.  The
    oracle replays the same sequence; F64 state to 1e-10."""
    api = vpm.api
    P = vpm.workloads.ring_cfg0()[::8].copy()          # 1,000 particles
    n = P.shape[0]
    rng = np.random.default_rng(40)
    rl = api.Relaxation(1, 2, 0.3)
    pf = api.ParticleField(n + 8, kernel="gaussianerf", formulation=api.rVPM, relaxation=rl, precision="f64")
    api.add_particles(pf, P)
    dt = 0.01

    def statics(k):
        X = P[rng.integers(0, n, 3), 0:3] + 0.05
        S = vpm.fields.make_particles(X, np.full((3, 3), 1e-3 * (k + 1)), np.full(3, float(P[0, 6])))
        S[:, 42] = 1.0
        return S

    added = [statics(k) for k in range(4)]
    seen = []

    def static_fn(pfield, t, dt_):
        api.add_particles(pfield, added[pfield.nt])
        return None

    def runtime(pfield, t, dt_):
        seen.append((pfield.nt, pfield.np))
        return pfield.nt == 3

    api.run_vpm(pf, dt, 10, runtime_function=runtime, static_particles_function=static_fn, resident=resident)
    assert seen == [(0, n), (1, n), (2, n), (3, n)]
    s = orc.make_settings(kernel="gaussianerf", f=0.0, g=0.2, relax=1, rlxf=0.3)
    R = P.copy()
    for i in range(1, 4):
        Q = np.ascontiguousarray(np.vstack([R, added[i - 1]]))
        orc.rk3_step(Q, s, dt, relax_now=(i % 2 == 0))
        R = Q[:n].copy()
    e = _state_err(vpm, pf.particles[:pf.np], R)
    _say("run_vpm statics resident=%s" % resident, e)
    assert max(e.values()) < 1e-10, e
    assert np.all(pf.particles[n:n + 3, 9:27] == 0.0)   # removed slots reset


@pytest.mark.parametrize("resident", [True, False])
def test_run_vpm_runtime_function_edits_field_vs_oracle(vpm, orc, resident):
    """A runtime_function that adds particles, removes one and scales a
    strength (the round jet's inflow does the first, examples/roundjet/
    roundjet_simulation.jl:229) acts on the live field, and the next step
    sees its changes on the device too.  The oracle replays the same
    edits; F64 state to 1e-10."""
    api = vpm.api
    P = vpm.workloads.ring_cfg0()[::10].copy()          # 800 particles
    n = P.shape[0]
    rng = np.random.default_rng(41)
    pf = api.ParticleField(n + 16, kernel="gaussianerf", formulation=api.rVPM, precision="f64")
    api.add_particles(pf, P)
    dt = 0.01
    new = [vpm.fields.make_particles(P[rng.integers(0, n, 2), 0:3] + 0.03, np.full((2, 3), 2e-3 * (k + 1)),
                                     np.full(2, float(P[0, 6]))) for k in range(3)]

    def edit(F, k):
        """the edits of step k on a record matrix F of m rows; returns m"""
        m = F.shape[0]
        if k == 1:
            F = np.vstack([F, new[0]])
        elif k == 2:
            F = np.vstack([F, new[1]])
            F[5] = F[-1]              # remove_particle(5): the last record moves in
            F = F[:-1]
            F[7, 3:6] *= 1.5
        elif k == 3:
            F = np.vstack([F, new[2]])
            F[1, 3:6] *= 0.5
        return F

    def runtime(pfield, t, dt_):
        k = pfield.nt
        if k == 1:
            api.add_particles(pfield, new[0])
        elif k == 2:
            api.add_particles(pfield, new[1])
            api.remove_particle(pfield, 5)
            pfield.particles[7, 3:6] *= 1.5
        elif k == 3:
            api.add_particles(pfield, new[2])
            pfield.particles[1, 3:6] *= 0.5
        return k == 4

    api.run_vpm(pf, dt, 10, runtime_function=runtime, resident=resident)
    assert pf.nt == 4 and pf.np == n + 5           # +2, +2 - 1, +2
    s = orc.make_settings(kernel="gaussianerf", f=0.0, g=0.2)
    R = P.copy()
    for k in range(1, 5):
        orc.rk3_step(R, s, dt)
        R = np.ascontiguousarray(edit(R, k))
    e = _state_err(vpm, pf.particles[:pf.np], R)
    _say("run_vpm runtime edits resident=%s" % resident, e)
    assert max(e.values()) < 1e-10, e


def test_run_vpm_saves_and_restarts(vpm, tmp_path):
    """run_vpm with save_path on a device-resident field (FLOWVPM_utils.jl:
    124-129): each saved file holds the field as it was at that step, and a
    restart with read! (load_time, overwrite) from step 2 takes the next step
    to the same state, bit for bit, as the run that went on (a step reads
    only X, Gamma, sigma, vol and static, all in the file).  Skipped without
    libhdf5."""
    from flowvpm_amd import h5io
    if h5io.hdf5_library_path() is None:
        pytest.skip("no libhdf5 on this machine")
    api = vpm.api
    P = vpm.workloads.ring_cfg0()[::8].copy()
    n = P.shape[0]
    out = str(tmp_path / "run")
    kw = dict(kernel="gaussianerf", formulation=api.rVPM, relaxation=api.pedrizzetti, precision="f32c")
    pf = _pfield(vpm, P, **kw)
    snaps = {}

    def runtime(pfield, t, dt_):
        snaps[pfield.nt] = pfield.particles[:pfield.np].copy()
        return False

    api.run_vpm(pf, 0.01, 3, runtime_function=runtime, save_path=out, resident=True)
    assert sorted(os.listdir(out)) == sorted("pfield.%d.%s" % (k, e) for k in range(4) for e in ("h5", "xmf"))
    for k in range(4):
        q = api.ParticleField(n)
        h5io.read_into(q, "pfield.%d" % k, path=out)
        assert q.nt == k and abs(q.t - 0.01 * k) < 1e-15
        cols = [0, 1, 2, 3, 4, 5, 6, 7, 8, 42]
        assert np.array_equal(q.particles[:, cols], snaps[k][:, cols]), k
    r = api.ParticleField(n, **kw)
    h5io.read_into(r, "pfield.2", path=out, overwrite=True, load_time=True)
    api.nextstep(r, 0.01, relax=True)
    assert r.nt == 3
    _say("restart from step 2 vs the run: max |dX| %.1e" % np.abs(r.particles[:n, 0:3] - snaps[3][:, 0:3]).max())
    assert np.array_equal(r.particles[:n, 0:9], snaps[3][:, 0:9])
