"""GPU parity of the time-step adjoint (vpm_rk3_step_taped / vpm_rk3_step_adjoint).

The step adjoint restores each recorded stage state, pulls the cotangents
back through update_particle_states (FP64, per particle) and through the
stage's U/J evaluation (the adjoint kernel), in reverse stage order.
References: the oracle's FP64 step pullback (pinned to central differences
in test_oracle.py) and central differences of the GPU FP64 rollout.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# FP32 pair arithmetic inside the step: normwise error of each cotangent
# group over all particles (per-particle errors are printed; a particle whose
# gradient cancels has a large relative error at FP32 rounding, see
# test_gpu_adjoint.py)
ROLL_F32C_NORM_TOL = 1e-5


def _say(*a):
    print(*a, flush=True)


def _pfield(vpm, P, **kw):
    api = vpm.api
    pf = api.ParticleField(P.shape[0], **kw)
    api.add_particles(pf, P)
    return pf


def _cloud(vpm, n, seed):
    P = vpm.workloads.random_cloud(n, seed=seed)
    P[:, 6] *= 4.0                  # sigma in [0.08, 0.16]: many near pairs
    P[:, 3:6] *= 50.0               # O(0.05) strengths: visible stretching
    return P


def _errs(vpm, A, R):
    out = {}
    for name, sl in (("X", slice(0, 3)), ("Gamma", slice(3, 6)), ("sigma", slice(6, 7))):
        st = vpm.metrics.err_stats(A[:, sl], R[:, sl])
        st["normwise"] = float(np.linalg.norm(A[:, sl] - R[:, sl]) / max(np.linalg.norm(R[:, sl]), 1e-300))
        out[name] = st
    return out


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_rk3_step_adjoint_vs_oracle(vpm, orc, kernel, precision):
    api = vpm.api
    n = 700
    P = _cloud(vpm, n, seed=41)
    P[[5, 77], 42] = 1.0            # static particles pass their cotangents through
    rng = np.random.default_rng(42)
    bar0 = np.zeros_like(P)
    bar0[:, 0:7] = rng.normal(0, 1, (n, 7))
    s = orc.make_settings(kernel=kernel, f=0.0, g=0.2, relax=0, Uinf=(0.1, 0.0, -0.05))
    dt = 0.02
    R = bar0.copy()
    orc.rk3_step_adjoint(P, s, dt, R)
    pf = _pfield(vpm, P, kernel=kernel, formulation=api.rVPM, relaxation=api.norelaxation,
                 Uinf=lambda t: (0.1, 0.0, -0.05), precision=precision)
    api.pfield_upload(pf)
    tape = api.Tape(pf.device_field())
    api.rungekutta3_taped(pf, dt, tape)
    # the forward half of the taped step is the ordinary rk3 step
    Q = P.copy()
    orc.rk3_step(Q, s, dt, relax_now=False)
    out = np.zeros_like(P)
    pf.device_field().get_particles(out)
    ef = _errs(vpm, out, Q)
    assert max(st["max"] for st in ef.values()) < (1e-10 if precision == "f64" else 1e-5), ef
    G = api.rk3_adjoint(tape, bar0.copy())
    tape.close()
    e = _errs(vpm, G, R)
    _say("rk3 step adjoint", kernel, precision, e)
    assert np.all(G[:, 9:36][P[:, 42] == 0] == 0.0)
    for name, st in e.items():
        if precision == "f64":
            assert st["max"] < 1e-9, (name, st)
        else:
            assert st["normwise"] < ROLL_F32C_NORM_TOL, (name, st)


@pytest.mark.parametrize("precision", ["f64", "f32c"])
def test_kinetic_energy_rollout_gradient_fd(vpm, precision):
    """L = sum |U|^2 after three rk3 steps; the gradient from the taped
    rollout against central differences (h = 1e-6) of the GPU FP64 rollout."""
    api = vpm.api
    n = 400
    P = _cloud(vpm, n, seed=51)
    dt = 0.02

    def pf_of(Q, prec):
        return _pfield(vpm, Q, kernel="gaussianerf", formulation=api.rVPM, relaxation=api.norelaxation,
                       precision=prec)

    def loss(Q):
        pf = pf_of(Q, "f64")
        api.pfield_upload(pf)
        for _ in range(3):
            api.rungekutta3(pf, dt, resident=True)
        f = pf.device_field()
        f.uj_direct("gaussianerf", "f64")
        out = np.zeros((n, 46))
        f.get_particles(out, rows=(9, 12))
        return float(np.sum(out[:, 9:12] ** 2))

    pf = pf_of(P, precision)
    L, bar = api.kinetic_energy_rollout_gradient(pf, dt, 3)
    L64 = loss(P)
    assert abs(L - L64) <= (1e-10 if precision == "f64" else 1e-5) * L64, (L, L64)
    rng = np.random.default_rng(52)
    h = 1e-6
    scale = {r: np.abs(bar[:, r]).max() for r in range(7)}
    worst = 0.0
    for k, r in zip(rng.choice(n, 12, replace=False), rng.integers(0, 7, 12)):
        Pp, Pm = P.copy(), P.copy()
        Pp[k, r] += h
        Pm[k, r] -= h
        fd = (loss(Pp) - loss(Pm)) / (2 * h)
        an = bar[k, r]
        e = abs(fd - an) / max(abs(an), 1e-3 * scale[r])
        worst = max(worst, e)
        assert e < (1e-5 if precision == "f64" else 1e-3), (k, r, fd, an)
    _say("rollout FD", precision, "worst", worst)


def test_taped_step_rejects_sfs_and_empty_tape(vpm):
    api = vpm.api
    P = _cloud(vpm, 64, seed=3)
    pf = _pfield(vpm, P, kernel="gaussianerf", formulation=api.rVPM, relaxation=api.norelaxation,
                 SFS=api.SFS_Cd_twolevel_nobackscatter)
    api.pfield_upload(pf)
    tape = api.Tape(pf.device_field())
    with pytest.raises(api.VPMError):
        api.rk3_adjoint(tape, np.zeros((64, 46)))
    with pytest.raises(api.VPMError):
        api.rungekutta3_taped(pf, 0.01, tape)
    tape.close()


def test_tape_outlives_its_field(vpm):
    """Destroying a field releases its tapes' device memory and detaches
    them: a detached tape's adjoint returns VPM_EINVAL, it can no longer
    record, and destroying it afterwards is safe."""
    api = vpm.api
    P = _cloud(vpm, 200, seed=5)
    f = api.Field(200, 0)
    f.set_particles(P)
    s = api._lib.StepConfig()
    s.uj = api.make_config("gaussianerf", "f32c", True)
    s.g = 0.2
    t1, t2 = api.Tape(f), api.Tape(f)
    api.check(api.lib().vpm_rk3_step_taped(f.handle, api.C.byref(s), 0.01, t1.handle))
    t2.close()                       # a tape closed before its field
    f.close()
    with pytest.raises(api.VPMError) as ei:
        api.rk3_adjoint(t1, np.zeros((200, 46)))
    assert ei.value.status == -1
    t1.close()
