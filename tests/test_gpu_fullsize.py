"""The headline workload at its full size: the round-jet column, N = 2^20,
Gaussian-erf (about 10 s of GPU time in all).

Forward: F32C against the FP64 GPU path with the per-particle relative
error of the measurement (vector 2-norm for U, Frobenius for J, floor
1e-6 of the field's largest), within the 1e-5 bound; two F32C evaluations
agree bit for bit; tr J vanishes (the F32C accumulators are traceless).

Adjoint: identities that hold exactly at any size, for the fixed-weight
loss L = sum_i Ubar_i . U_i + sum_i Jbar_i : J_i with the weights set to
the F32C outputs (Ubar = U, Jbar = J, then held fixed), so that L has no
cancellation and an error in any part of the gradient shows at its size:
  * L is linear in Gamma:               sum_j dL/dGamma_j . Gamma_j = L
  * L depends on differences of X:      sum_i dL/dX_i = 0
  * (X, sigma) -> lambda (X, sigma) scales U by lambda^-2, J by lambda^-3:
        sum_i dL/dX_i . X_i + sum_j dL/dsigma_j sigma_j = -2 sum Ubar.U - 3 sum Jbar:J
each checked to ADJ_IDENTITY_TOL of the summation scale of its own sum.  A
dropped term, a wrong sign or a missing source-side sum breaks one of them.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FWD_TOL = 1e-5
ADJ_IDENTITY_TOL = 1e-5


def _say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def jet(vpm):
    P = vpm.workloads.jet_column(1 << 20)
    f = vpm.api.Field(P.shape[0], 0)
    f.set_particles(P)
    yield P, f
    f.close()


def test_jet_forward_f32c_vs_f64(vpm, jet):
    P, f = jet
    n = P.shape[0]
    f.uj_direct("gaussianerf", "f32c")
    A = np.zeros_like(P)
    f.get_particles(A, rows=(9, 24))
    f.uj_direct("gaussianerf", "f32c")
    A2 = np.zeros_like(P)
    f.get_particles(A2, rows=(9, 24))
    assert np.array_equal(A[:, 9:24], A2[:, 9:24])
    f.uj_direct("gaussianerf", "f64")
    B = np.zeros_like(P)
    f.get_particles(B, rows=(9, 24))
    eu = vpm.metrics.err_stats(A[:, 9:12], B[:, 9:12])
    ej = vpm.metrics.err_stats(A[:, 15:24], B[:, 15:24])
    _say("jet n=%d f32c vs f64: U %s J %s" % (n, eu, ej))
    assert eu["max"] < FWD_TOL and ej["max"] < FWD_TOL, (eu, ej)
    tr = A[:, 15] + A[:, 19] + A[:, 23]
    jn = np.linalg.norm(A[:, 15:24], axis=1)
    assert np.all(np.abs(tr) <= 1e-12 * np.maximum(jn, 1e-300)), np.abs(tr / np.maximum(jn, 1e-300)).max()


@pytest.mark.parametrize("with_j", [True, False], ids=["loss_U_and_J", "loss_U"])
def test_jet_adjoint_identities(vpm, jet, with_j):
    """The identities below, for L = sum U.U + sum J:J (the general kernel)
    and for L = sum U.U alone (Jbar = 0: the kernel without Jbar terms)."""
    P, f = jet
    n = P.shape[0]
    f.uj_direct("gaussianerf", "f32c")
    A = np.zeros_like(P)
    f.get_particles(A, rows=(9, 24))
    U, J = A[:, 9:12], A[:, 15:24]
    Ubar, Jbar = U.copy(), (J.copy() if with_j else np.zeros_like(J))
    tu, tj = np.sum(Ubar * U, axis=1), np.sum(Jbar * J, axis=1)
    L = float(np.sum(tu) + np.sum(tj))
    Xb, Gb, Sb = f.uj_adjoint(Ubar, Jbar, "gaussianerf", "f32c")
    X, G, S = P[:, 0:3], P[:, 3:6], P[:, 6]
    # linearity in Gamma
    lhs = np.sum(Gb * G, axis=1)
    scale = np.sum(np.abs(lhs)) + np.sum(np.abs(tu)) + np.sum(np.abs(tj))
    e_lin = abs(float(np.sum(lhs)) - L) / scale
    # translation invariance, per component
    e_tr = np.abs(np.sum(Xb, axis=0)) / np.sum(np.abs(Xb), axis=0)
    # homogeneity in (X, sigma); X relative to the column's centre so that
    # the identity's terms keep their size (translation invariance makes the
    # origin free)
    Xc = X - X.mean(axis=0)
    hx, hs = np.sum(Xb * Xc, axis=1), Sb * S
    rhs = -2.0 * tu - 3.0 * tj
    scale_h = np.sum(np.abs(hx)) + np.sum(np.abs(hs)) + np.sum(np.abs(rhs))
    e_hom = abs(float(np.sum(hx) + np.sum(hs) - np.sum(rhs))) / scale_h
    _say("jet adjoint identities (%s): linearity %.2e translation %s homogeneity %.2e (L %.4e)"
         % ("U and J" if with_j else "U", e_lin, e_tr, e_hom, L))
    assert e_lin < ADJ_IDENTITY_TOL, e_lin
    assert np.all(e_tr < ADJ_IDENTITY_TOL), e_tr
    assert e_hom < ADJ_IDENTITY_TOL, e_hom
