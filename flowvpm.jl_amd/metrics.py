"""Per-particle relative error used for every parity statement.

e_i = ||F_i - R_i||_2 / max(||R_i||_2, floor * max_k ||R_k||_2), with the
vector 2-norm for U, dGamma/dt and SFS and the Frobenius norm for J.
Component-wise relative error is not used: tr J = 0 and many components of U
and J vanish by symmetry.
"""
import numpy as np


def rel_err(F, R, floor=1e-6):
    F = np.asarray(F, dtype=np.float64).reshape(len(F), -1)
    R = np.asarray(R, dtype=np.float64).reshape(len(R), -1)
    num = np.linalg.norm(F - R, axis=1)
    den = np.linalg.norm(R, axis=1)
    den = np.maximum(den, floor * (den.max() if den.size else 0.0))
    with np.errstate(invalid="ignore", divide="ignore"):
        e = np.where(den > 0, num / np.where(den > 0, den, 1.0), np.where(num > 0, np.inf, 0.0))
    return e


def err_stats(F, R, floor=1e-6):
    e = rel_err(F, R, floor)
    if e.size == 0:
        return {"max": 0.0, "rms": 0.0}
    return {"max": float(e.max()), "rms": float(np.sqrt(np.mean(e * e)))}


def stretching(J, G, transposed=True):
    """(Gamma . grad^T) U per particle from the J rows (flat column-major,
    J(r, c) = J[3 c + r]): transposed J^T Gamma, classic J Gamma
    (src/FLOWVPM_timeintegration.jl:489-499).  Host-side, for error reports."""
    J = np.asarray(J, dtype=np.float64).reshape(-1, 3, 3)   # J[i][c][r]
    G = np.asarray(G, dtype=np.float64).reshape(-1, 3)
    if transposed:
        return np.einsum("icr,ir->ic", J, G)
    return np.einsum("icr,ic->ir", J, G)
