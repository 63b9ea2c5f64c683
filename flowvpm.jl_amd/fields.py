"""Particle record layout, identical to FLOWVPM's ParticleField storage.

The reference keeps particles in a `46 x maxparticles` column-major Float64
matrix (src/FLOWVPM_particlefield.jl:11,134), i.e. one contiguous 46-double
record per particle.  A C-ordered numpy array of shape (n, 46) has the same
bytes, so it can be handed to the C-ABI (or from Julia) without reshuffling.

Indices here are 0-based; the reference's 1-based ranges are cited.
"""
import numpy as np

NFIELDS = 46                       # particlefield.jl:11
X_INDEX = slice(0, 3)              # 1:3
GAMMA_INDEX = slice(3, 6)          # 4:6
SIGMA_INDEX = 6                    # 7
VOL_INDEX = 7                      # 8
CIRCULATION_INDEX = 8              # 9
U_INDEX = slice(9, 12)             # 10:12
VORTICITY_INDEX = slice(12, 15)    # 13:15
J_INDEX = slice(15, 24)            # 16:24
PSE_INDEX = slice(24, 27)          # 25:27
M_INDEX = slice(27, 36)            # 28:36
C_INDEX = slice(36, 39)            # 37:39
SFS_INDEX = slice(39, 42)          # 40:42
STATIC_INDEX = 42                  # 43
U_PREV_INDEX = 43                  # 44


def empty(n):
    """A zeroed particle matrix for n particles."""
    return np.zeros((n, NFIELDS), dtype=np.float64)


def make_particles(X, Gamma, sigma, vol=0.0, circulation=1.0, C=0.0, static=None):
    """Particle matrix from positions, strengths and core sizes, with the
    defaults of add_particle (particlefield.jl:167-194)."""
    X = np.asarray(X, dtype=np.float64).reshape(-1, 3)
    n = X.shape[0]
    P = empty(n)
    P[:, X_INDEX] = X
    P[:, GAMMA_INDEX] = np.asarray(Gamma, dtype=np.float64).reshape(n, 3)
    P[:, SIGMA_INDEX] = np.broadcast_to(np.asarray(sigma, dtype=np.float64), (n,))
    P[:, VOL_INDEX] = vol
    P[:, CIRCULATION_INDEX] = circulation
    P[:, C_INDEX] = C
    if static is not None:
        P[:, STATIC_INDEX] = np.asarray(static, dtype=np.float64)
    return P


def get_W(P):
    """Vorticity from J (particlefield.jl:275-279): (J6-J8, J7-J3, J2-J4)."""
    J = P[:, J_INDEX]
    return np.stack([J[:, 5] - J[:, 7], J[:, 6] - J[:, 2], J[:, 1] - J[:, 3]], axis=1)


def J_matrix(P):
    """J as (n, 3, 3) with [i, a, b] = du_a/dx_b (column-major flat storage)."""
    return P[:, J_INDEX].reshape(-1, 3, 3).transpose(0, 2, 1)
