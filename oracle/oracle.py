"""ctypes front end of the FP64 CPU oracle (oracle/vpm_oracle.c).

TEST INFRASTRUCTURE ONLY.  This is the checker: tests/, the smoke() check in
__graft_entry__.py and the cpu_baseline leg of bench.py load it; the product
package never does.

Particle arrays are numpy float64 arrays of shape (n, 46), C order, which is
byte-for-byte the reference's 46 x n column-major particle matrix
(src/FLOWVPM_particlefield.jl:11,134).
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

SINGULAR, GAUSSIAN, GAUSSIANERF, WINCKELMANS = 0, 1, 2, 3
KERNELS = {"singular": SINGULAR, "gaussian": GAUSSIAN,
           "gaussianerf": GAUSSIANERF, "winckelmans": WINCKELMANS}
NFIELDS = 46

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class Settings(C.Structure):
    _fields_ = [("kernel", C.c_int32), ("transposed", C.c_int32),
                ("f", C.c_double), ("g", C.c_double), ("Uinf", C.c_double * 3),
                ("sfs", C.c_int32), ("sfs_force_positive", C.c_int32), ("sfs_clip", C.c_int32),
                ("relax", C.c_int32),
                ("sfs_alpha", C.c_double), ("sfs_rlxf", C.c_double),
                ("sfs_minC", C.c_double), ("sfs_maxC", C.c_double),
                ("rlxf", C.c_double),
                ("viscous", C.c_int32), ("nthreads", C.c_int32),
                ("nu", C.c_double), ("sgm0", C.c_double), ("beta", C.c_double),
                ("t_sgm", C.c_double),
                ("rbf_itmax", C.c_int32), ("rbf_iterror", C.c_int32), ("rbf_tol", C.c_double)]


def build():
    """Compile liboracle.so with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.vpmo_erf64.restype = C.c_double
        L.vpmo_erf64.argtypes = [C.c_double]
        L.vpmo_zeta.restype = C.c_double
        L.vpmo_zeta.argtypes = [C.c_int, C.c_double]
        L.vpmo_g_dgdr.argtypes = [C.c_int, C.c_double, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.vpmo_g_derivs.argtypes = [C.c_int, C.c_double] + [C.POINTER(C.c_double)] * 3
        L.vpmo_direct2.argtypes = [_dp, C.c_int64, C.c_int64, C.c_int64,
                                   _dp, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_int]
        L.vpmo_direct_ranges.argtypes = [_dp, C.c_int64, C.c_int, _ip, _ip, _ip, _ip, C.c_int64]
        L.vpmo_reset_particles.argtypes = [_dp, C.c_int64]
        L.vpmo_reset_particles_sfs.argtypes = [_dp, C.c_int64]
        L.vpmo_uj_direct.argtypes = [_dp, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int]
        L.vpmo_estr_direct.argtypes = [_dp, C.c_int64, C.c_int, C.c_int, C.c_int]
        L.vpmo_estr_direct_targets.argtypes = [_dp, C.c_int64, C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int]
        L.vpmo_update_particle_states.argtypes = [_dp, C.c_int64, C.c_double, C.c_double, C.c_double,
                                                  _dp, C.c_double, C.c_double, C.c_double, C.c_int]
        L.vpmo_relax.argtypes = [_dp, C.c_int64, C.c_double, C.c_int]
        L.vpmo_update_particle_states_adjoint.argtypes = [_dp, _dp, C.c_int64, C.c_double, C.c_double,
                                                          C.c_double, C.c_double, C.c_double, C.c_int]
        L.vpmo_uj_adjoint_scale.argtypes = [_dp, C.c_int64, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.vpmo_estr_ranges.argtypes = [_dp, C.c_int64, C.c_int, C.c_int, _ip, _ip, _ip, _ip, C.c_int64]
        L.vpmo_estr_scale.argtypes = [_dp, C.c_int64, C.c_int, C.c_int, _dp, C.c_int]
        L.vpmo_estr_scale_ranges.argtypes = [_dp, C.c_int64, C.c_int, C.c_int, _ip, _ip, _ip, _ip, C.c_int64, _dp]
        L.vpmo_u_scale_ranges.argtypes = [_dp, C.c_int64, C.c_int, _ip, _ip, _ip, _ip, C.c_int64, _dp]
        L.vpmo_core_spreading_reset.restype = C.c_int
        L.vpmo_core_spreading_reset.argtypes = [_dp, C.c_int64, C.c_int, C.c_double, C.c_int, C.c_double,
                                                C.c_int, C.c_int, C.POINTER(C.c_int), _dp]
        L.vpmo_zeta_direct.argtypes = [_dp, C.c_int64, C.c_int, C.c_int]
        L.vpmo_update_particle_states_pullback_rrules.argtypes = [_dp, _dp, C.c_int64, C.c_double, C.c_double,
                                                                  C.c_double, C.c_double, C.c_double, C.c_double,
                                                                  C.c_int]
        L.vpmo_uj_adjoint_rrules.argtypes = [_dp, C.c_int64, C.c_int, _dp, _dp, _dp, _dp, _dp]
        L.vpmo_zeta_ranges.argtypes = [_dp, C.c_int64, C.c_int, _ip, _ip, _ip, _ip, C.c_int64]
        L.vpmo_rk3_step_adjoint.restype = C.c_int
        L.vpmo_rk3_step_adjoint.argtypes = [_dp, C.c_int64, C.c_void_p, C.c_double, _dp]
        L.vpmo_core_spreading_rk3.restype = C.c_int
        L.vpmo_core_spreading_rk3.argtypes = [_dp, C.c_int64, C.c_double, C.c_double, C.c_double,
                                              C.POINTER(C.c_double), C.c_double, C.c_double, C.c_double]
        L.vpmo_sfs_before.restype = C.c_int
        L.vpmo_sfs_before.argtypes = [_dp, C.c_int64, C.c_int, C.c_int, C.c_double,
                                      C.c_double, C.c_double, C.c_int]
        L.vpmo_sfs_after.restype = C.c_int
        L.vpmo_sfs_after.argtypes = [_dp, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double,
                                     C.c_double, C.c_double, C.c_int, C.c_int]
        L.vpmo_rk3_step.restype = C.c_int
        L.vpmo_rk3_step.argtypes = [_dp, C.c_int64, C.POINTER(Settings), C.c_double, C.c_int]
        L.vpmo_uj_adjoint.argtypes = [_dp, C.c_int64, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_int]
        L.vpmo_loss_local.restype = C.c_double
        L.vpmo_loss_local.argtypes = [_dp, C.c_int64, C.c_int, _dp, _dp, C.c_int64]
        _lib = L
    return _lib


def _kid(kernel):
    return KERNELS[kernel] if isinstance(kernel, str) else int(kernel)


def _check(P):
    assert P.dtype == np.float64 and P.flags.c_contiguous and P.shape[1] == NFIELDS
    return P


def nthreads_default():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


# ---- scalar functions ----
def erf64(x):
    return lib().vpmo_erf64(float(x))


def zeta(kernel, q):
    return lib().vpmo_zeta(_kid(kernel), float(q))


def g_dgdr(kernel, q):
    g, dg = C.c_double(), C.c_double()
    lib().vpmo_g_dgdr(_kid(kernel), float(q), C.byref(g), C.byref(dg))
    return g.value, dg.value


def g_derivs(kernel, q):
    g, dg, d2g = C.c_double(), C.c_double(), C.c_double()
    lib().vpmo_g_derivs(_kid(kernel), float(q), C.byref(g), C.byref(dg), C.byref(d2g))
    return g.value, dg.value, d2g.value


# ---- U/J ----
def direct(P, kernel, t_range=None, s_range=None, nthreads=None, source=None):
    """fmm.direct! over 0-based half-open ranges; accumulates into P's U, J."""
    _check(P)
    S = P if source is None else _check(source)
    t0, t1 = t_range if t_range is not None else (0, P.shape[0])
    s0, s1 = s_range if s_range is not None else (0, S.shape[0])
    lib().vpmo_direct2(P, NFIELDS, t0, t1, S, NFIELDS, s0, s1, _kid(kernel),
                       nthreads or nthreads_default())


def direct_ranges(P, kernel, tb, te, sb, se):
    _check(P)
    a = [np.ascontiguousarray(x, dtype=np.int64) for x in (tb, te, sb, se)]
    lib().vpmo_direct_ranges(P, NFIELDS, _kid(kernel), a[0], a[1], a[2], a[3], len(a[0]))


def reset_particles(P, np_=None):
    lib().vpmo_reset_particles(_check(P), np_ if np_ is not None else P.shape[0])


def reset_particles_sfs(P, np_=None):
    lib().vpmo_reset_particles_sfs(_check(P), np_ if np_ is not None else P.shape[0])


def uj_direct(P, kernel, reset=True, reset_sfs=False, sfs=False, transposed=True,
              np_=None, nthreads=None):
    lib().vpmo_uj_direct(_check(P), np_ if np_ is not None else P.shape[0], _kid(kernel),
                         int(reset), int(reset_sfs), int(sfs), int(transposed),
                         nthreads or nthreads_default())


def estr_ranges(P, kernel, tb, te, sb, se, transposed=True):
    """Estr_fmm! over a direct list, in place on the SFS rows."""
    a = [np.ascontiguousarray(x, dtype=np.int64) for x in (tb, te, sb, se)]
    lib().vpmo_estr_ranges(_check(P), NFIELDS, _kid(kernel), int(transposed), *a, len(a[0]))


def estr_scale(P, kernel, transposed=True, nthreads=None):
    """Summation scale of E_str per particle, s_i = sum zeta |(J_i - J_j)^T
    Gamma_j| over Estr_direct!'s pairs (static skips included)."""
    s = np.zeros(P.shape[0])
    lib().vpmo_estr_scale(_check(P), P.shape[0], _kid(kernel), int(transposed), s, nthreads or nthreads_default())
    return s


def estr_scale_ranges(P, kernel, tb, te, sb, se, transposed=True):
    """The E_str summation scale over the pairs of a direct list."""
    a = [np.ascontiguousarray(x, dtype=np.int64) for x in (tb, te, sb, se)]
    s = np.zeros(P.shape[0])
    lib().vpmo_estr_scale_ranges(_check(P), NFIELDS, _kid(kernel), int(transposed), *a, len(a[0]), s)
    return s


def u_scale_ranges(P, kernel, tb, te, sb, se):
    """The U summation scale over the pairs of a direct list: sum over each
    target's pairs of |U term| (the rounding bound of any evaluation)."""
    a = [np.ascontiguousarray(x, dtype=np.int64) for x in (tb, te, sb, se)]
    s = np.zeros(P.shape[0])
    lib().vpmo_u_scale_ranges(_check(P), NFIELDS, _kid(kernel), *a, len(a[0]), s)
    return s


def core_spreading_reset(P, kernel, sgm0, itmax=15, tol=1e-3, iterror=True, nthreads=None):
    """Core-spreading reset + RBF conjugate gradient, in place; returns
    (status, iterations, residuals)."""
    it = C.c_int(0)
    res = np.zeros(3)
    rc = lib().vpmo_core_spreading_reset(_check(P), P.shape[0], _kid(kernel), sgm0, int(itmax), tol,
                                         int(iterror), nthreads or nthreads_default(), C.byref(it), res)
    return rc, it.value, res


def zeta_direct(P, kernel, nthreads=None):
    """zeta_direct (viscous.jl:486-515), in place on J rows 15:18."""
    lib().vpmo_zeta_direct(_check(P), P.shape[0], _kid(kernel), nthreads or nthreads_default())


def zeta_ranges(P, kernel, tb, te, sb, se):
    """zeta_fmm's pair loop over a direct list, accumulated onto J rows 15:18."""
    a = [np.ascontiguousarray(x, dtype=np.int64) for x in (tb, te, sb, se)]
    lib().vpmo_zeta_ranges(_check(P), NFIELDS, _kid(kernel), *a, len(a[0]))


def estr_direct(P, kernel, transposed=True, np_=None, nthreads=None, t_range=None):
    n = np_ if np_ is not None else P.shape[0]
    if t_range is None:
        lib().vpmo_estr_direct(_check(P), n, _kid(kernel), int(transposed), nthreads or nthreads_default())
    else:
        lib().vpmo_estr_direct_targets(_check(P), n, int(t_range[0]), int(t_range[1]), _kid(kernel),
                                       int(transposed), nthreads or nthreads_default())


# ---- per-particle schemes ----
def update_particle_states(P, a, b, dt, Uinf, f, g, zeta0, transposed=True):
    lib().vpmo_update_particle_states(_check(P), P.shape[0], a, b, dt,
                                      np.ascontiguousarray(Uinf, dtype=np.float64),
                                      f, g, zeta0, int(transposed))


def relax(P, rlxf, corrected=False):
    lib().vpmo_relax(_check(P), P.shape[0], rlxf, int(corrected))


def core_spreading_rk3(P, nu, sgm0, beta, t_sgm, a, b, dt):
    t = C.c_double(t_sgm)
    rc = lib().vpmo_core_spreading_rk3(_check(P), P.shape[0], nu, sgm0, beta, C.byref(t), a, b, dt)
    return rc, t.value


def sfs_before(P, kernel, transposed, alpha, minC, maxC, nthreads=None):
    return lib().vpmo_sfs_before(_check(P), P.shape[0], _kid(kernel), int(transposed), alpha,
                                 minC, maxC, nthreads or nthreads_default())


def sfs_after(P, kernel, transposed, alpha, rlxf, minC, maxC, force_positive=True, clip=True):
    return lib().vpmo_sfs_after(_check(P), P.shape[0], _kid(kernel), int(transposed), alpha, rlxf,
                                minC, maxC, int(force_positive), int(clip))


def make_settings(kernel="gaussianerf", transposed=True, f=0.0, g=0.2, Uinf=(0, 0, 0),
                  sfs=0, sfs_alpha=0.999, sfs_rlxf=0.005, sfs_minC=0.0, sfs_maxC=1.0,
                  sfs_force_positive=True, sfs_clip=True, relax=1, rlxf=0.3,
                  viscous=0, nu=0.0, sgm0=1.0, beta=1.5, t_sgm=0.0, nthreads=None,
                  rbf_itmax=15, rbf_tol=1e-3, rbf_iterror=True):
    s = Settings()
    s.kernel, s.transposed, s.f, s.g = _kid(kernel), int(transposed), f, g
    s.Uinf[:] = list(Uinf)
    s.sfs, s.sfs_force_positive, s.sfs_clip = int(sfs), int(sfs_force_positive), int(sfs_clip)
    s.relax, s.sfs_alpha, s.sfs_rlxf, s.sfs_minC, s.sfs_maxC = int(relax), sfs_alpha, sfs_rlxf, sfs_minC, sfs_maxC
    s.rlxf, s.viscous, s.nthreads = rlxf, int(viscous), int(nthreads or nthreads_default())
    s.nu, s.sgm0, s.beta, s.t_sgm = nu, sgm0, beta, t_sgm
    s.rbf_itmax, s.rbf_tol, s.rbf_iterror = int(rbf_itmax), float(rbf_tol), int(bool(rbf_iterror))
    return s


def rk3_step(P, settings, dt, relax_now=True):
    return lib().vpmo_rk3_step(_check(P), P.shape[0], C.byref(settings), dt, int(relax_now))


# ---- adjoint ----
def uj_adjoint(P, kernel, Ubar, Jbar, nthreads=None):
    n = P.shape[0]
    Xb, Gb, Sb = np.zeros(3 * n), np.zeros(3 * n), np.zeros(n)
    lib().vpmo_uj_adjoint(_check(P), n, _kid(kernel),
                          np.ascontiguousarray(Ubar, dtype=np.float64).reshape(-1),
                          np.ascontiguousarray(Jbar, dtype=np.float64).reshape(-1),
                          Xb, Gb, Sb, nthreads or nthreads_default())
    return Xb.reshape(n, 3), Gb.reshape(n, 3), Sb


def update_particle_states_pullback_rrules(P, B, a, b, dt, f, g, zeta0, full=False):
    """The reference's state_pullback (rrules.jl:436-547) restated literally,
    in place on the cotangent records B (layout of update_particle_states_adjoint)."""
    lib().vpmo_update_particle_states_pullback_rrules(_check(P), _check(B), P.shape[0], a, b, dt, f, g, zeta0,
                                                      int(bool(full)))


def uj_adjoint_rrules(P, kernel, Ubar, Jbar):
    """The reference's UJS_pullback (rrules.jl:183-326) restated literally:
    (dL/dX, dL/dGamma, dL/dsigma) for a field acting on itself."""
    n = P.shape[0]
    Xb, Gb, Sb = np.zeros(3 * n), np.zeros(3 * n), np.zeros(n)
    lib().vpmo_uj_adjoint_rrules(_check(P), n, _kid(kernel),
                                 np.ascontiguousarray(Ubar, dtype=np.float64).reshape(-1),
                                 np.ascontiguousarray(Jbar, dtype=np.float64).reshape(-1), Xb, Gb, Sb)
    return Xb.reshape(n, 3), Gb.reshape(n, 3), Sb


def uj_adjoint_scale(P, kernel, Ubar, Jbar, nthreads=None):
    """Per-output sums of |pair contribution| of uj_adjoint (the rounding
    error scale of the adjoint sums)."""
    n = P.shape[0]
    Xm, Gm, Sm = np.zeros(3 * n), np.zeros(3 * n), np.zeros(n)
    lib().vpmo_uj_adjoint_scale(_check(P), n, _kid(kernel),
                                np.ascontiguousarray(Ubar, dtype=np.float64).reshape(-1),
                                np.ascontiguousarray(Jbar, dtype=np.float64).reshape(-1),
                                Xm, Gm, Sm, nthreads or nthreads_default())
    return Xm.reshape(n, 3), Gm.reshape(n, 3), Sm


def loss_local(P, kernel, Ubar, Jbar, k):
    return lib().vpmo_loss_local(_check(P), P.shape[0], _kid(kernel),
                                 np.ascontiguousarray(Ubar, dtype=np.float64).reshape(-1),
                                 np.ascontiguousarray(Jbar, dtype=np.float64).reshape(-1), int(k))


def update_particle_states_adjoint(P, B, a, b, dt, f, g, transposed=True):
    """In place on B (cotangent records, shape (n, 46)); P is the state before
    the update."""
    lib().vpmo_update_particle_states_adjoint(_check(P), _check(B), P.shape[0], a, b, dt, f, g,
                                              int(transposed))


def rk3_step_adjoint(P0, settings, dt, B):
    """In place on B: cotangents after one rk3 step -> before it."""
    rc = lib().vpmo_rk3_step_adjoint(_check(P0), P0.shape[0], C.addressof(settings), dt, _check(B))
    if rc != 0:
        raise ValueError("rk3_step_adjoint: unsupported settings (SFS or viscous) rc=%d" % rc)
