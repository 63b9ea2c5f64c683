"""Host-side mirror of FLOWVPM's call surface for the P2P path, over the
C-ABI of lib/libflowvpm_hip.so.

Reference interfaces mirrored here (src/ paths of FLOWVPM.jl):
  ParticleField(maxparticles; formulation, viscous, kernel, UJ, SFS, ...)
                                               FLOWVPM_particlefield.jl:120-141
  add_particle / remove_particle / get_np     FLOWVPM_particlefield.jl:167-212,401-419
  UJ_direct(pfield; reset, reset_sfs, sfs)    FLOWVPM_UJ.jl:21-37  -> UJ_hip
  UJ_direct(source, target)                   FLOWVPM_UJ.jl:48-50  -> UJ_hip_fields
  fmm.direct!(tgt, tgt_range, _, src, _, src_range)  FLOWVPM_fmm.jl:102-168 -> p2p_ranges
  Estr_direct!(pfield)                        FLOWVPM_subfilterscale_models.jl:43-92
  Estr_fmm!(... direct_list)                  FLOWVPM_subfilterscale_models.jl:157-188 -> estr_ranges
  rungekutta3(pfield, dt; relax) / nextstep   FLOWVPM_timeintegration.jl:388-461,
                                              FLOWVPM_particlefield.jl:435-460
  run_vpm!(pfield, dt, nsteps)                FLOWVPM_utils.jl:41-139 (relax and save cadence)
  save(pfield, name) / read!(pfield, h5)      FLOWVPM_utils.jl:148-328,493-571 -> h5io
  zeta_direct(pfield)                         FLOWVPM_viscous.jl:486-515
  kernel / formulation / relaxation / SFS / viscous aliases   FLOWVPM.jl:80-256

Indices are 0-based (particle i of Julia is i-1 here); particle storage is
the reference's own 46-double record per particle.  Every numerical call
goes through the HIP library; nothing here computes U or J on the CPU.
"""
import ctypes as C
import math
from dataclasses import dataclass

import numpy
import numpy as np

from . import _lib, h5io
from .h5io import read_into, save  # noqa: F401  (save / read! of FLOWVPM_utils.jl)
from .fields import NFIELDS, STATIC_INDEX, make_particles
from ._lib import VPMError, check, lib  # noqa: F401

# ---------------------------------------------------------------- options
SINGULAR, GAUSSIAN, GAUSSIANERF, WINCKELMANS = "singular", "gaussian", "gaussianerf", "winckelmans"
# kernel aliases (FLOWVPM.jl:129-164)
singular, gaussian, gaussianerf, winckelmans = SINGULAR, GAUSSIAN, GAUSSIANERF, WINCKELMANS
kernel_singular, kernel_gaussian, kernel_gaussianerf, kernel_winckelmans = singular, gaussian, gaussianerf, winckelmans
kernel_default = gaussianerf


@dataclass(frozen=True)
class ReformulatedVPM:
    """ReformulatedVPM{f,g} (FLOWVPM_formulation.jl:31-37)."""
    f: float = 0.0
    g: float = 0.2

    @property
    def h(self):
        return (1 - 3 * self.g) / (1 + 3 * self.f)


formulation_cVPM = cVPM = ReformulatedVPM(0.0, 0.0)                 # FLOWVPM.jl:84
formulation_rVPM = rVPM = ReformulatedVPM(0.0, 1 / 5)               # :85
formulation_tube_continuity = ReformulatedVPM(1 / 2, 0.0)           # :93
formulation_tube_momentum = ReformulatedVPM(1 / 4, 1 / 4)           # :101
formulation_sphere_momentum = ReformulatedVPM(0.0, 1 / 5 + 1e-8)    # :102
formulation_default = rVPM


@dataclass(frozen=True)
class Relaxation:
    """Relaxation(relax, nsteps_relax, rlxf) (FLOWVPM_relaxation.jl:24-32)."""
    method: int          # 0 none, 1 pedrizzetti, 2 corrected pedrizzetti
    nsteps_relax: int
    rlxf: float


relaxation_none = norelaxation = Relaxation(0, -1, 0.0)              # FLOWVPM.jl:168
relaxation_pedrizzetti = pedrizzetti = Relaxation(1, 1, 0.3)         # :169,177
relaxation_correctedpedrizzetti = correctedpedrizzetti = Relaxation(2, 1, 0.3)
relaxation_default = pedrizzetti


@dataclass(frozen=True)
class NoSFS:
    """NoSFS (FLOWVPM_subfilterscale.jl:63-82)."""


@dataclass(frozen=True)
class DynamicSFS:
    """DynamicSFS with the pseudo-3-level procedure (FLOWVPM_subfilterscale.jl:172-268)."""
    alpha: float = 0.667
    rlxf: float = 0.005
    minC: float = 0.0
    maxC: float = 1.0
    force_positive: bool = False
    clip_backscatter: bool = False


SFS_none = noSFS = NoSFS()
# FLOWVPM.jl:223 and :231
SFS_Cd_twolevel_nobackscatter = DynamicSFS(alpha=0.999, force_positive=True, clip_backscatter=True)
SFS_Cd_threelevel_nobackscatter = DynamicSFS(alpha=0.667, force_positive=True, clip_backscatter=True)


@dataclass(frozen=True)
class Inviscid:
    """Inviscid (FLOWVPM_viscous.jl:40-58)."""


@dataclass
class CoreSpreading:
    """CoreSpreading(nu, sgm0; beta=1.5, itmax=15, tol=1e-3, iterror=True)
    (FLOWVPM_viscous.jl:63-223): core spreading, and the core-size reset by
    RBF interpolation with the conjugate gradient (:202-220, :314-478) when
    the growth reaches beta, with zeta = zeta_direct.  itmax=0 leaves the
    reset to the caller: the step then raises VPMError(VPM_ERESET).  An
    unconverged CG with iterror raises VPMError(VPM_ECONVERGE)."""
    nu: float
    sgm0: float
    beta: float = 1.5
    itmax: int = 15
    tol: float = 1e-3
    iterror: bool = True
    t_sgm: float = 0.0


def nofreestream(t):
    return (0.0, 0.0, 0.0)


# ---------------------------------------------------------------- device field
class Field:
    """A device-resident particle field (vpm_field of the C-ABI)."""

    def __init__(self, maxparticles, device=0):
        h = C.c_void_p()
        check(lib().vpm_create(int(device), int(maxparticles), C.byref(h)))
        self._h = h
        self.maxparticles = int(maxparticles)
        self.device = int(device)

    @property
    def handle(self):
        if self._h is None:
            raise VPMError(_lib.VPM_EINVAL, "field is closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().vpm_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_particles(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64)
        assert P.ndim == 2 and P.shape[1] == NFIELDS
        check(lib().vpm_set_particles(self.handle, P.ctypes.data_as(_lib._dptr), P.shape[0]))

    def add_particles(self, P):
        P = np.ascontiguousarray(P, dtype=np.float64).reshape(-1, NFIELDS)
        check(lib().vpm_add_particles(self.handle, P.ctypes.data_as(_lib._dptr), P.shape[0]))

    def remove_particle(self, i):
        check(lib().vpm_remove_particle(self.handle, int(i)))

    def get_particles(self, out, rows=(0, NFIELDS), np_=None):
        assert out.dtype == np.float64 and out.flags.c_contiguous and out.shape[1] == NFIELDS
        n = out.shape[0] if np_ is None else np_
        check(lib().vpm_get_particles(self.handle, out.ctypes.data_as(_lib._dptr), n, rows[0], rows[1]))
        return out

    @property
    def np(self):
        v = C.c_int64()
        check(lib().vpm_get_np(self.handle, C.byref(v)))
        return v.value

    def synchronize(self):
        check(lib().vpm_synchronize(self.handle))

    def uj_direct(self, kernel=kernel_default, precision="f32c", reset=True, reset_sfs=False, sfs=False,
                  transposed=True, sort=True):
        """vpm_uj_direct.  precision "f64far" is FP64 with the far form
        (VPM_PREC_F64 + VPM_UJ_F64_FAR); E_str (sfs) then runs in FP64."""
        flags = ((_lib.UJ_RESET if reset else 0) | (_lib.UJ_RESET_SFS if reset_sfs else 0) |
                 (_lib.UJ_SFS if sfs else 0) | (0 if sort else _lib.UJ_NO_SORT))
        if precision == "f64far":
            flags |= _lib.UJ_F64_FAR
            precision = "f64"
        cfg = make_config(kernel, precision, transposed)
        check(lib().vpm_uj_direct(self.handle, C.byref(cfg), flags))

    def estr_direct(self, kernel=kernel_default, precision="f32c", transposed=True):
        cfg = make_config(kernel, precision, transposed)
        check(lib().vpm_estr_direct(self.handle, C.byref(cfg)))

    def zeta_direct(self, kernel=kernel_default, precision="f32c"):
        """vpm_zeta_direct: J[1:3] of every particle <- the zeta field."""
        cfg = make_config(kernel, precision, True)
        check(lib().vpm_zeta_direct(self.handle, C.byref(cfg)))

    def core_spreading_reset(self, sgm0, kernel=kernel_default, precision="f32c", itmax=15, tol=1e-3,
                             iterror=True):
        """vpm_core_spreading_reset on the resident particles; returns
        (iterations, residuals sqrt(rr/rr0) per component)."""
        cfg = make_config(kernel, precision, True)
        it = C.c_int32(0)
        res = np.zeros(3)
        check(lib().vpm_core_spreading_reset(self.handle, C.byref(cfg), float(sgm0), int(itmax), float(tol),
                                             int(bool(iterror)), C.byref(it), res.ctypes.data_as(_lib._dptr)))
        return it.value, res

    def rbf_conjugategradient(self, kernel=kernel_default, precision="f32c", itmax=15, tol=1e-3, iterror=True):
        """vpm_rbf_conjugategradient on the resident particles (target
        vorticity in M[7:9]); returns (iterations, residuals)."""
        cfg = make_config(kernel, precision, True)
        it = C.c_int32(0)
        res = np.zeros(3)
        check(lib().vpm_rbf_conjugategradient(self.handle, C.byref(cfg), int(itmax), float(tol),
                                              int(bool(iterror)), C.byref(it), res.ctypes.data_as(_lib._dptr)))
        return it.value, res

    def uj_adjoint(self, Ubar, Jbar, kernel=kernel_default, precision="f32c"):
        """vpm_uj_adjoint on the resident particles: returns (dL/dX (n,3),
        dL/dGamma (n,3), dL/dsigma (n,)) for L = sum Ubar.U + sum Jbar:J."""
        n = self.np
        ub = np.ascontiguousarray(Ubar, dtype=np.float64).reshape(n * 3)
        jb = np.ascontiguousarray(Jbar, dtype=np.float64).reshape(n * 9)
        Xb, Gb, Sb = np.zeros(3 * n), np.zeros(3 * n), np.zeros(n)
        cfg = make_config(kernel, precision, True)
        d = _lib._dptr
        check(lib().vpm_uj_adjoint(self.handle, C.byref(cfg), ub.ctypes.data_as(d), jb.ctypes.data_as(d),
                                   Xb.ctypes.data_as(d), Gb.ctypes.data_as(d), Sb.ctypes.data_as(d)))
        return Xb.reshape(n, 3), Gb.reshape(n, 3), Sb

    def plan_ranges(self, tb, te, sb, se):
        """vpm_plan_ranges: the pair list planned once on this field (see
        RangePlan)."""
        return RangePlan(self, tb, te, sb, se)

    def profile(self, enable=True):
        check(lib().vpm_profile(self.handle, int(bool(enable))))

    def timing(self, which):
        ms, n = C.c_double(), C.c_int64()
        check(lib().vpm_timing(self.handle, int(which), C.byref(ms), C.byref(n)))
        return ms.value, n.value


def _range_arrays(tb, te, sb, se):
    arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (tb, te, sb, se)]
    if not (len(arrs[0]) == len(arrs[1]) == len(arrs[2]) == len(arrs[3])):
        raise VPMError(_lib.VPM_EINVAL, "range arrays differ in length")
    return arrs


class RangePlan:
    """A direct list of 0-based half-open (target, source) range pairs,
    checked and planned once for a device field (vpm_plan_ranges) and served
    to the U/J, E_str and zeta list kernels without re-planning or
    re-uploading it: UJ_fmm hands one direct_list to the near field and then
    to Estr_fmm! (FLOWVPM_UJ.jl:90-125).  Valid while the field holds the
    particle count it was made for."""

    def __init__(self, field, tb, te, sb, se):
        arrs = _range_arrays(tb, te, sb, se)
        ptr = [a.ctypes.data_as(_lib._i64p) for a in arrs]
        h = C.c_void_p()
        check(lib().vpm_plan_ranges(field.handle, ptr[0], ptr[1], ptr[2], ptr[3], len(arrs[0]), C.byref(h)))
        self._h = h
        self.field = field
        self.npairs = len(arrs[0])

    @property
    def handle(self):
        if self._h is None:
            raise VPMError(_lib.VPM_EINVAL, "plan is closed")
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().vpm_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def p2p(self, kernel=kernel_default, precision="f32c"):
        """vpm_p2p_plan: U and J of every pair accumulated (fmm.direct!)."""
        cfg = make_config(kernel, precision, True)
        check(lib().vpm_p2p_plan(self.field.handle, C.byref(cfg), self.handle))

    def estr(self, kernel=kernel_default, precision="f32c", transposed=True):
        """vpm_estr_plan: E_str of every pair onto SFS (Estr_fmm!)."""
        cfg = make_config(kernel, precision, transposed)
        check(lib().vpm_estr_plan(self.field.handle, C.byref(cfg), self.handle))

    def zeta(self, kernel=kernel_default, precision="f32c"):
        """vpm_zeta_plan: J[1:3] += the zeta field of every pair's sources."""
        cfg = make_config(kernel, precision, True)
        check(lib().vpm_zeta_plan(self.field.handle, C.byref(cfg), self.handle))


def make_config(kernel=kernel_default, precision="f32c", transposed=True):
    cfg = _lib.Config()
    cfg.kernel = _lib.KERNELS[kernel]
    if precision == "f64far":     # FP64; the far form is a flag of the U/J calls
        precision = "f64"
    if precision not in _lib.PRECISIONS:
        raise ValueError("precision %r: one of %s or \"f64far\"" % (precision, sorted(_lib.PRECISIONS)))
    cfg.precision = _lib.PRECISIONS[precision]
    cfg.transposed = int(bool(transposed))
    return cfg


_field_cache = {}


def _scratch_field(n, device):
    """A cached device field with capacity >= n (avoids re-allocating per call)."""
    f = _field_cache.get(device)
    if f is None or f.maxparticles < n:
        if f is not None:
            f.close()
        f = Field(max(n, 1), device)
        _field_cache[device] = f
    return f


def release_scratch():
    for f in _field_cache.values():
        f.close()
    _field_cache.clear()


# ---------------------------------------------------------------- ParticleField
class ParticleField:
    """ParticleField (FLOWVPM_particlefield.jl:67-141) with the device as the
    compute target.  `particles` is the (maxparticles, 46) record matrix."""

    def __init__(self, maxparticles, *, formulation=formulation_default, viscous=None, np=0, nt=0, t=0.0,
                 transposed=True, SFS=noSFS, kernel=kernel_default, UJ=None, Uinf=nofreestream,
                 relaxation=relaxation_default, integration=None, precision="f32c", device=0):
        self.maxparticles = int(maxparticles)
        self.particles = numpy.zeros((self.maxparticles, NFIELDS))
        self.formulation = formulation
        self.viscous = viscous if viscous is not None else Inviscid()
        self.np, self.nt, self.t = int(np), int(nt), float(t)
        self.kernel = kernel
        self.UJ = UJ if UJ is not None else UJ_hip
        self.Uinf = Uinf
        self.SFS = SFS
        self.integration = integration if integration is not None else rungekutta3
        self.transposed = bool(transposed)
        self.relaxation = relaxation
        self.precision = precision
        self.device = int(device)
        self._field = None

    def device_field(self):
        if self._field is None or self._field.maxparticles < self.maxparticles:
            self._field = Field(self.maxparticles, self.device)
        return self._field

    def close(self):
        """Release the device copy (and its scratch buffers) now."""
        if self._field is not None:
            self._field.close()
            self._field = None


def isLES(pfield):
    return isinstance(pfield.SFS, DynamicSFS)


def get_np(pfield):
    return pfield.np


def add_particle(pfield, X, Gamma, sigma, *, vol=0.0, circulation=1.0, C=0.0, static=False):
    """add_particle (FLOWVPM_particlefield.jl:167-194)."""
    if pfield.np == pfield.maxparticles:
        raise VPMError(_lib.VPM_ERANGE, "PARTICLE OVERFLOW. Max number of particles %d has been reached"
                       % pfield.maxparticles)
    rec = make_particles([X], [Gamma], sigma, vol=vol, circulation=circulation, C=C,
                         static=[float(bool(static))])[0]
    pfield.particles[pfield.np] = rec
    pfield.np += 1


def add_particles(pfield, P):
    """Append many 46-double records at once."""
    P = np.asarray(P, dtype=np.float64).reshape(-1, NFIELDS)
    if pfield.np + P.shape[0] > pfield.maxparticles:
        raise VPMError(_lib.VPM_ERANGE, "PARTICLE OVERFLOW. Max number of particles %d has been reached"
                       % pfield.maxparticles)
    pfield.particles[pfield.np:pfield.np + P.shape[0]] = P
    pfield.np += P.shape[0]


def remove_particle(pfield, i):
    """remove_particle (FLOWVPM_particlefield.jl:401-419): the last particle
    moves into slot i (0-based), and the last slot is reset."""
    if i < 0:
        raise VPMError(_lib.VPM_EINVAL, "Requested removal of invalid particle index %d" % i)
    if i >= pfield.np:
        raise VPMError(_lib.VPM_EINVAL, "Requested removal of particle %d, but there is only %d particles "
                       "in the field." % (i, pfield.np))
    last = pfield.np - 1
    if i != last:
        pfield.particles[i] = pfield.particles[last]
    pfield.particles[last, 9:27] = 0.0          # _reset_particle (particlefield.jl:485-490)
    pfield.np -= 1


def _as_matrix(target):
    if isinstance(target, ParticleField):
        return target.particles[:target.np]
    P = target
    if not (isinstance(P, np.ndarray) and P.dtype == np.float64 and P.ndim == 2 and P.shape[1] == NFIELDS
            and P.flags.c_contiguous):
        raise VPMError(_lib.VPM_EINVAL, "particles must be a C-contiguous (n, 46) float64 array")
    return P


# ---------------------------------------------------------------- UJ functors
def UJ_hip(pfield, *, reset=True, reset_sfs=False, sfs=False, kernel=None, precision=None,
           transposed=None, device=None, sort=True, rbf=False):
    """The UJ functor with UJ_direct semantics (FLOWVPM_UJ.jl:21-37), computed
    on the GPU.  Accepts a ParticleField or an (n, 46) record matrix; mutates
    U, vorticity, J, PSE (reset) and SFS rows in place.  `rbf` is ignored, as
    in the reference."""
    P = _as_matrix(pfield)
    pf = pfield if isinstance(pfield, ParticleField) else None
    kernel = kernel or (pf.kernel if pf else kernel_default)
    precision = precision or (pf.precision if pf else "f32c")
    transposed = (pf.transposed if pf else True) if transposed is None else transposed
    device = (pf.device if pf else 0) if device is None else device
    n = P.shape[0]
    if n == 0:
        return None
    f = pf.device_field() if pf else _scratch_field(n, device)
    f.set_particles(P)
    f.uj_direct(kernel, precision, reset, reset_sfs, sfs, transposed, sort)
    _copy_back(f, P, sfs or reset_sfs)
    return None


def _copy_back(f, P, with_sfs):
    tmp = np.empty_like(P)
    f.get_particles(tmp, rows=(9, 27))
    P[:, 9:27] = tmp[:, 9:27]
    if with_sfs:
        f.get_particles(tmp, rows=(39, 42))
        P[:, 39:42] = tmp[:, 39:42]


def UJ_hip_fields(source, target, *, kernel=None, precision=None, device=None):
    """UJ_direct(source, target) (FLOWVPM_UJ.jl:48-50): accumulates the
    influence of every particle of `source` on every particle of `target`."""
    S, T = _as_matrix(source), _as_matrix(target)
    kernel = kernel or (source.kernel if isinstance(source, ParticleField) else kernel_default)
    precision = precision or (source.precision if isinstance(source, ParticleField) else "f32c")
    device = device if device is not None else 0
    fs, ft = Field(max(S.shape[0], 1), device), Field(max(T.shape[0], 1), device)
    try:
        fs.set_particles(S)
        ft.set_particles(T)
        cfg = make_config(kernel, precision, True)
        check(lib().vpm_uj_direct_fields(ft.handle, fs.handle, C.byref(cfg)))
        _copy_back(ft, T, False)
    finally:
        fs.close()
        ft.close()


def p2p_ranges(pfield, tb, te, sb, se, *, kernel=None, precision=None, device=None):
    """fmm.direct! over a batch of 0-based half-open (target, source) range
    pairs of one field (FLOWVPM_fmm.jl:102-168); accumulates U and J."""
    P = _as_matrix(pfield)
    kernel = kernel or (pfield.kernel if isinstance(pfield, ParticleField) else kernel_default)
    precision = precision or (pfield.precision if isinstance(pfield, ParticleField) else "f32c")
    f = _scratch_field(P.shape[0], 0 if device is None else device)
    f.set_particles(P)
    arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (tb, te, sb, se)]
    if not (len(arrs[0]) == len(arrs[1]) == len(arrs[2]) == len(arrs[3])):
        raise VPMError(_lib.VPM_EINVAL, "range arrays differ in length")
    cfg = make_config(kernel, precision, True)
    ptr = [a.ctypes.data_as(_lib._i64p) for a in arrs]
    check(lib().vpm_p2p_ranges(f.handle, C.byref(cfg), ptr[0], ptr[1], ptr[2], ptr[3], len(arrs[0])))
    tmp = np.empty_like(P)
    f.get_particles(tmp, rows=(9, 24))
    P[:, 9:24] = tmp[:, 9:24]


def estr_ranges(pfield, tb, te, sb, se, *, kernel=None, precision=None, transposed=None, device=None):
    """Estr_fmm! over a direct list of 0-based half-open (target, source)
    range pairs (FLOWVPM_subfilterscale_models.jl:157-188): E_str pairs onto
    SFS, no static skip, no reset."""
    P = _as_matrix(pfield)
    pf = pfield if isinstance(pfield, ParticleField) else None
    kernel = kernel or (pf.kernel if pf else kernel_default)
    precision = precision or (pf.precision if pf else "f32c")
    transposed = (pf.transposed if pf else True) if transposed is None else transposed
    f = _scratch_field(P.shape[0], 0 if device is None else device)
    f.set_particles(P)
    arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (tb, te, sb, se)]
    if not (len(arrs[0]) == len(arrs[1]) == len(arrs[2]) == len(arrs[3])):
        raise VPMError(_lib.VPM_EINVAL, "range arrays differ in length")
    cfg = make_config(kernel, precision, transposed)
    ptr = [a.ctypes.data_as(_lib._i64p) for a in arrs]
    check(lib().vpm_estr_ranges(f.handle, C.byref(cfg), ptr[0], ptr[1], ptr[2], ptr[3], len(arrs[0])))
    tmp = np.empty_like(P)
    f.get_particles(tmp, rows=(39, 42))
    P[:, 39:42] = tmp[:, 39:42]


def zeta_ranges(pfield, tb, te, sb, se, *, kernel=None, precision=None, device=None):
    """zeta_fmm's pair loop (FLOWVPM_viscous.jl:523-558) over a direct list
    of 0-based half-open (target, source) range pairs: J[1:3] += the zeta
    field of the sources, no reset, no static skip."""
    P = _as_matrix(pfield)
    pf = pfield if isinstance(pfield, ParticleField) else None
    kernel = kernel or (pf.kernel if pf else kernel_default)
    precision = precision or (pf.precision if pf else "f32c")
    f = _scratch_field(P.shape[0], 0 if device is None else device)
    f.set_particles(P)
    arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (tb, te, sb, se)]
    if not (len(arrs[0]) == len(arrs[1]) == len(arrs[2]) == len(arrs[3])):
        raise VPMError(_lib.VPM_EINVAL, "range arrays differ in length")
    cfg = make_config(kernel, precision, True)
    ptr = [a.ctypes.data_as(_lib._i64p) for a in arrs]
    check(lib().vpm_zeta_ranges(f.handle, C.byref(cfg), ptr[0], ptr[1], ptr[2], ptr[3], len(arrs[0])))
    tmp = np.empty_like(P)
    f.get_particles(tmp, rows=(15, 18))
    P[:, 15:18] = tmp[:, 15:18]


def Estr_hip(pfield, *, kernel=None, precision=None, transposed=None, device=None):
    """Estr_direct! (FLOWVPM_subfilterscale_models.jl:43-92) on the GPU."""
    P = _as_matrix(pfield)
    pf = pfield if isinstance(pfield, ParticleField) else None
    kernel = kernel or (pf.kernel if pf else kernel_default)
    precision = precision or (pf.precision if pf else "f32c")
    transposed = (pf.transposed if pf else True) if transposed is None else transposed
    f = _scratch_field(P.shape[0], 0 if device is None else device)
    f.set_particles(P)
    f.estr_direct(kernel, precision, transposed)
    tmp = np.empty_like(P)
    f.get_particles(tmp, rows=(39, 42))
    P[:, 39:42] = tmp[:, 39:42]


# ---------------------------------------------------------------- time stepping
def step_config(pfield):
    s = _lib.StepConfig()
    s.uj = make_config(pfield.kernel, pfield.precision, pfield.transposed)
    s.uj_flags = _lib.UJ_F64_FAR if pfield.precision == "f64far" else 0
    s.f, s.g = pfield.formulation.f, pfield.formulation.g
    s.Uinf[:] = [float(u) for u in pfield.Uinf(pfield.t)]
    sfs = pfield.SFS
    if isinstance(sfs, DynamicSFS):
        s.sfs, s.sfs_alpha, s.sfs_rlxf, s.sfs_minC, s.sfs_maxC = 1, sfs.alpha, sfs.rlxf, sfs.minC, sfs.maxC
        s.sfs_force_positive, s.sfs_clip = int(sfs.force_positive), int(sfs.clip_backscatter)
    s.relax, s.rlxf = pfield.relaxation.method, pfield.relaxation.rlxf
    if isinstance(pfield.viscous, CoreSpreading):
        s.viscous, s.nu, s.sgm0 = 1, pfield.viscous.nu, pfield.viscous.sgm0
        s.beta, s.t_sgm = pfield.viscous.beta, pfield.viscous.t_sgm
        s.rbf_itmax, s.rbf_tol = int(pfield.viscous.itmax), float(pfield.viscous.tol)
        s.rbf_iterror = int(bool(pfield.viscous.iterror))
    return s


def rungekutta3(pfield, dt, *, relax=False, resident=False):
    """rungekutta3 for ReformulatedVPM (FLOWVPM_timeintegration.jl:388-461),
    including nextstep's U_prev update, run on the device.  With
    resident=True the host copy is not re-uploaded/downloaded (the caller
    syncs explicitly with pfield_upload / pfield_download)."""
    f = pfield.device_field()
    if not resident:
        f.set_particles(pfield.particles[:pfield.np])
    s = step_config(pfield)
    rc = lib().vpm_rk3_step(f.handle, C.byref(s), float(dt), int(bool(relax)))
    if isinstance(pfield.viscous, CoreSpreading):
        pfield.viscous.t_sgm = s.t_sgm
    if not resident:
        f.get_particles(pfield.particles, np_=pfield.np)
    check(rc)


def pfield_upload(pfield):
    pfield.device_field().set_particles(pfield.particles[:pfield.np])


def pfield_download(pfield):
    pfield.device_field().get_particles(pfield.particles, np_=pfield.np)


def nextstep(pfield, dt, *, relax=False, resident=False):
    """nextstep (FLOWVPM_particlefield.jl:435-460): one integration step (the
    device step already updates U_prev), then t and nt."""
    if pfield.np != 0:
        pfield.integration(pfield, dt, relax=relax, resident=resident)
    pfield.t += dt
    pfield.nt += 1


def _resident_callback(pfield, fun, dt):
    """Calls fun(pfield, t, dt) on a device-resident field as the reference
    calls it on the live field (FLOWVPM_utils.jl:100,117): the host copy is
    downloaded first, and whatever the callback changed (particles appended,
    removed or edited) is brought back to the device afterwards.  Appending
    only takes vpm_add_particles; any other change re-uploads the field.  The
    callback's return value is left in pfield._last_callback_result."""
    np0 = pfield.np
    if np0:
        pfield_download(pfield)
    before = pfield.particles[:np0].copy()
    pfield._last_callback_result = fun(pfield, pfield.t, dt)
    n = pfield.np
    kept = pfield.particles[:min(n, np0)]
    same_prefix = n >= np0 and np.array_equal(kept, before, equal_nan=True)
    if same_prefix and n == np0:
        return
    f = pfield.device_field()
    if same_prefix:
        f.add_particles(pfield.particles[np0:n])
    else:
        f.set_particles(pfield.particles[:n])


def run_vpm(pfield, dt, nsteps, *, runtime_function=None, static_particles_function=None, resident=True,
            save_path=None, save_pfield=True, create_savepath=True, run_name="pfield", nsteps_save=1,
            prompt=True, save_time=True):
    """The stepping loop of run_vpm! (FLOWVPM_utils.jl:86-136): relaxation
    every nsteps_relax steps, starting at step 1; before each step
    static_particles_function(pfield, t, dt) may append particles, which are
    removed after the step unless it returns False (:100-112); then
    runtime_function(pfield, t, dt) ends the run by returning True.  The
    viscous scheme must be compatible with the kernel (:56-64,
    FLOWVPM.jl:259-271).  State stays on the device between steps when
    resident=True.  Both callbacks see the live field: it is downloaded
    before each call, and what the call changed (appended, removed or edited
    particles) goes back to the device after it.  Particles appended by
    static_particles_function are removed after the step with
    vpm_remove_particle.

    With save_path, the folder is made (an existing one is replaced when
    create_savepath, :66-68) and the field is saved as <run_name>.<nt>.h5 +
    .xmf at step 0, every nsteps_save steps, at the last step and at a
    runtime_function break (:124-129); save_time=False stamps the files
    with nt instead of t.  The settings file (BSON) is not written."""
    if isinstance(pfield.viscous, CoreSpreading) and pfield.kernel != "gaussianerf":
        raise VPMError(_lib.VPM_EINVAL, "Kernel %s is not compatible with viscous scheme CoreSpreading; "
                       "compatible kernels are ['gaussianerf']" % pfield.kernel)
    if save_path is not None and create_savepath:
        h5io.create_path(save_path, prompt)
    if resident:
        pfield_upload(pfield)
    for i in range(0, nsteps + 1):
        rl = pfield.relaxation
        relax = rl.method != 0 and rl.nsteps_relax >= 1 and i > 0 and i % rl.nsteps_relax == 0
        org_np = pfield.np
        downloaded = False
        if i != 0:
            remove = None
            if static_particles_function is not None:
                if resident:
                    # the callback reads and may change the live field
                    _resident_callback(pfield, static_particles_function, dt)
                    remove = pfield._last_callback_result
                else:
                    remove = static_particles_function(pfield, pfield.t, dt)
            nextstep(pfield, dt, relax=relax, resident=resident)
            if remove is None or remove:
                for pi in range(pfield.np - 1, org_np - 1, -1):
                    remove_particle(pfield, pi)
                    if resident:
                        pfield.device_field().remove_particle(pi)
        breakflag = False
        if runtime_function is not None:
            if resident:
                _resident_callback(pfield, runtime_function, dt)
                downloaded = True
                breakflag = bool(pfield._last_callback_result)
            else:
                breakflag = bool(runtime_function(pfield, pfield.t, dt))
        if save_pfield and save_path is not None and (i % nsteps_save == 0 or i == nsteps or breakflag):
            if resident and pfield.np and not downloaded:
                pfield_download(pfield)
            h5io.save(pfield, run_name, path=save_path, add_num=True,
                      overwrite_time=None if save_time else pfield.nt)
        if breakflag:
            break
    if resident and pfield.np:
        pfield_download(pfield)


# ---------------------------------------------------------------- adjoint
def uj_adjoint(pfield, Ubar, Jbar, *, kernel=None, precision="f32c", device=None):
    """Gradient of L = sum Ubar.U + sum Jbar:J through one all-pairs U/J
    evaluation: returns (dL/dX (n,3), dL/dGamma (n,3), dL/dsigma (n,)).
    precision "f32c": fused FP32 pair kernel with FP64 partial sums;
    "f64": FP64 pair arithmetic."""
    P = _as_matrix(pfield)
    kernel = kernel or (pfield.kernel if isinstance(pfield, ParticleField) else kernel_default)
    n = P.shape[0]
    f = _scratch_field(n, 0 if device is None else device)
    f.set_particles(P)
    return f.uj_adjoint(Ubar, Jbar, kernel, precision)


def zeta_direct(pfield, *, precision=None):
    """zeta_direct(pfield) (FLOWVPM_viscous.jl:486-493): J[1:3] of every
    particle (static included) <- sum_j Gamma_j zeta(r/sigma_j)/sigma_j^3, on
    the device; the host copy is updated."""
    if pfield.np == 0:
        return
    f = pfield.device_field()
    f.set_particles(pfield.particles[:pfield.np])
    f.zeta_direct(pfield.kernel, precision or pfield.precision)
    f.get_particles(pfield.particles, rows=(15, 18), np_=pfield.np)


def core_spreading_reset(pfield, scheme=None, *, precision=None):
    """The CoreSpreading reset (FLOWVPM_viscous.jl:202-220): target vorticity
    M[7:9] <- zeta_direct field, sigma <- sgm0, and Gamma by
    rbf_conjugategradient (:314-478), on the device; the host copy is
    updated and scheme.t_sgm set to 0.  Returns (iterations, residuals)."""
    scheme = scheme or pfield.viscous
    if not isinstance(scheme, CoreSpreading):
        raise TypeError("core_spreading_reset needs a CoreSpreading scheme")
    out = (0, np.zeros(3))
    if pfield.np > 0:
        f = pfield.device_field()
        f.set_particles(pfield.particles[:pfield.np])
        try:
            out = f.core_spreading_reset(scheme.sgm0, pfield.kernel, precision or pfield.precision,
                                         scheme.itmax, scheme.tol, scheme.iterror)
        finally:
            f.get_particles(pfield.particles, np_=pfield.np)
    scheme.t_sgm = 0.0
    return out


def rbf_conjugategradient(pfield, cs=None, *, precision=None):
    """rbf_conjugategradient(pfield, cs) (FLOWVPM_viscous.jl:314-478) on the
    device: Gamma of the non-static particles by CG on the target vorticity
    in M[7:9]; the host copy is updated.  Returns (iterations, residuals)."""
    cs = cs or pfield.viscous
    out = (0, np.zeros(3))
    if pfield.np > 0:
        f = pfield.device_field()
        f.set_particles(pfield.particles[:pfield.np])
        try:
            out = f.rbf_conjugategradient(pfield.kernel, precision or pfield.precision, cs.itmax, cs.tol,
                                          cs.iterror)
        finally:
            f.get_particles(pfield.particles, np_=pfield.np)
    return out


class Tape:
    """Per-stage checkpoints of one rk3 step of a device field (vpm_tape):
    the particle state after each stage's UJ evaluation, in device memory."""

    def __init__(self, field):
        self.field = field
        h = C.c_void_p()
        check(lib().vpm_tape_create(field.handle, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        if self._h is None:
            raise ValueError("tape is closed")
        return self._h

    def close(self):
        if self._h is not None:
            lib().vpm_tape_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def rungekutta3_taped(pfield, dt, tape):
    """rungekutta3 on the device-resident state (ReformulatedVPM without SFS,
    inviscid, no relaxation) that records its stage states into tape."""
    f = pfield.device_field()
    s = step_config(pfield)
    check(lib().vpm_rk3_step_taped(f.handle, C.byref(s), float(dt), tape.handle))


def rk3_adjoint(tape, bar):
    """Pullback of the step recorded in tape, in place on bar ((n, 46)
    cotangent records: after the step in, before the step out)."""
    bar = np.ascontiguousarray(bar, dtype=np.float64)
    check(lib().vpm_rk3_step_adjoint(tape.handle, bar.ctypes.data_as(_lib._dptr)))
    return bar


def kinetic_energy_rollout_gradient(pfield, dt, nsteps, timings=None):
    """Gradient of L = sum_i |U_i|^2 after nsteps rk3 steps with respect to
    X, Gamma and sigma of the initial state (pfield.particles): nsteps taped
    steps (per-stage checkpoints on the device), one U/J evaluation of the
    final state, then the adjoint back through the final U/J evaluation and
    every recorded stage.  Returns (L, bar) with bar the (n, 46) cotangent
    records (rows X, Gamma, sigma).  pfield is left at the initial state.
    With a dict `timings`, the wall times of the forward and the adjoint
    phases (ms, every call synchronizes) and the tape size are stored in it."""
    import time
    n = pfield.np
    f = pfield.device_field()
    f.set_particles(pfield.particles[:n])
    tapes = []
    try:
        t0 = time.perf_counter()
        marks = []
        for _ in range(nsteps):
            tapes.append(Tape(f))
            rungekutta3_taped(pfield, dt, tapes[-1])
            marks.append(time.perf_counter())
        f.uj_direct(pfield.kernel, pfield.precision, transposed=pfield.transposed)
        out = np.zeros((n, NFIELDS))
        f.get_particles(out, rows=(9, 12))
        t1 = time.perf_counter()
        U = out[:, 9:12]
        L = float(np.sum(U * U))
        Xb, Gb, Sb = f.uj_adjoint(2.0 * U, np.zeros((n, 9)), pfield.kernel, pfield.precision)
        bar = np.zeros((n, NFIELDS))
        bar[:, 0:3], bar[:, 3:6], bar[:, 6] = Xb, Gb, Sb
        t_adj0 = time.perf_counter()
        for t in reversed(tapes):
            bar = rk3_adjoint(t, bar)
        t2 = time.perf_counter()
        if timings is not None:
            steps = np.diff([t0] + marks)
            timings.update(forward_ms=1e3 * (t1 - t0), adjoint_ms=1e3 * (t2 - t1),
                           forward_step_ms=[1e3 * x for x in steps], final_uj_ms=1e3 * (t1 - marks[-1]),
                           final_adjoint_ms=1e3 * (t_adj0 - t1), step_adjoint_ms=1e3 * (t2 - t_adj0) / nsteps,
                           tape_bytes=3 * nsteps * n * NFIELDS * 8)
    finally:
        for t in tapes:
            t.close()
    f.set_particles(pfield.particles[:n])
    return L, bar


def zeta0(kernel):
    """kernel.zeta(0) (FLOWVPM_kernel.jl:45,51,60,69-74)."""
    return {singular: 1.0, gaussian: 3 / (4 * math.pi), gaussianerf: 1 / (2 * math.pi) ** 1.5,
            winckelmans: 7.5 / (4 * math.pi)}[kernel]
