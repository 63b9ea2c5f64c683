"""Particle-field files: FLOWVPM's HDF5 + XDMF format (src/FLOWVPM_utils.jl).

  save(pfield, file_name; path, add_num, num, createpath, overwrite_time)
                                                    FLOWVPM_utils.jl:148-328
  read!(pfield, h5_fname; path, overwrite, load_time)  -> read_into
                                                    FLOWVPM_utils.jl:493-571
  create_path(save_path, prompt)                    FLOWVPM_utils.jl:578-600

What the reference writes, and so what is written and read here:
  - scalars "np" (Int64), "nt" (Int64), "t" (Float64);
  - the rows X, Gamma, sigma, circulation, vol, static, velocity,
    velocity_gradient_x/y/z (J[1:3], J[4:6], J[7:9]), vorticity, and C when
    the field has an SFS model, all Float64.
HDF5.jl stores a Julia `k x np` column-major matrix with the dataspace
dimensions reversed, so in the file (C order) these are (np, k) datasets
whose row i is particle i's k values; single rows are (np,) vectors.  That
is exactly a column block of the (n, 46) record matrix.

The file is written and read through the HDF5 C library (libhdf5, 1.10 or
later) over ctypes.  Set VPM_HDF5_LIB to its path when it is not on the
loader path; a missing library raises an error, there is no other writer.
"""
import ctypes as C
import ctypes.util
import glob
import os
import shutil

import numpy as np

from .fields import (C_INDEX, CIRCULATION_INDEX, GAMMA_INDEX, J_INDEX, NFIELDS, SIGMA_INDEX, STATIC_INDEX,
                     U_INDEX, VOL_INDEX, VORTICITY_INDEX, X_INDEX)

# (dataset, record columns); the order of save (FLOWVPM_utils.jl:189-214)
_ROWS = (
    ("X", X_INDEX),
    ("Gamma", GAMMA_INDEX),
    ("sigma", SIGMA_INDEX),
    ("circulation", CIRCULATION_INDEX),
    ("vol", VOL_INDEX),
    ("static", STATIC_INDEX),
    ("velocity", U_INDEX),
    ("velocity_gradient_x", slice(J_INDEX.start, J_INDEX.start + 3)),
    ("velocity_gradient_y", slice(J_INDEX.start + 3, J_INDEX.start + 6)),
    ("velocity_gradient_z", slice(J_INDEX.start + 6, J_INDEX.stop)),
    ("vorticity", VORTICITY_INDEX),
)

# XDMF attributes after the geometry (FLOWVPM_utils.jl:272-366):
# (name shown, dataset, type, data type)
_XDMF_ATTRIBUTES = (
    ("Gamma", "Gamma", "Vector", "Float"),
    ("velocity", "velocity", "Vector", "Float"),
    ("velocity gradient x", "velocity_gradient_x", "Vector", "Float"),
    ("velocity gradient y", "velocity_gradient_y", "Vector", "Float"),
    ("velocity gradient z", "velocity_gradient_z", "Vector", "Float"),
    ("vorticity", "vorticity", "Vector", "Float"),
    ("sigma", "sigma", "Scalar", "Float"),
    ("circulation", "circulation", "Scalar", "Float"),
    ("vol", "vol", "Scalar", "Float"),
    ("static", "static", "Scalar", "Int"),
)


class HDF5Error(RuntimeError):
    pass


# ---------------------------------------------------------------- libhdf5
_hid = C.c_int64          # hid_t (HDF5 >= 1.10)
_hsize = C.c_ulonglong    # hsize_t
_H5P_DEFAULT = 0
_H5S_ALL = 0
_H5S_SCALAR = 0
_H5F_ACC_RDONLY = 0
_H5F_ACC_TRUNC = 2
_H5T_INTEGER, _H5T_FLOAT = 0, 1


class _H5:
    def __init__(self, path):
        L = C.CDLL(path)
        self.L = L
        for name, res, args in (
                ("H5open", C.c_int, []),
                ("H5get_libversion", C.c_int, [C.POINTER(C.c_uint)] * 3),
                ("H5Eset_auto2", C.c_int, [_hid, C.c_void_p, C.c_void_p]),
                ("H5Fcreate", _hid, [C.c_char_p, C.c_uint, _hid, _hid]),
                ("H5Fopen", _hid, [C.c_char_p, C.c_uint, _hid]),
                ("H5Fclose", C.c_int, [_hid]),
                ("H5Screate", _hid, [C.c_int]),
                ("H5Screate_simple", _hid, [C.c_int, C.POINTER(_hsize), C.POINTER(_hsize)]),
                ("H5Sget_simple_extent_ndims", C.c_int, [_hid]),
                ("H5Sget_simple_extent_dims", C.c_int, [_hid, C.POINTER(_hsize), C.POINTER(_hsize)]),
                ("H5Sclose", C.c_int, [_hid]),
                ("H5Dcreate2", _hid, [_hid, C.c_char_p, _hid, _hid, _hid, _hid, _hid]),
                ("H5Dopen2", _hid, [_hid, C.c_char_p, _hid]),
                ("H5Dwrite", C.c_int, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
                ("H5Dread", C.c_int, [_hid, _hid, _hid, _hid, _hid, C.c_void_p]),
                ("H5Dget_space", _hid, [_hid]),
                ("H5Dget_type", _hid, [_hid]),
                ("H5Dclose", C.c_int, [_hid]),
                ("H5Tget_class", C.c_int, [_hid]),
                ("H5Tclose", C.c_int, [_hid]),
                ("H5Lexists", C.c_int, [_hid, C.c_char_p, _hid])):
            fn = getattr(L, name)
            fn.restype, fn.argtypes = res, args
        if L.H5open() < 0:
            raise HDF5Error("H5open failed")
        L.H5Eset_auto2(0, None, None)       # errors come back as statuses, unprinted
        v = [C.c_uint() for _ in range(3)]
        L.H5get_libversion(*[C.byref(x) for x in v])
        self.version = tuple(x.value for x in v)
        if self.version < (1, 10, 0):
            raise HDF5Error("libhdf5 %d.%d.%d at %s: 1.10 or later is needed (64-bit hid_t)"
                            % (self.version + (path,)))
        self.f64 = _hid.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
        self.i64 = _hid.in_dll(L, "H5T_NATIVE_LLONG_g").value
        self.f64le = _hid.in_dll(L, "H5T_IEEE_F64LE_g").value
        self.i64le = _hid.in_dll(L, "H5T_STD_I64LE_g").value

    def ok(self, status, what):
        if status < 0:
            raise HDF5Error("HDF5: %s failed" % what)
        return status


_h5 = None


def hdf5_library_path():
    """Where libhdf5 is taken from: VPM_HDF5_LIB, the loader's search, then
    the conda tree of this image (/opt/conda/lib)."""
    env = os.environ.get("VPM_HDF5_LIB")
    if env:
        return env
    found = ctypes.util.find_library("hdf5")
    if found:
        return found
    for pattern in ("/opt/conda/lib/libhdf5.so", "/opt/conda/lib/libhdf5.so.*",
                    "/usr/lib/x86_64-linux-gnu/libhdf5_serial.so*"):
        hits = sorted(glob.glob(pattern))
        if hits:
            return hits[0]
    return None


def hdf5():
    """The loaded libhdf5 (raises HDF5Error when none is found)."""
    global _h5
    if _h5 is None:
        path = hdf5_library_path()
        if path is None:
            raise HDF5Error("libhdf5 not found: set VPM_HDF5_LIB to the path of libhdf5.so")
        try:
            _h5 = _H5(path)
        except OSError as e:
            raise HDF5Error("cannot load libhdf5 from %s: %s" % (path, e)) from None
    return _h5


def _write(h5, fid, name, data, filetype, memtype):
    data = np.asarray(data)                  # a 0-d array stays 0-d: a scalar dataspace
    if not data.flags.c_contiguous:
        data = data.copy()
    if data.ndim == 0:
        space = h5.ok(h5.L.H5Screate(_H5S_SCALAR), "H5Screate")
    else:
        dims = (_hsize * data.ndim)(*data.shape)
        space = h5.ok(h5.L.H5Screate_simple(data.ndim, dims, None), "H5Screate_simple")
    try:
        d = h5.ok(h5.L.H5Dcreate2(fid, name.encode(), filetype, space, _H5P_DEFAULT, _H5P_DEFAULT,
                                  _H5P_DEFAULT), "H5Dcreate2(%s)" % name)
        try:
            h5.ok(h5.L.H5Dwrite(d, memtype, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, data.ctypes.data),
                  "H5Dwrite(%s)" % name)
        finally:
            h5.L.H5Dclose(d)
    finally:
        h5.L.H5Sclose(space)


def _read(h5, fid, name):
    """A dataset as float64 (or int64 for integer data), in its own shape."""
    d = h5.L.H5Dopen2(fid, name.encode(), _H5P_DEFAULT)
    if d < 0:
        raise HDF5Error("dataset %s missing" % name)
    try:
        space = h5.ok(h5.L.H5Dget_space(d), "H5Dget_space(%s)" % name)
        try:
            nd = h5.ok(h5.L.H5Sget_simple_extent_ndims(space), "ndims(%s)" % name)
            dims = (_hsize * max(nd, 1))()
            h5.ok(h5.L.H5Sget_simple_extent_dims(space, dims, None), "dims(%s)" % name)
            shape = tuple(int(dims[k]) for k in range(nd))
        finally:
            h5.L.H5Sclose(space)
        t = h5.ok(h5.L.H5Dget_type(d), "H5Dget_type(%s)" % name)
        cls = h5.L.H5Tget_class(t)
        h5.L.H5Tclose(t)
        if cls == _H5T_INTEGER:
            out, mem = np.empty(shape, dtype=np.int64), h5.i64
        elif cls == _H5T_FLOAT:
            out, mem = np.empty(shape, dtype=np.float64), h5.f64
        else:
            raise HDF5Error("dataset %s: type class %d is neither integer nor float" % (name, cls))
        if out.size:
            h5.ok(h5.L.H5Dread(d, mem, _H5S_ALL, _H5S_ALL, _H5P_DEFAULT, out.ctypes.data), "H5Dread(%s)" % name)
        return out
    finally:
        h5.L.H5Dclose(d)


# ---------------------------------------------------------------- files
def create_path(save_path, prompt=True):
    """create_path (FLOWVPM_utils.jl:578-600): an existing folder is removed
    and made again.  The reference's prompt always answers yes, so `prompt`
    changes nothing."""
    if os.path.isdir(save_path):
        shutil.rmtree(save_path)
    os.mkdir(save_path)


def _xdmf(h5fname, np_, time, les):
    """XDMF 3.0 description of the HDF5 datasets, with the element structure
    of FLOWVPM_utils.jl:253-375 (a Polyvertex grid over the particles)."""
    def item(dset, dims, dtype="Float"):
        return ('\t\t\t\t\t<DataItem DataType="%s" Dimensions="%s" Format="HDF" Precision="8">%s:%s</DataItem>\n'
                % (dtype, dims, h5fname, dset))

    vec, sca = "%d 3" % np_, "%d" % np_
    out = ['<?xml version="1.0" encoding="utf-8"?>\n',
           '<Xdmf xmlns:xi="http://www.w3.org/2001/XInclude" Version="3.0">\n',
           '\t<Domain>\n',
           '\t\t<Grid Name="particles" GridType="Uniform">\n',
           '\t\t\t\t<Time Value="%s" />\n' % time,
           '\t\t\t\t<Geometry Type="XYZ">\n', item("X", vec), '\t\t\t\t</Geometry>\n',
           '\t\t\t\t<Topology Dimensions="%d" Type="Polyvertex"/>\n' % np_]
    attrs = _XDMF_ATTRIBUTES + ((("C", "C", "Vector", "Float"),) if les else ())
    for name, dset, kind, dtype in attrs:
        out += ['\t\t\t\t<Attribute Center="Node" Name="%s" Type="%s">\n' % (name, kind),
                item(dset, vec if kind == "Vector" else sca, dtype),
                '\t\t\t\t</Attribute>\n']
    out += ['\t\t</Grid>\n', '\t</Domain>\n', '</Xdmf>\n']
    return "".join(out)


def _time_text(t):
    """A time stamp as Julia prints it: integers plain, floats with a
    decimal point."""
    if isinstance(t, (int, np.integer)):
        return "%d" % t
    s = repr(float(t))
    if "e" in s:
        m, e = s.split("e")
        return (m if "." in m else m + ".0") + "e" + str(int(e))
    return s


def save(pfield, file_name, *, path="", add_num=True, num=-1, createpath=False, overwrite_time=None):
    """save (FLOWVPM_utils.jl:148-328): `<file_name>.<nt>.h5` (or `.<num>`,
    or no number with add_num=False) plus the `.xmf` that describes it.
    Returns "<name>.xmf;".  The records are taken from pfield.particles (the
    host copy: download a device-resident field first).  An empty field is
    saved as one particle at the origin with zero strength (:157-166)."""
    from .api import isLES

    np_ = int(pfield.np)
    if np_ == 0:
        P = np.zeros((1, NFIELDS))
        P[0, CIRCULATION_INDEX] = 1.0            # add_particle's default circulation
        les, np_ = False, 1                      # the stand-in field has no SFS model
    else:
        P = pfield.particles[:np_]
        les = isLES(pfield)
    if createpath:
        create_path(path, True)
    fname = file_name + ((".%d" % pfield.nt if num == -1 else ".%d" % num) if add_num else "")
    h5fname = fname + ".h5"
    time = overwrite_time if overwrite_time is not None else pfield.t

    h5 = hdf5()
    fid = h5.L.H5Fcreate(os.path.join(path, h5fname).encode(), _H5F_ACC_TRUNC, _H5P_DEFAULT, _H5P_DEFAULT)
    if fid < 0:
        raise HDF5Error("cannot create %s" % os.path.join(path, h5fname))
    try:
        _write(h5, fid, "np", np.int64(np_), h5.i64le, h5.i64)
        _write(h5, fid, "nt", np.int64(pfield.nt), h5.i64le, h5.i64)
        if isinstance(time, (int, np.integer)):
            _write(h5, fid, "t", np.int64(time), h5.i64le, h5.i64)
        else:
            _write(h5, fid, "t", np.float64(time), h5.f64le, h5.f64)
        for name, cols in _ROWS:
            _write(h5, fid, name, P[:, cols], h5.f64le, h5.f64)
        if les:
            _write(h5, fid, "C", P[:, C_INDEX], h5.f64le, h5.f64)
    finally:
        h5.ok(h5.L.H5Fclose(fid), "H5Fclose")

    with open(os.path.join(path, fname + ".xmf"), "w") as f:
        f.write(_xdmf(h5fname, np_, _time_text(time), les))
    return fname + ".xmf;"


def read_into(pfield, h5_fname, *, path="", overwrite=True, load_time=True):
    """read! (FLOWVPM_utils.jl:493-571): append the particles of a file made
    by save (after dropping the current ones with overwrite=True), and take
    its t and nt with load_time=True.  Only X, Gamma, sigma, vol,
    circulation, static and C are loaded; every other row of the new
    records keeps what the slot held.  Without "C" in the file C is zeroed;
    without "static" the particles are not static."""
    if overwrite:
        pfield.np = 0
    h5fname = h5_fname if h5_fname.endswith(".h5") else h5_fname + ".h5"
    h5 = hdf5()
    full = os.path.join(path, h5fname)
    fid = h5.L.H5Fopen(full.encode(), _H5F_ACC_RDONLY, _H5P_DEFAULT)
    if fid < 0:
        raise HDF5Error("cannot open %s" % full)
    try:
        np_ = int(_read(h5, fid, "np"))
        if pfield.np + np_ > pfield.maxparticles:
            raise HDF5Error("The field to be read (%s) contains %d particles but max number of particles is %d."
                            % (h5_fname, np_, pfield.maxparticles))
        if load_time:
            pfield.t = float(_read(h5, fid, "t"))      # converted to the field's float type
            pfield.nt = int(_read(h5, fid, "nt"))

        def rows(name, k):
            a = _read(h5, fid, name).astype(np.float64, copy=False)
            want = (np_, k) if k > 1 else (np_,)
            if a.shape != want:
                raise HDF5Error("dataset %s has shape %s, expected %s" % (name, a.shape, want))
            return a

        cols = [(X_INDEX, rows("X", 3)), (GAMMA_INDEX, rows("Gamma", 3)), (SIGMA_INDEX, rows("sigma", 1)),
                (VOL_INDEX, rows("vol", 1)), (CIRCULATION_INDEX, rows("circulation", 1))]
        if h5.L.H5Lexists(fid, b"static", _H5P_DEFAULT) > 0:
            s = rows("static", 1)
            if not np.all((s == 0) | (s == 1)):
                raise HDF5Error("dataset static holds values other than 0 and 1 (InexactError in Bool)")
            cols.append((STATIC_INDEX, s))
        else:
            cols.append((STATIC_INDEX, np.zeros(np_)))
        if h5.L.H5Lexists(fid, b"C", _H5P_DEFAULT) > 0:
            cols.append((C_INDEX, rows("C", 3)))
        else:
            cols.append((C_INDEX, np.zeros((np_, 3))))
    finally:
        h5.ok(h5.L.H5Fclose(fid), "H5Fclose")
    a, b = pfield.np, pfield.np + np_
    for c, v in cols:
        pfield.particles[a:b, c] = v
    pfield.np = b
    return pfield
