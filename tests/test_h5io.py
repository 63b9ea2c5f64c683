"""Particle-field files (save / read! of src/FLOWVPM_utils.jl:148-328,
493-571) on the CPU: what a file holds, that read_into restores what save
wrote, the XDMF description against the HDF5 file it names, and the reading
rules of read! (overwrite, load_time, overflow, static as Bool, missing C).

The reference holds no .h5 file, so the on-disk layout is pinned by the
reference's own XDMF text (FLOWVPM_utils.jl:262-366): it declares X as an
"np 3" HDF item and sigma as "np", i.e. (np, 3) and (np,) datasets in C
order, which is how HDF5.jl stores a Julia 3 x np matrix and an np vector.
An independent reader (h5py or h5dump of the conda tree, when present)
checks the file against that.
"""
import os
import shutil
import subprocess
import xml.etree.ElementTree as ET

import numpy as np
import pytest

H5PY_PYTHON = "/opt/conda/bin/python3.9"


@pytest.fixture(scope="module")
def h5io(vpm):
    from flowvpm_amd import h5io as m
    if m.hdf5_library_path() is None:
        pytest.skip("no libhdf5 on this machine")
    m.hdf5()
    return m


def _field(vpm, n, maxp=None, les=False, seed=0):
    api = vpm.api
    pf = api.ParticleField(maxp or n + 10, SFS=api.SFS_Cd_twolevel_nobackscatter if les else api.noSFS)
    rng = np.random.default_rng(seed)
    P = rng.normal(size=(n, 46))
    P[:, 42] = rng.integers(0, 2, n)
    api.add_particles(pf, P)
    pf.nt, pf.t = 12, 0.0625
    return pf, P


SAVED = {"X": slice(0, 3), "Gamma": slice(3, 6), "sigma": 6, "circulation": 8, "vol": 7, "static": 42,
         "velocity": slice(9, 12), "velocity_gradient_x": slice(15, 18), "velocity_gradient_y": slice(18, 21),
         "velocity_gradient_z": slice(21, 24), "vorticity": slice(12, 15)}
LOADED = [0, 1, 2, 3, 4, 5, 6, 7, 8, 42, 36, 37, 38]


@pytest.mark.parametrize("les", [False, True])
def test_round_trip(vpm, h5io, tmp_path, les):
    pf, P = _field(vpm, 257, les=les)
    assert h5io.save(pf, "pfield", path=str(tmp_path)) == "pfield.12.xmf;"
    assert sorted(os.listdir(tmp_path)) == ["pfield.12.h5", "pfield.12.xmf"]
    q = vpm.api.ParticleField(400, SFS=pf.SFS)
    q.particles[:] = -7.0                        # rows read! does not load keep this
    h5io.read_into(q, "pfield.12", path=str(tmp_path))
    assert (q.np, q.nt, q.t) == (257, 12, 0.0625)
    cols = LOADED if les else LOADED[:10]
    assert np.array_equal(q.particles[:257, cols], P[:, cols])
    if not les:
        assert np.all(q.particles[:257, 36:39] == 0)          # no "C" in the file: C zeroed
    other = np.setdiff1d(np.arange(46), LOADED)
    assert np.all(q.particles[:257, other] == -7.0)
    assert np.all(q.particles[257:] == -7.0)


def test_read_appends_and_keeps_time(vpm, h5io, tmp_path):
    pf, P = _field(vpm, 40)
    h5io.save(pf, "a", path=str(tmp_path), add_num=False)
    q, Q = _field(vpm, 5, maxp=50, seed=1)
    q.nt, q.t = 3, 9.5
    h5io.read_into(q, "a.h5", path=str(tmp_path), overwrite=False, load_time=False)
    assert (q.np, q.nt, q.t) == (45, 3, 9.5)
    assert np.array_equal(q.particles[:5], Q)
    assert np.array_equal(q.particles[5:45, LOADED[:10]], P[:, LOADED[:10]])
    with pytest.raises(h5io.HDF5Error, match="max number of particles is 50"):
        h5io.read_into(q, "a", path=str(tmp_path), overwrite=False)
    assert q.np == 45


def test_file_numbering_and_time_stamp(vpm, h5io, tmp_path):
    pf, _ = _field(vpm, 3)
    assert h5io.save(pf, "f", path=str(tmp_path), num=4) == "f.4.xmf;"
    assert h5io.save(pf, "g", path=str(tmp_path), overwrite_time=pf.nt) == "g.12.xmf;"
    assert '<Time Value="12" />' in (tmp_path / "g.12.xmf").read_text()
    assert '<Time Value="0.0625" />' in (tmp_path / "f.4.xmf").read_text()
    q = vpm.api.ParticleField(3)
    h5io.read_into(q, "g.12", path=str(tmp_path))
    assert q.t == 12.0 and isinstance(q.t, float)


def test_empty_field_saves_one_stand_in(vpm, h5io, tmp_path):
    """save of an empty field writes one particle at the origin with zero
    strength and add_particle's defaults (FLOWVPM_utils.jl:157-166)."""
    pf = vpm.api.ParticleField(8, SFS=vpm.api.SFS_Cd_twolevel_nobackscatter)
    pf.nt, pf.t = 2, 0.5
    h5io.save(pf, "e", path=str(tmp_path))
    q = vpm.api.ParticleField(8)
    h5io.read_into(q, "e.2", path=str(tmp_path))
    assert q.np == 1 and q.nt == 2 and q.t == 0.5
    ref = np.zeros(46)
    ref[8] = 1.0
    assert np.array_equal(q.particles[0], ref)
    assert "C" not in (tmp_path / "e.2.xmf").read_text().split("Name=")[-1]


def test_static_must_be_boolean(vpm, h5io, tmp_path):
    pf, _ = _field(vpm, 4)
    pf.particles[2, 42] = 0.5
    h5io.save(pf, "s", path=str(tmp_path))
    with pytest.raises(h5io.HDF5Error, match="static"):
        h5io.read_into(vpm.api.ParticleField(4), "s.12", path=str(tmp_path))


def test_missing_file_and_create_path(vpm, h5io, tmp_path):
    with pytest.raises(h5io.HDF5Error, match="cannot open"):
        h5io.read_into(vpm.api.ParticleField(4), "nothing", path=str(tmp_path))
    d = tmp_path / "out"
    d.mkdir()
    (d / "old.txt").write_text("x")
    h5io.create_path(str(d), True)
    assert os.listdir(d) == []


def _xdmf_items(path):
    root = ET.parse(path).getroot()
    grid = root.find("Domain/Grid")
    items = {}
    geo = grid.find("Geometry/DataItem")
    items["geometry"] = geo
    for a in grid.findall("Attribute"):
        items[a.get("Name")] = a.find("DataItem")
    return grid, items


@pytest.mark.parametrize("les", [False, True])
def test_xdmf_names_the_datasets(vpm, h5io, tmp_path, les):
    """Every XDMF DataItem points at a dataset of the .h5 with the declared
    dimensions and the record columns of save."""
    pf, P = _field(vpm, 33, les=les)
    h5io.save(pf, "x", path=str(tmp_path))
    grid, items = _xdmf_items(tmp_path / "x.12.xmf")
    assert grid.find("Time").get("Value") == "0.0625"
    assert grid.find("Topology").get("Type") == "Polyvertex" and grid.find("Topology").get("Dimensions") == "33"
    names = ["geometry", "Gamma", "velocity", "velocity gradient x", "velocity gradient y", "velocity gradient z",
             "vorticity", "sigma", "circulation", "vol", "static"] + (["C"] if les else [])
    assert list(items) == names
    h5 = h5io.hdf5()
    fid = h5.L.H5Fopen(str(tmp_path / "x.12.h5").encode(), 0, 0)
    try:
        for name, it in items.items():
            fname, dset = it.text.split(":")
            assert fname == "x.12.h5" and it.get("Format") == "HDF" and it.get("Precision") == "8"
            dims = tuple(int(v) for v in it.get("Dimensions").split())
            a = h5io._read(h5, fid, dset)
            assert a.shape == dims, (name, a.shape, dims)
            cols = SAVED.get(dset, slice(36, 39))
            assert np.array_equal(a, P[:, cols])
    finally:
        h5.L.H5Fclose(fid)


def test_independent_reader(vpm, h5io, tmp_path):
    """h5py (conda's Python) reads the file as HDF5.jl would have written
    it: scalar int64 np and nt, scalar float64 t, float64 (np, k) and (np,)
    datasets."""
    if not os.path.exists(H5PY_PYTHON):
        pytest.skip("no conda Python with h5py")
    pf, P = _field(vpm, 21, les=True)
    h5io.save(pf, "r", path=str(tmp_path))
    np.save(tmp_path / "P.npy", P)
    code = r"""
import sys, h5py, numpy as np
f = h5py.File(sys.argv[1], "r")
P = np.load(sys.argv[2])
assert f["np"].shape == () and f["np"].dtype == np.int64 and f["np"][()] == 21
assert f["nt"].shape == () and f["nt"].dtype == np.int64 and f["nt"][()] == 12
assert f["t"].shape == () and f["t"].dtype == np.float64 and f["t"][()] == 0.0625
cols = {"X": (0, 3), "Gamma": (3, 6), "velocity": (9, 12), "vorticity": (12, 15),
        "velocity_gradient_x": (15, 18), "velocity_gradient_y": (18, 21), "velocity_gradient_z": (21, 24),
        "C": (36, 39)}
for k, (a, b) in cols.items():
    d = f[k]
    assert d.shape == (21, 3) and d.dtype == "<f8", (k, d.shape, d.dtype)
    assert np.array_equal(d[...], P[:, a:b]), k
for k, c in {"sigma": 6, "vol": 7, "circulation": 8, "static": 42}.items():
    d = f[k]
    assert d.shape == (21,) and d.dtype == "<f8", (k, d.shape)
    assert np.array_equal(d[...], P[:, c]), k
assert sorted(f.keys()) == sorted(list(cols) + ["sigma", "vol", "circulation", "static", "np", "nt", "t"])
print("ok")
"""
    env = {k: v for k, v in os.environ.items() if not k.startswith("PYTHON")}
    r = subprocess.run([H5PY_PYTHON, "-c", code, str(tmp_path / "r.12.h5"), str(tmp_path / "P.npy")],
                       capture_output=True, text=True, env=env, timeout=120)
    if r.returncode != 0 and "No module named 'h5py'" in r.stderr:
        pytest.skip("conda Python has no h5py")
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr[-2000:]


def test_reads_a_file_written_by_another_writer(vpm, h5io, tmp_path):
    """read_into takes a file that h5py wrote with HDF5.jl's layout (scalar
    np/nt/t, (np,3) and (np,) float64 datasets, no static and no C)."""
    if not os.path.exists(H5PY_PYTHON):
        pytest.skip("no conda Python with h5py")
    rng = np.random.default_rng(5)
    X, G, s = rng.normal(size=(6, 3)), rng.normal(size=(6, 3)), rng.uniform(0.1, 0.2, 6)
    np.save(tmp_path / "in.npy", np.concatenate([X, G, s[:, None]], 1))
    code = r"""
import sys, h5py, numpy as np
A = np.load(sys.argv[2])
with h5py.File(sys.argv[1], "w") as f:
    f["np"] = np.int64(6); f["nt"] = np.int64(9); f["t"] = 1.25
    f["X"] = A[:, 0:3]; f["Gamma"] = A[:, 3:6]; f["sigma"] = A[:, 6]
    f["vol"] = np.full(6, 0.5); f["circulation"] = np.full(6, 2.0)
"""
    env = {k: v for k, v in os.environ.items() if not k.startswith("PYTHON")}
    r = subprocess.run([H5PY_PYTHON, "-c", code, str(tmp_path / "w.h5"), str(tmp_path / "in.npy")],
                       capture_output=True, text=True, env=env, timeout=120)
    if r.returncode != 0 and "No module named 'h5py'" in r.stderr:
        pytest.skip("conda Python has no h5py")
    assert r.returncode == 0, r.stderr[-2000:]
    q = vpm.api.ParticleField(6)
    q.particles[:] = 3.0
    h5io.read_into(q, "w", path=str(tmp_path))
    assert (q.np, q.nt, q.t) == (6, 9, 1.25)
    assert np.array_equal(q.particles[:, 0:3], X) and np.array_equal(q.particles[:, 3:6], G)
    assert np.array_equal(q.particles[:, 6], s)
    assert np.all(q.particles[:, 7] == 0.5) and np.all(q.particles[:, 8] == 2.0)
    assert np.all(q.particles[:, 42] == 0.0) and np.all(q.particles[:, 36:39] == 0.0)


def test_run_vpm_save_cadence(vpm, h5io, tmp_path, monkeypatch):
    """run_vpm's save cadence (FLOWVPM_utils.jl:124-129): step 0, every
    nsteps_save, the last step and a break; the stepping itself is replaced
    by a stand-in here (it needs the GPU), so only the cadence is checked."""
    api = vpm.api
    pf, _ = _field(vpm, 4)
    pf.nt, pf.t = 0, 0.0
    calls = []

    def fake_step(p, dt, relax=False, resident=False):
        calls.append(relax)
        p.particles[:p.np, 0] += 1.0

    pf.integration = fake_step
    out = str(tmp_path / "run")
    os.mkdir(out)
    open(os.path.join(out, "stale"), "w").close()
    api.run_vpm(pf, 0.1, 5, save_path=out, nsteps_save=2, resident=False)
    assert sorted(os.listdir(out)) == sorted("pfield.%d.%s" % (k, e) for k in (0, 2, 4, 5) for e in ("h5", "xmf"))
    q = api.ParticleField(4)
    h5io.read_into(q, "pfield.5", path=out)
    assert np.array_equal(q.particles[:4, 0], pf.particles[:4, 0]) and q.nt == 5
    shutil.rmtree(out)
    api.run_vpm(pf, 0.1, 5, save_path=out, nsteps_save=10, resident=False, run_name="b", save_time=False,
                runtime_function=lambda p, t, dt: p.nt == 7)
    assert sorted(os.listdir(out)) == sorted("b.%d.%s" % (k, e) for k in (5, 7) for e in ("h5", "xmf"))
    assert '<Time Value="7" />' in open(os.path.join(out, "b.7.xmf")).read()
