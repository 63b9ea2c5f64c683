"""CPU-only checks of the C-ABI library and the host-side logic.

No call here launches GPU work: the library is loaded, its exports are
compared with include/flowvpm_hip.h, and only calls that fail before any
device work (argument checks, no device present) are made.
"""
import ctypes as C
import os

import numpy as np
import pytest


def test_library_exports_header(vpm):
    L = vpm.api.lib()
    names = vpm._lib.header_symbols()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), n
    assert L.vpm_version().startswith(b"flowvpm_hip")


def test_error_paths_without_device(vpm):
    L = vpm.api.lib()
    assert L.vpm_device_count(None) == vpm._lib.VPM_EINVAL
    assert b"null" in L.vpm_last_error()
    h = C.c_void_p()
    assert L.vpm_create(0, 0, C.byref(h)) == vpm._lib.VPM_EINVAL
    assert L.vpm_create(0, 10, None) == vpm._lib.VPM_EINVAL
    if vpm._lib.device_count() == 0:
        assert L.vpm_create(0, 10, C.byref(h)) == vpm._lib.VPM_EDEVICE
    assert L.vpm_destroy(None) == vpm._lib.VPM_OK
    assert L.vpm_uj_direct(None, None, 0) == vpm._lib.VPM_EINVAL
    assert L.vpm_get_np(None, None) == vpm._lib.VPM_EINVAL


def test_no_cpu_fallback(vpm, tmp_path, monkeypatch):
    """The product path fails loudly when the library is missing."""
    monkeypatch.setattr(vpm._lib, "_lib", None)
    monkeypatch.setattr(vpm._lib, "LIB_PATH", str(tmp_path / "missing.so"))
    with pytest.raises(OSError):
        vpm._lib.lib()


def test_particle_layout(vpm):
    P = vpm.fields.make_particles([[1, 2, 3]], [[4, 5, 6]], 0.5, static=[1])
    assert P.shape == (1, 46)
    assert list(P[0, 0:7]) == [1, 2, 3, 4, 5, 6, 0.5]
    assert P[0, 8] == 1.0 and P[0, 42] == 1.0
    J = np.zeros((1, 46))
    J[0, 15:24] = np.arange(9)
    Jm = vpm.fields.J_matrix(J)
    # column-major: J[3b + a] = du_a/dx_b
    assert Jm[0, 1, 0] == 1 and Jm[0, 0, 1] == 3
    W = vpm.fields.get_W(J)
    assert list(W[0]) == [5 - 7, 6 - 2, 1 - 3]


def test_particlefield_add_remove(vpm):
    api = vpm.api
    pf = api.ParticleField(3)
    for k in range(3):
        api.add_particle(pf, (k, 0, 0), (0, 0, 1), 0.1, static=(k == 1))
    with pytest.raises(api.VPMError):
        api.add_particle(pf, (9, 0, 0), (0, 0, 1), 0.1)
    assert api.get_np(pf) == 3 and pf.particles[1, 42] == 1.0
    pf.particles[2, 9:27] = 7.0
    api.remove_particle(pf, 0)            # last particle moves into slot 0
    assert api.get_np(pf) == 2
    assert pf.particles[0, 0] == 2.0 and np.all(pf.particles[0, 9:27] == 7.0)
    assert np.all(pf.particles[2, 9:27] == 0.0)
    with pytest.raises(api.VPMError):
        api.remove_particle(pf, 5)


def test_options_mirror_reference(vpm):
    api = vpm.api
    assert api.rVPM.f == 0 and api.rVPM.g == pytest.approx(0.2) and api.rVPM.h == pytest.approx(0.4)
    assert api.cVPM.g == 0
    assert api.SFS_Cd_twolevel_nobackscatter.alpha == 0.999
    assert api.SFS_Cd_twolevel_nobackscatter.force_positive
    assert api.pedrizzetti.rlxf == 0.3 and api.pedrizzetti.nsteps_relax == 1
    assert api.zeta0("gaussianerf") == pytest.approx((2 * np.pi) ** -1.5)


def test_workloads_shapes(vpm):
    w = vpm.workloads
    P0 = w.ring_cfg0()
    assert P0.shape == (8000, 46)
    np.testing.assert_allclose(P0[:, 6], 1.5 * 2 * np.pi / 200)
    # total circulation of a station: Gamma0 * 2 pi R / Nphi along t_hat
    np.testing.assert_allclose(np.linalg.norm(P0[:40, 3:6].sum(axis=0)), 2 * np.pi / 200, rtol=1e-12)
    P1 = w.random_cloud(4096)
    assert P1.shape == (4096, 46) and P1[:, 6].min() >= 0.02 and P1[:, 6].max() <= 0.04
    P2 = w.leapfrog_cfg2(Nphi=64)
    assert P2.shape == (2 * 64 * 64, 46)
    Pr = w.ring_test()
    assert Pr.shape == (100, 46)
    np.testing.assert_allclose(np.linalg.norm(Pr[:, 3:6], axis=1), 2 * np.pi / 100)


@pytest.mark.slow
def test_jet_column_size(vpm):
    P = vpm.workloads.jet_column()
    assert P.shape == (1 << 20, 46)
    r = np.hypot(P[:, 0], P[:, 1])
    assert r.max() < 0.5 + 0.01 and abs(P[:, 2]).max() < 2.01
    assert np.all(P[:, 6] == 0.015)


def test_metrics(vpm):
    R = np.array([[1.0, 0, 0], [0, 2.0, 0], [0, 0, 1e-9]])
    F = R + np.array([[1e-6, 0, 0], [0, 0, 0], [0, 0, 1e-9]])
    e = vpm.metrics.rel_err(F, R)
    assert e[0] == pytest.approx(1e-6) and e[1] == 0
    # the floor: 1e-6 * max norm
    assert e[2] == pytest.approx(1e-9 / (1e-6 * 2.0))


def test_gen_kernel_fits_reproducible(vpm):
    """kernel_fits.h is what its generator writes."""
    import subprocess
    import sys
    csrc = os.path.join(os.path.dirname(vpm.__file__), "csrc")
    out = subprocess.run([sys.executable, os.path.join(csrc, "gen_kernel_fits.py")], check=True,
                         capture_output=True, text=True).stdout
    with open(os.path.join(csrc, "kernel_fits.h")) as fh:
        assert fh.read().strip() == out.strip()


def test_stretching_helper_matches_oracle_update(vpm, orc):
    """metrics.stretching against the stretching term inside the oracle's
    update_particle_states (f = g = 0 so MM4 = 0, a = 0, dt = 1: the second
    M column becomes exactly the stretching term)."""
    rng = np.random.default_rng(3)
    P = vpm.fields.make_particles(rng.normal(size=(7, 3)), rng.normal(size=(7, 3)), np.full(7, 0.1))
    P[:, 15:24] = rng.normal(size=(7, 9))
    for transposed in (True, False):
        Q = P.copy()
        orc.update_particle_states(Q, 0.0, 1.0, 1.0, np.zeros(3), 0.0, 0.0, 1.0, transposed)
        np.testing.assert_allclose(vpm.metrics.stretching(P[:, 15:24], P[:, 3:6], transposed), Q[:, 30:33],
                                   rtol=1e-14, atol=1e-14)


def test_leaf_direct_list_shape(vpm):
    """morton_order is a permutation; the leaf list covers every leaf with
    itself, is symmetric and is grouped by target leaf (the linear-time plan
    of vpm_p2p_ranges relies on that grouping)."""
    P = vpm.workloads.random_cloud(3000, seed=4)
    Q = vpm.workloads.morton_order(P)
    assert np.array_equal(np.sort(Q[:, 0]), np.sort(P[:, 0]))
    tb, te, sb, se = vpm.workloads.leaf_direct_list(Q, leaf=64)
    pairs = set(zip(tb.tolist(), sb.tolist()))
    for a in range(0, 3000, 64):
        assert (a, a) in pairs
    assert all((s, t) in pairs for (t, s) in pairs)
    first = {}
    for k, t in enumerate(tb.tolist()):
        first.setdefault(t, k)
    assert all(tb[first[t]:first[t] + int(np.sum(tb == t))].tolist() == [t] * int(np.sum(tb == t)) for t in first)
    assert np.all(te - tb <= 64) and np.all(te > tb)


def _header_array(path, name):
    import re
    with open(path) as fh:
        text = fh.read()
    body = re.search(r"%s\[\d+\] = \{([^}]*)\}" % name, text).group(1)
    return np.array([float(v.strip().rstrip("f")) for v in body.split(",")], dtype=np.float32)


def _g_fp32(S, q, drop_erfc=False):
    """The device form of g = erf(q/sqrt2) - sqrt(2/pi) q e^{-q^2/2} above the
    series: 1 - E (S(1/q)/q + sqrt(2/pi) q), E = e^{-q^2/2}, evaluated in
    float32 (Horner in float32, each fma one rounding)."""
    f32 = np.float32
    q = q.astype(f32)
    t = (q * q).astype(f32)
    E = np.exp((f32(-0.5) * t).astype(f32)).astype(f32)
    w = (f32(1) / np.sqrt(t)).astype(f32)
    acc = np.full(q.shape, S[-1], dtype=f32)
    for c in S[-2::-1]:
        acc = (acc.astype(np.float64) * w + c).astype(f32)
    c2q = (f32(np.sqrt(2 / np.pi)) * q).astype(f32)
    inner = (acc.astype(np.float64) * w + c2q).astype(f32)
    if drop_erfc:
        inner = c2q
    g = (1.0 - E.astype(np.float64) * inner).astype(f32)
    hm = (g.astype(np.float64) - (f32(1 / 3) * t).astype(np.float64) * (c2q * E).astype(np.float64)).astype(f32)
    return g, hm


def test_erfc_fit_gives_g_to_fp32(vpm):
    """kErfcS (kernel_fits.h) in the device's closed form of the Gaussian-erf
    near field reproduces g and g - q g'/3 (the numerators of F' and H'') to
    FP32 accuracy against 30-digit mpmath, for every q from the series'
    threshold (1.5) past the far cut (7), where the fit is only extrapolated
    and E < 2.3e-11.  Dropping the erfc term fails the check."""
    import mpmath as mp
    mp.mp.dps = 30
    csrc = os.path.join(os.path.dirname(vpm.__file__), "csrc")
    S = _header_array(os.path.join(csrc, "kernel_fits.h"), "kErfcS")
    q = np.concatenate([np.linspace(1.5, 7.0, 1500), np.linspace(7.0, 12.0, 200)])
    g32, hm32 = _g_fp32(S, q)
    c2 = mp.sqrt(2 / mp.pi)
    g_ref = np.array([float(mp.erf(mp.mpf(float(x)) / mp.sqrt(2)) - c2 * float(x) * mp.exp(-mp.mpf(float(x)) ** 2 / 2))
                      for x in q.astype(np.float32)])
    hm_ref = np.array([float(gr - c2 * mp.mpf(float(x)) ** 3 * mp.exp(-mp.mpf(float(x)) ** 2 / 2) / 3)
                       for gr, x in zip(g_ref, q.astype(np.float32))])
    eg = np.abs(g32 - g_ref) / np.abs(g_ref)
    eh = np.abs(hm32 - hm_ref) / np.abs(hm_ref)
    assert eg.max() < 4e-7, eg.max()
    assert eh.max() < 8e-7, eh.max()
    gbad, _ = _g_fp32(S, q, drop_erfc=True)
    assert (np.abs(gbad - g_ref) / np.abs(g_ref)).max() > 1e-3
