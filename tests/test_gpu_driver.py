"""Host-side argument checks of the driver entry points (vpm_add_particles,
vpm_remove_particle, vpm_zeta_ranges, vpm_core_spreading_reset) and their
error codes (particlefield.jl:171-173, 402-407 raise the same conditions).
Every case is rejected before any kernel launch."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_add_and_remove_particle_errors(vpm):
    api = vpm.api
    P = vpm.workloads.random_cloud(10, seed=1)
    f = api.Field(12, 0)
    f.set_particles(P)
    with pytest.raises(api.VPMError) as ei:
        f.add_particles(vpm.workloads.random_cloud(3, seed=2))      # 13 > 12
    assert ei.value.status == -2
    bad = vpm.workloads.random_cloud(1, seed=3)
    bad[0, 6] = 0.0
    with pytest.raises(api.VPMError) as ei:
        f.add_particles(bad)                                        # sigma <= 0
    assert ei.value.status == -1
    assert f.np == 10
    for i in (-1, 10):
        with pytest.raises(api.VPMError) as ei:
            f.remove_particle(i)
        assert ei.value.status == -1
    f.add_particles(vpm.workloads.random_cloud(2, seed=4))
    assert f.np == 12
    f.remove_particle(0)                                            # the last one moves into slot 0
    G = np.zeros((11, 46))
    f.get_particles(G)
    Q = np.vstack([P, vpm.workloads.random_cloud(2, seed=4)])
    assert np.array_equal(G[0], Q[11]) and np.array_equal(G[1:10], Q[1:10]) and np.array_equal(G[10], Q[10])
    f.close()


def test_zeta_ranges_and_reset_argument_errors(vpm):
    api = vpm.api
    n = 50
    P = vpm.workloads.random_cloud(n, seed=5)
    with pytest.raises(api.VPMError) as ei:
        api.zeta_ranges(P.copy(), [0], [n + 1], [0], [n], kernel="gaussianerf", precision="f32c")
    assert ei.value.status == -1
    with pytest.raises(api.VPMError) as ei:
        api.zeta_ranges(P.copy(), [0, 1], [n], [0], [n], kernel="gaussianerf", precision="f32c")
    assert ei.value.status == -1
    A = P.copy()
    api.zeta_ranges(A, [], [], [], [], kernel="gaussianerf", precision="f32c")   # empty list: no-op
    assert np.array_equal(A, P)
    f = api.Field(n, 0)
    f.set_particles(P)
    for sg0 in (0.0, -1.0, float("nan")):
        with pytest.raises(api.VPMError) as ei:
            f.core_spreading_reset(sg0, "gaussianerf", "f64")
        assert ei.value.status == -1
    with pytest.raises(api.VPMError) as ei:
        f.core_spreading_reset(0.05, "gaussianerf", "f64", itmax=-1)
    assert ei.value.status == -1
    G = np.zeros_like(P)
    f.get_particles(G)
    assert np.array_equal(G, P)                                     # nothing ran
    # itmax = 0: target and sigma are set, Gamma = target * vol, no iteration
    it, res = f.core_spreading_reset(0.05, "gaussianerf", "f64", itmax=0)
    f.get_particles(G)
    assert it == 0 and np.all(G[:, 6] == 0.05)
    np.testing.assert_array_equal(G[:, 3:6], G[:, 33:36] * G[:, 7:8])
    f.close()


def test_first_upload_survives_the_zero_fill(vpm):
    """vpm_create zeroes the record buffer on the field's own stream and
    waits for it, so an upload right after creation is never overwritten by
    the fill (the fill on the null stream was not ordered with the field's
    non-blocking stream).  Large capacities make the fill the longer
    operation."""
    api = vpm.api
    rng = np.random.default_rng(77)
    P = vpm.fields.make_particles(rng.normal(size=(1000, 3)), rng.normal(size=(1000, 3)), 0.1)
    for cap in (1 << 20, 1 << 22, 3000):
        f = api.Field(cap, 0)
        f.set_particles(P)
        G = np.zeros_like(P)
        f.get_particles(G)
        f.close()
        assert np.array_equal(G, P), cap


@pytest.mark.parametrize("entry", ["p2p", "estr", "zeta"])
def test_range_lists_checked_before_any_work(vpm, entry):
    """Each pair of a list is checked (0 <= begin <= end <= np) on both plan
    paths (leaf form: equal or disjoint target ranges listed together;
    general form: overlapping target ranges); a bad pair anywhere fails the
    call with VPM_EINVAL naming it, and nothing is computed."""
    api = vpm.api
    n = 300
    P = vpm.workloads.random_cloud(n, seed=12)
    P[:, 9:24] = 0.25
    P[:, 39:42] = 0.5
    fn = {"p2p": lambda A, *r: vpm.p2p_ranges(A, *r, kernel="gaussianerf", precision="f32c"),
          "estr": lambda A, *r: api.estr_ranges(A, *r, kernel="gaussianerf", precision="f32c"),
          "zeta": lambda A, *r: api.zeta_ranges(A, *r, kernel="gaussianerf", precision="f32c")}[entry]
    leaf = ([0, 0, 100, 100], [100, 100, 200, 200], [0, 100, 0, 100], [100, 200, 100, 200])
    general = ([0, 50, 100], [150, 250, 300], [0, 100, 200], [100, 200, 300])
    for tb, te, sb, se in (leaf, general):
        A = P.copy()
        fn(A, tb, te, sb, se)          # the valid list runs
        for k, col, bad in ((len(tb) - 1, 3, n + 1), (1, 2, -1), (len(tb) - 1, 1, n + 5), (1, 0, -3)):
            lists = [list(tb), list(te), list(sb), list(se)]
            lists[col][k] = bad
            A = P.copy()
            with pytest.raises(vpm.VPMError) as ei:
                fn(A, *lists)
            assert ei.value.status == -1 and ("range pair %d" % k) in str(ei.value), str(ei.value)
            assert np.array_equal(A, P)
            G = np.zeros_like(P)
            api._scratch_field(n, 0).get_particles(G)       # the device copy the call was given
            assert np.array_equal(G, P)


def _ranges_call(vpm, f, tb, te, sb, se, kernel="gaussianerf"):
    import ctypes as C
    arrs = [np.ascontiguousarray(a, dtype=np.int64) for a in (tb, te, sb, se)]
    cfg = vpm.api.make_config(kernel, "f32c", True)
    ptr = [a.ctypes.data_as(vpm._lib._i64p) for a in arrs]
    vpm._lib.check(vpm._lib.lib().vpm_p2p_ranges(f.handle, C.byref(cfg), *ptr, len(arrs[0])))


def test_concurrent_calls(vpm):
    """Threads calling vpm_p2p_ranges on one field at once, each with its own
    target leaves (as FastMultipole's threaded direct! does), give the same
    records bit for bit as one call over the whole list: each call holds the
    field's lock, and a target's pairs are summed in list order within one
    call.  Meanwhile other threads run whole U/J evaluations on fields of
    their own, which match a sequential run bit for bit."""
    import threading
    api = vpm.api
    P = vpm.workloads.morton_order(vpm.workloads.random_cloud(4096, seed=21))
    n = P.shape[0]
    tb, te, sb, se = vpm.workloads.leaf_direct_list(P, 64)
    ref = api.Field(n, 0)
    ref.set_particles(P)
    _ranges_call(vpm, ref, tb, te, sb, se)
    R = np.zeros_like(P)
    ref.get_particles(R)
    ref.close()
    others = [vpm.workloads.random_cloud(3000 + 500 * k, seed=30 + k) for k in range(3)]
    seq = []
    for Q in others:
        g = api.Field(Q.shape[0], 0)
        g.set_particles(Q)
        g.uj_direct("gaussianerf", "f32c")
        S = np.zeros_like(Q)
        g.get_particles(S)
        g.close()
        seq.append(S)

    f = api.Field(n, 0)
    f.set_particles(P)
    leaf = tb // 64
    errors = []
    outs = [None] * len(others)

    def leaves(k):
        try:
            m = leaf % 4 == k
            _ranges_call(vpm, f, tb[m], te[m], sb[m], se[m])
        except Exception as e:          # pragma: no cover - reported below
            errors.append(e)

    def whole(k):
        try:
            Q = others[k]
            g = api.Field(Q.shape[0], 0)
            g.set_particles(Q)
            g.uj_direct("gaussianerf", "f32c")
            S = np.zeros_like(Q)
            g.get_particles(S)
            g.close()
            outs[k] = S
        except Exception as e:          # pragma: no cover
            errors.append(e)

    threads = [threading.Thread(target=leaves, args=(k,)) for k in range(4)]
    threads += [threading.Thread(target=whole, args=(k,)) for k in range(len(others))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in threads)
    assert not errors, errors
    G = np.zeros_like(P)
    f.get_particles(G)
    f.close()
    assert np.array_equal(G, R)
    for S, Z in zip(outs, seq):
        assert np.array_equal(S, Z)
