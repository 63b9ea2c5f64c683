"""vpm_uj_direct_noalloc: vpm_uj_direct with its scratch in the library's
static per-GPU array instead of hipMalloc'd field buffers.

Compared with vpm_uj_direct on the same field, two calls in a row, the second
with fewer particles: bit-identical wherever both take the same source splits
(small fields); for a field whose splits do not fit the array (fewer splits,
another summation order) both stay within the 1e-5 F32C bound of the FP64
path.  Beyond the array's capacity the call fails with VPM_ERANGE before any
device work.
"""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)


def _uj(vpm, f, P, kernel, precision, noalloc, flags=None):
    api = vpm.api
    f.set_particles(P)
    cfg = api.make_config(kernel, precision, True)
    fl = api._lib.UJ_RESET if flags is None else flags
    fn = api.lib().vpm_uj_direct_noalloc if noalloc else api.lib().vpm_uj_direct
    rc = fn(f.handle, C.byref(cfg), fl)
    if rc != 0:
        return rc, None
    out = np.zeros_like(P)
    f.get_particles(out, rows=(9, 24))
    return 0, out[:, 9:24]


@pytest.mark.parametrize("kernel", ["gaussianerf", "winckelmans"])
@pytest.mark.parametrize("precision", ["f32c", "f64"])
def test_noalloc_equals_uj_direct_two_calls(vpm, orc, kernel, precision):
    api = vpm.api
    f = api.Field(6000, 0)
    for n, seed in ((6000, 90), (1500, 91)):       # the second call is smaller
        _say("noalloc", kernel, precision, "n", n)
        P = vpm.workloads.random_cloud(n, seed=seed)
        rc, ref = _uj(vpm, f, P, kernel, precision, False)
        assert rc == 0
        rc, got = _uj(vpm, f, P, kernel, precision, True)
        assert rc == 0, api.lib().vpm_last_error()
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        if n <= 1500:
            R = P.copy()
            orc.uj_direct(R, kernel)
            tol = 1e-5 if precision == "f32c" else 1e-12
            for sl in (slice(0, 3), slice(6, 15)):
                assert vpm.metrics.err_stats(got[:, sl], R[:, 9:24][:, sl])["max"] < tol
    f.close()


def test_noalloc_fewer_splits_and_limits(vpm):
    """A jet column of 200,000 particles: vpm_uj_direct's 42 source splits
    (0.94 GB of partial sums) do not fit the array, so the call takes fewer;
    then 100,000 on the same field.  Both calls stay within 1e-5 of the FP64
    path, as vpm_uj_direct does.  400,000 particles exceed the array: VPM_ERANGE with
    the sizes in the message.  VPM_UJ_SFS is refused."""
    api = vpm.api
    f = api.Field(400000, 0)
    for n in (200000, 100000):
        _say("noalloc fewer splits n", n)
        P = vpm.workloads.jet_column(n, seed=92)
        rc, got = _uj(vpm, f, P, "gaussianerf", "f32c", True)
        assert rc == 0, api.lib().vpm_last_error()
        rc, ref = _uj(vpm, f, P, "gaussianerf", "f32c", False)
        rc64, r64 = _uj(vpm, f, P, "gaussianerf", "f64", False)
        assert rc == 0 and rc64 == 0
        for sl in (slice(0, 3), slice(6, 15)):
            eg = vpm.metrics.err_stats(got[:, sl], r64[:, sl])["max"]
            er = vpm.metrics.err_stats(ref[:, sl], r64[:, sl])["max"]
            _say("  noalloc %.2e  uj_direct %.2e" % (eg, er))
            assert eg < 1e-5 and er < 1e-5
    P = vpm.workloads.random_cloud(400000, seed=93)
    rc, _ = _uj(vpm, f, P, "gaussianerf", "f32c", True)
    assert rc == api._lib.VPM_ERANGE
    msg = api.lib().vpm_last_error()
    assert b"400000 particles need" in msg and b"byte maximum" in msg, msg
    P = P[:1000].copy()
    rc, _ = _uj(vpm, f, P, "gaussianerf", "f32c", True,
                flags=api._lib.UJ_RESET | api._lib.UJ_SFS)
    assert rc == api._lib.VPM_EINVAL
    f.close()
