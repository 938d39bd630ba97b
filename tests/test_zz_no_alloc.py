"""The allocation-free tile factorisation (sa_dpotrf_tile_noalloc) against
the entry point it varies (sa_dpotrf_tile with a caller-provided
workspace), called through the C ABI: the same kernels, so the factor, both
inverses and the info flag must be bit-identical.  Two calls in a row, the
second on a smaller tile (the shared scratch keeps the first call's
contents, which must not leak into the second); a not-SPD tile; the size
limit."""
import numpy as np
import pytest

import oracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def K():
    from starpu_amd import kernels
    if kernels.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    return kernels


def spd(n, seed):
    rng = np.random.default_rng(seed)
    B = rng.uniform(-1, 1, (n, n))
    return np.asfortranarray(B @ B.T + n * np.eye(n))


def with_workspace(K, A0):
    n = A0.shape[0]
    dA, dLi, dW = K.DeviceArray.from_host(A0), K.DeviceArray(n, n), K.DeviceArray(n, n)
    info = K.DeviceInt(-7)
    wb = K.lib().sa_dpotrf_tile_worksize(n)
    ws = K.DeviceArray(wb, 1, dtype=np.uint8)
    K.check(K.lib().sa_dpotrf_tile(None, n, dA.ptr, n, dLi.ptr, n, dW.ptr, n, info.ptr, ws.ptr, wb), "tile")
    K.sync()
    return dA.download(), dLi.download(), dW.download(), info.get()


def no_alloc(K, A0):
    n = A0.shape[0]
    dA, dLi, dW = K.DeviceArray.from_host(A0), K.DeviceArray(n, n), K.DeviceArray(n, n)
    info = K.DeviceInt(-7)
    K.check(K.lib().sa_dpotrf_tile_noalloc(None, n, dA.ptr, n, dLi.ptr, n, dW.ptr, n, info.ptr), "noalloc")
    K.sync()
    return dA.download(), dLi.download(), dW.download(), info.get()


def same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a[:3], b[:3])) and a[3] == b[3]


def test_two_calls_second_smaller_match_the_workspace_entry_point(K):
    assert K.lib().sa_dpotrf_tile_noalloc_maxn() == 2048
    for n, seed in ((2048, 1), (700, 2), (64, 3)):
        print(f"noalloc tile factor n={n}", flush=True)
        A0 = spd(n, seed)
        ref = with_workspace(K, A0)
        got = no_alloc(K, A0)
        assert got[3] == 0
        assert same(got, ref), n
        # and the factor is the oracle's to the FP64 kernel bar
        L, _ = oracle.dpotrf_l(A0)
        assert oracle.rel_diff_lower(np.tril(got[0]), np.tril(L)) < 1e-13


def test_not_spd_reports_the_same_column(K):
    A0 = spd(500, 4)
    A0[321, 321] = -1e3
    ref = with_workspace(K, A0)
    got = no_alloc(K, A0)
    assert got[3] == ref[3] == 322


def test_tiles_above_the_limit_are_rejected(K):
    n = 2049
    d = K.DeviceArray(n, n)
    info = K.DeviceInt(-7)
    rc = K.lib().sa_dpotrf_tile_noalloc(None, n, d.ptr, n, d.ptr, n, d.ptr, n, info.ptr)
    assert rc == -75            # -EOVERFLOW
    K.sync()
    assert info.get() == -7     # nothing ran
