"""The solve server (chol_serve) on the GPU: continuous batching of SPD
solve requests of mixed sizes with re-solves, HBM-budget admission and
eviction, and the same result for a request alone and in the mix.

Each request must reach normwise backward error <= 1e-14 (the FP64 bar)
and agree with an FP64 solve to cond(A) * 1e-14; alone versus in the mix
the solutions must agree to 1e-13 relative (the arithmetic of a request
does not depend on what else runs, so they are expected to be equal)."""
import os
import tempfile

import numpy as np
import pytest

PERF_MODEL_DIR = os.path.join(tempfile.gettempdir(), "starpu_amd_perfmodels")

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def env():
    from starpu_amd import kernels as K
    from starpu_amd import runtime as R
    from starpu_amd import serving as S
    if K.device_count() < 1:
        pytest.fail("no GPU visible to the HIP runtime")
    R.init(STARPU_NHIP=1, STARPU_NCPU=0, STARPU_SCHED="dmda", STARPU_PERF_MODEL_DIR=PERF_MODEL_DIR)
    R.set_cpu_kernels()
    yield K, R, S
    R.shutdown()


def check_solutions(K, inp, tr, reqs, idx=None):
    for i in (idx if idx is not None else range(len(tr.n))):
        q = reqs[i]
        assert q.status == 0, (i, q.status)
        assert q.berr <= 1e-14, (i, q.berr, q.iterations)
        A = inp.A[tr.matrix[i]].download()
        b = inp.B[i].download()
        x = inp.X[i].download()
        ref = np.linalg.solve(A, b)
        assert np.linalg.norm(x - ref) / np.linalg.norm(ref) < 10 * np.linalg.cond(A) * 1e-14


def test_trace_with_resolves_and_overlap(env):
    K, R, S = env
    tr = S.make_trace(14, qps=400.0, seed=3, nmin=256, nmax=2048, resolve_frac=0.35)
    print("trace n:", tr.n, "matrix:", tr.matrix, flush=True)
    inp = S.Inputs(tr, shift=0.01, seed=5)
    reqs = inp.requests(tr)
    st = R.serve(S.config(budget_gb=20, slots=4, nb=256), reqs)
    check_solutions(K, inp, tr, reqs)
    n_distinct = len(set(tr.matrix))
    assert st.factorizations == n_distinct           # every re-solve reused a resident factor
    assert st.cache_hits == len(tr.n) - n_distinct
    assert st.max_inflight >= 2                       # requests overlapped in service
    for i, q in enumerate(reqs):
        first = tr.matrix.index(tr.matrix[i]) == i
        assert q.factored == (1 if first else 0)
        assert q.finish_us >= q.start_us >= q.arrival_us - 1.0
    inp.free()


def test_same_result_alone_and_in_the_mix(env):
    K, R, S = env
    tr = S.make_trace(10, qps=0.0, seed=9, nmin=512, nmax=3072, resolve_frac=0.3)
    inp = S.Inputs(tr, shift=0.01, seed=11)
    # alone: one request in service at a time
    reqs = inp.requests(tr, alone=True)
    R.serve(S.config(budget_gb=20, slots=1, nb=512), reqs)
    alone = [x.download() for x in inp.X]
    # in the mix: all arrive together, four in service
    for x in inp.X:
        K.check(K.lib().sa_dlaset(None, x.m, x.n, 0.0, 0.0, x.ptr, x.ld))
    reqs = inp.requests(tr, alone=True)
    st = R.serve(S.config(budget_gb=20, slots=4, nb=512), reqs)
    assert st.max_inflight >= 2
    check_solutions(K, inp, tr, reqs)
    for i, x in enumerate(inp.X):
        mixed = x.download()
        assert np.linalg.norm(mixed - alone[i]) / np.linalg.norm(alone[i]) <= 1e-13, i
    inp.free()


def test_budget_admission_evicts_and_waits(env):
    K, R, S = env
    # six distinct 2048 matrices, then re-solves of the first ones: with a
    # budget for about two factors, requests wait for memory and resident
    # factors are evicted, yet every answer is right
    tr = S.Trace(n=[2048] * 8, nrhs=[4] * 8, matrix=[0, 1, 2, 3, 4, 5, 0, 1],
                 arrival_us=[0.0] * 8, matrix_n=[2048] * 6)
    inp = S.Inputs(tr, shift=0.01, seed=21)
    reqs = inp.requests(tr)
    lib = R.chol()
    lib.chol_mixed_factor_bytes.restype = __import__("ctypes").c_size_t
    fbytes = lib.chol_mixed_factor_bytes(2048, 4)
    lib.chol_mixed_refine_worksize.restype = __import__("ctypes").c_size_t
    wbytes = lib.chol_mixed_refine_worksize(2048, 4, 4)
    budget = 2.5 * (fbytes + wbytes)
    cfg = S.config(budget_gb=budget / 1e9, slots=4, nb=512)
    st = R.serve(cfg, reqs)
    check_solutions(K, inp, tr, reqs)
    assert st.peak_bytes <= budget
    assert st.memory_waits > 0
    assert st.evictions > 0
    # the re-solves of matrices 0 and 1 found their factors evicted
    assert st.factorizations == 8
    inp.free()


def test_request_larger_than_budget_is_rejected(env):
    K, R, S = env
    tr = S.Trace(n=[1024, 256], nrhs=[2, 2], matrix=[0, 1], arrival_us=[0.0, 0.0], matrix_n=[1024, 256])
    inp = S.Inputs(tr, shift=0.01, seed=31)
    reqs = inp.requests(tr)
    st = R.serve(S.config(budget_gb=2e-3, slots=2, nb=256), reqs)    # 2 MB: the 256 fits, the 1024 cannot
    assert reqs[0].status == -12                                       # -ENOMEM
    assert reqs[1].status == 0 and reqs[1].berr <= 1e-14
    assert st.factorizations == 1
    inp.free()
