"""Workload driver of the solve server (chol_serve): a seeded trace of SPD
solve requests, the matrices and right-hand sides generated on the GPU
before the trace starts (they stand for the clients' inputs), and the
latency / throughput summary.

Trace: Poisson arrivals at a fixed rate; N log-uniform in [nmin, nmax]
rounded to 256; nrhs uniform in 1..32; a fraction of requests re-solves a
matrix an earlier request used, with new right-hand sides;
A = B B^T / N + shift I with B i.i.d. N(0,1).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from . import kernels as K
from . import runtime as R


@dataclass
class Trace:
    n: list[int]
    nrhs: list[int]
    matrix: list[int]
    arrival_us: list[float]
    matrix_n: list[int] = field(default_factory=list)


def make_trace(nreq: int, qps: float, seed: int, nmin: int = 256, nmax: int = 16384,
               resolve_frac: float = 0.3, max_nrhs: int = 32) -> Trace:
    rng = np.random.default_rng(seed)
    # unit-rate gaps are drawn whatever qps is, so that traces of one seed
    # hold the same requests at every arrival rate
    unit = rng.exponential(1.0, nreq)
    gaps = unit / qps if qps > 0 else np.zeros(nreq)
    arrival = np.cumsum(gaps) * 1e6
    arrival -= arrival[0]
    n, nrhs, matrix, matrix_n = [], [], [], []
    for i in range(nreq):
        if matrix_n and rng.random() < resolve_frac:
            m = int(rng.integers(0, len(matrix_n)))
        else:
            v = math.exp(rng.uniform(math.log(nmin), math.log(nmax)))
            matrix_n.append(int(min(nmax, max(nmin, 256 * round(v / 256)))))
            m = len(matrix_n) - 1
        matrix.append(m)
        n.append(matrix_n[m])
        nrhs.append(int(rng.integers(1, max_nrhs + 1)))
    return Trace(n, nrhs, matrix, [float(a) for a in arrival], matrix_n)


class Inputs:
    """Device-resident matrices (one per distinct matrix of the trace),
    right-hand sides and solution buffers of every request.

    With ``share_by_size`` the distinct matrices of one size share one
    device array (generated from a seed of that size): the server still sees
    distinct matrices (its factor cache is keyed by the request's matrix
    index), while the inputs of a long trace fit in HBM -- N is a multiple
    of 256, so there are at most 64 sizes in [256, 16384], about 47 GB at
    most, against about 180 GB for 700 separate matrices."""

    def __init__(self, tr: Trace, shift: float, seed: int, share_by_size: bool = False):
        L = K.lib()
        self.A = []
        self._own = []
        by_size = {}
        for m, nm in enumerate(tr.matrix_n):
            if share_by_size and nm in by_size:
                self.A.append(by_size[nm])
                continue
            A = K.DeviceArray(nm, nm)
            Bm = K.DeviceArray(nm, nm)
            gen_seed = seed + 7919 * (nm if share_by_size else m)
            K.check(L.sa_fill_random(None, nm, nm, Bm.ptr, nm, gen_seed, 1), "fill")
            K.dsyrk_ln(nm, nm, 1.0 / nm, Bm.ptr, nm, 0.0, A.ptr, nm)
            K.check(L.sa_symmetrize_lower(None, nm, A.ptr, nm), "symmetrize")
            K.check(L.sa_add_diagonal(None, nm, A.ptr, nm, shift), "diag")
            K.sync()
            Bm.free()
            self.A.append(A)
            self._own.append(A)
            by_size[nm] = A
        self.B, self.X = [], []
        for i, (n, r) in enumerate(zip(tr.n, tr.nrhs)):
            b = K.DeviceArray(n, r)
            K.check(L.sa_fill_random(None, n, r, b.ptr, n, seed + 104729 * (i + 1), 1), "fill b")
            self.B.append(b)
            self.X.append(K.DeviceArray(n, r))
        K.sync()

    def requests(self, tr: Trace, alone: bool = False):
        reqs = (R.ServeRequest * len(tr.n))()
        for i in range(len(tr.n)):
            q = reqs[i]
            q.n, q.nrhs, q.matrix = tr.n[i], tr.nrhs[i], tr.matrix[i]
            q.arrival_us = 0.0 if alone else tr.arrival_us[i]
            A = self.A[tr.matrix[i]]
            q.A, q.lda = A.ptr, A.ld
            q.B, q.ldb = self.B[i].ptr, self.B[i].ld
            q.X, q.ldx = self.X[i].ptr, self.X[i].ld
        return reqs

    def free(self):
        for a in self._own + self.B + self.X:
            a.free()


def config(budget_gb: float = 200.0, slots: int = 4, nb: int = 512, prec: str = "fp32", tol: float = 1e-14,
           maxiter: int = 40) -> R.ServeConfig:
    cfg = R.ServeConfig()
    cfg.budget_bytes = budget_gb * 1e9
    cfg.slots, cfg.nb, cfg.prec = slots, nb, R.PREC[prec]
    cfg.maxiter, cfg.prio, cfg.tol = maxiter, 1, tol
    return cfg


FP32_PEAK_TFLOPS = 157.3   # MI355X dense FP32 matrix peak (AMD public spec)


def factor_share(reqs, prec: str = "fp32") -> dict:
    """Roofline of the factorisations: their flops (FLOPS_POTRF(N) each)
    over the wall time during which at least one was running (the union of
    their intervals), against the matrix peak of the update precision."""
    iv = sorted((q.factor_start_us, q.factor_end_us) for q in reqs if q.factored and q.factor_end_us > 0)
    busy, cur_s, cur_e = 0.0, None, None
    for a, b in iv:
        if cur_e is None or a > cur_e:
            if cur_e is not None:
                busy += cur_e - cur_s
            cur_s, cur_e = a, b
        else:
            cur_e = max(cur_e, b)
    if cur_e is not None:
        busy += cur_e - cur_s
    flops = sum(q.n ** 3 / 3.0 + q.n ** 2 / 2.0 + q.n / 6.0 for q in reqs if q.factored)
    peak = FP32_PEAK_TFLOPS if prec == "fp32" else 2516.6 / 3 if prec == "bf16x3" else 2516.6
    ach = flops / (busy * 1e-6) / 1e12 if busy > 0 else 0.0
    return {"bound": "mfma", "achieved": round(ach, 3), "peak": round(peak, 1), "unit": "TFLOP/s",
            "frac": round(ach / peak, 4), "factor_busy_s": round(busy / 1e6, 3),
            "factorizations": len(iv),
            "note": "factor flops (N^3/3 each) over the union of the factorisations' intervals; "
                    f"peak: the {prec} update precision's dense MFMA peak"}


def summarise(reqs, st: R.ServeStats, tr: Trace) -> dict:
    lat = np.array([(q.finish_us - q.arrival_us) / 1e3 for q in reqs])
    svc = np.array([(q.finish_us - q.start_us) / 1e3 for q in reqs])
    wait = np.array([(q.start_us - q.arrival_us) / 1e3 for q in reqs])
    ok = [q.status == 0 for q in reqs]
    span_s = st.wall_us / 1e6
    solves = int(sum(ok))
    flops = sum((q.n ** 3 / 3.0 if q.factored else 0.0) + 2.0 * q.n * q.n * q.nrhs * max(1, q.iterations)
                for q in reqs)
    return {
        "requests": len(reqs), "ok": solves,
        "p50_ms": round(float(np.percentile(lat, 50)), 3), "p99_ms": round(float(np.percentile(lat, 99)), 3),
        "mean_ms": round(float(lat.mean()), 3), "max_ms": round(float(lat.max()), 3),
        "p50_queue_ms": round(float(np.percentile(wait, 50)), 3),
        "p99_queue_ms": round(float(np.percentile(wait, 99)), 3),
        "p50_service_ms": round(float(np.percentile(svc, 50)), 3),
        "solves_per_s": round(solves / span_s, 2), "wall_s": round(span_s, 3),
        "tflops": round(flops / span_s / 1e12, 3),
        "max_backward_error": float(f"{max(q.berr for q in reqs):.3e}"),
        "factorizations": st.factorizations, "factor_cache_hits": st.cache_hits, "evictions": st.evictions,
        "memory_waits": st.memory_waits, "max_in_service": st.max_inflight,
        "peak_gb": round(st.peak_bytes / 1e9, 3),
    }
