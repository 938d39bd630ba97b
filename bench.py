#!/usr/bin/env python3
"""Benchmark: tiled FP64 Cholesky factorisation, N=32768, tile 1024
(32 x 32 tiles), on one MI355X through the StarPU-style runtime.

Workload (the GPU configuration of the reference's Cholesky benchmark):
A = B*B^T/N + I with B i.i.d. N(0,1) from a seed, generated on the GPU,
factored by the POTRF/TRSM/SYRK/GEMM task DAG of
examples/cholesky/cholesky_implicit.c under dmda with HIP workers.  A is
resident in HBM before the timed region; a step is one factorisation,
timed from the first task submission to starpu_task_wait_for_all
(cholesky_implicit.c:233,325-326), after restoring A from a pristine copy
(untimed).  value = FLOPS_DPOTRF(N) / mean step time.

Prints one JSON line (see the keys below).  --gpus above 1 is not
supported: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# Eight HIP streams per GPU (STARPU_NWORKER_PER_HIP) need as many hardware
# queues to run concurrently; HIP reads this once, at its initialisation.
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

FP64_PEAK_TFLOPS = 78.6     # MI355X dense FP64 matrix peak (AMD public spec)


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def make_spd(K, n: int, seed: int):
    """A = B B^T / n + I on the device (B ~ N(0,1), counter-based, seeded)."""
    L = K.lib()
    A = K.DeviceArray(n, n)
    B = K.DeviceArray(n, n)
    K.check(L.sa_fill_random(None, n, n, B.ptr, n, seed, 1), "fill")
    K.dsyrk_ln(n, n, 1.0 / n, B.ptr, n, 0.0, A.ptr, n)
    K.check(L.sa_symmetrize_lower(None, n, A.ptr, n), "symmetrize")
    K.check(L.sa_add_diagonal(None, n, A.ptr, n, 1.0), "diag")
    K.sync()
    B.free()
    return A


def cpu_baseline(A_host_lead: np.ndarray, nb: int, cores: int) -> dict:
    """The same DAG on `cores` CPU workers under dmda, each running the
    oracle's FP64 CPU codelets, on a leading principal submatrix."""
    import oracle
    from oracle.codelets import addresses
    from starpu_amd import runtime as R
    n = A_host_lead.shape[0]
    R.init(STARPU_NHIP=0, STARPU_NCPU=cores, STARPU_SCHED="dmda")
    R.set_cpu_kernels(**addresses())
    A = np.asfortranarray(A_host_lead.copy())
    st = R.factor(A.ctypes.data, n, n, max(1, n // nb))
    R.set_cpu_kernels()
    R.shutdown()
    # sanity: the CPU run is a correct factorisation
    be = oracle.backward_error_llt(A_host_lead[:512, :512], np.tril(A[:512, :512]))
    assert st.info == 0 and be < 1e-13, (st.info, be)
    sec = st.elapsed_us / 1e6
    return {"value": round(st.flops / sec / 1e12, 6), "unit": "TFLOP/s", "cores": cores,
            "kind": "port",
            "sample": f"leading {n}x{n} principal submatrix of the same A, tile {nb} "
                      f"({n // nb}x{n // nb} tiles), one factorisation on {cores} CPU workers "
                      f"(dmda, oracle FP64 CPU codelets): {sec:.2f} s"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=32768)
    ap.add_argument("--nb", type=int, default=1024)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--workers-per-gpu", type=int, default=int(os.environ.get("STARPU_NWORKER_PER_HIP", "8")))
    ap.add_argument("--cpu-n", type=int, default=8192, help="CPU-baseline sample size (0: skip)")
    ap.add_argument("--cpu-cores", type=int, default=0)
    ap.add_argument("--trace", default="", help="write a Chrome trace of the last step here")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported by this build "
              f"(multi-GPU sharding is described in DESIGN.md but not built)", file=sys.stderr)
        return 2
    if os.environ.get("WORLD_SIZE", "1") not in ("", "1"):
        print("bench.py runs as a single process on one GPU", file=sys.stderr)
        return 2

    from starpu_amd import kernels as K
    from starpu_amd import runtime as R
    if K.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        return 1
    n, nb = args.n, args.nb
    assert n % nb == 0
    nblocks = n // nb
    log(f"generating A (N={n}) on the GPU")
    t0 = time.perf_counter()
    A0 = make_spd(K, n, args.seed)
    Aw = K.DeviceArray(n, n)
    log(f"generated in {time.perf_counter() - t0:.2f} s")

    pmdir = os.path.join(ROOT, "gpurun_out", "perfmodels") if os.path.isdir(os.path.join(ROOT, "gpurun_out")) \
        else "/tmp/starpu_amd_perfmodels"
    R.init(STARPU_NHIP=1, STARPU_NCPU=0, STARPU_SCHED="dmda", STARPU_PERF_MODEL_DIR=pmdir,
           STARPU_NWORKER_PER_HIP=args.workers_per_gpu)
    node = R.hip_node(0)
    L = K.lib()
    flops = R.chol().chol_flops_potrf(float(n))

    def step():
        K.check(L.sa_memcpy_d2d(Aw.ptr, A0.ptr, n * n * 8), "restore A")
        K.sync()
        st = R.factor(Aw.ptr, n, n, nblocks, home_node=node)
        K.sync()
        if st.info != 0:
            raise RuntimeError(f"POTRF info {st.info}")
        return st

    for w in range(args.warmup):
        st = step()
        log(f"warmup {w}: {st.elapsed_us / 1e3:.1f} ms  {flops / st.elapsed_us / 1e6:.2f} TFLOP/s")
    times, gemm_us, gemm_n, kinds = [], 0.0, 0, np.zeros(4)
    kn = np.zeros(4)
    for s in range(args.steps):
        st = step()
        times.append(st.elapsed_us / 1e6)
        gemm_us += st.kernel_us[3]
        gemm_n += st.kernel_n[3]
        kinds += np.array(st.kernel_us[:])
        kn += np.array(st.kernel_n[:])
        log(f"step {s}: {st.elapsed_us / 1e3:.1f} ms  {flops / st.elapsed_us / 1e6:.2f} TFLOP/s")
    if args.trace:
        R.rt().starpu_trace_dump_chrome(args.trace.encode())
    R.shutdown()

    # factor check on a leading block against the CPU oracle: L11 of the
    # tiled GPU factor equals the Cholesky factor of A11
    import oracle
    m = min(2048, n)
    lead = np.empty((m, m), order="F")
    host = np.empty((n, m), order="F")   # first m columns
    K.check(L.sa_memcpy_d2h(host.ctypes.data, Aw.ptr, n * m * 8), "d2h")
    L11 = np.asfortranarray(host[:m, :m])
    K.check(L.sa_memcpy_d2h(host.ctypes.data, A0.ptr, n * m * 8), "d2h")
    lead[:] = host[:m, :m]
    ref, _ = oracle.tiled_cholesky(lead, max(1, m // nb))
    rel = oracle.rel_diff_lower(np.tril(L11), np.tril(ref))
    log(f"leading {m}x{m} factor vs oracle: rel diff {rel:.2e}")

    mean_t = float(np.mean(times))
    value = flops / mean_t / 1e12
    gemm_avg_us = gemm_us / max(1, gemm_n)
    gemm_flops = 2.0 * nb ** 3
    achieved = gemm_flops / (gemm_avg_us * 1e-6) / 1e12 if gemm_n else None

    cpu = None
    if args.cpu_n > 0:
        cores = args.cpu_cores or min(16, len(os.sched_getaffinity(0)))
        cn = min(args.cpu_n, n)
        lead = np.empty((cn, cn), order="F")
        host = np.empty((n, cn), order="F")
        K.check(L.sa_memcpy_d2h(host.ctypes.data, A0.ptr, n * cn * 8), "d2h")
        lead[:] = host[:cn, :cn]
        del host
        log(f"CPU baseline: N={cn} on {cores} cores")
        cpu = cpu_baseline(lead, nb, cores)
        log(f"CPU baseline: {cpu['value']:.4f} TFLOP/s")

    out = {
        "metric": "Cholesky TFLOP/s N=32768 FP64 vs mixed+IR at FP64 accuracy; p99 solve latency",
        "value": round(value, 4),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mean_t * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: A = B*B^T/N + I, B i.i.d. N(0,1) (seeded, generated on the GPU)",
        "config": {"workload": "tiled FP64 Cholesky factorisation (POTRF/TRSM/SYRK/GEMM DAG, dmda, "
                               "HIP codelets), device-resident A",
                   "N": n, "tile": nb, "tiles": f"{nblocks}x{nblocks}", "tasks": int(kn.sum()),
                   "hip_workers": args.workers_per_gpu, "sched": "dmda"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3) if achieved else None,
                     "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / FP64_PEAK_TFLOPS, 4) if achieved else None,
                     "traffic": None,
                     "kernel": "k_dgemm_nt (tile GEMM, 2*1024^3 flop per launch)",
                     "avg_launch_us": round(gemm_avg_us, 2)},
        "kernel_time_share": {k: round(float(v / max(kinds.sum(), 1e-9)), 4)
                              for k, v in zip(("potrf", "trsm", "syrk", "gemm"), kinds)},
        "parity": {"leading_block_rel_diff_vs_oracle": float(f"{rel:.3e}")},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
