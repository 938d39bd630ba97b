"""Builds the native libraries of starpu_amd in-tree.

* ``starpu_amd/libstarpu_amd_kernels.so`` -- hand-written gfx950 HIP tile
  kernels behind the C ABI of ``include/starpu_amd_kernels.h`` (hipcc).
* ``starpu_amd/libstarpu_amd.so`` -- the C host runtime implementing the
  StarPU API subset of ``include/starpu*.h`` plus the HIP worker driver and
  memory node (gcc, linked against the HIP runtime and the kernels).
* ``starpu_amd/libstarpu_amd_chol.so`` -- the tiled Cholesky application
  (FP64 restatement of the reference's examples/cholesky DAG) with its HIP
  codelets, solve and refinement (gcc).
"""
from __future__ import annotations

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "starpu_amd")
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
ARCH = "gfx950"

HIPCC_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed-FP32 VALU forms stay disabled on this hardware.
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]
CFLAGS = ["-O2", "-g", "-fPIC", "-std=gnu11", "-Wall", "-Wextra", "-Werror=implicit-function-declaration",
          "-Wno-unused-parameter", "-Wno-missing-field-initializers",
          "-D__HIP_PLATFORM_AMD__", "-pthread"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError(f"build step failed ({r.returncode}): {' '.join(cmd[:3])} ...")


def _srcs(sub: str, ext: str) -> list[str]:
    d = os.path.join(CSRC, sub)
    return sorted(os.path.join(d, f) for f in os.listdir(d) if f.endswith(ext))


def build_kernels() -> str:
    out = os.path.join(PKG, "libstarpu_amd_kernels.so")
    _run(["hipcc", *HIPCC_FLAGS, f"-I{INCLUDE}", "-shared", "-o", out, *_srcs("hip", ".hip")])
    return out


def build_runtime() -> str:
    out = os.path.join(PKG, "libstarpu_amd.so")
    srcs = _srcs("runtime", ".c")
    if not srcs:
        return out
    _run(["gcc", *CFLAGS, f"-I{INCLUDE}", f"-I{ROCM}/include", f"-I{os.path.join(CSRC, 'runtime')}",
          "-shared", "-o", out, *srcs,
          f"-L{ROCM}/lib", "-lamdhip64", f"-Wl,-rpath,{ROCM}/lib", "-lm", "-ldl", "-lpthread"])
    return out


def build_apps() -> str:
    out = os.path.join(PKG, "libstarpu_amd_chol.so")
    srcs = _srcs("apps", ".c")
    if not srcs:
        return out
    _run(["gcc", *CFLAGS, f"-I{INCLUDE}", "-shared", "-o", out, *srcs,
          f"-L{PKG}", "-lstarpu_amd", "-lstarpu_amd_kernels", "-Wl,-rpath,$ORIGIN", "-lm", "-lpthread"])
    return out


def build_oracle() -> str:
    """Builds the CPU oracle (test infrastructure) with its own Makefile."""
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_all() -> None:
    build_kernels()
    build_runtime()
    build_apps()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built")
