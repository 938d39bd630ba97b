"""starpu_amd: the StarPU task-based runtime model on AMD Instinct MI355X.

Native libraries (built in-tree by ``python -m starpu_amd._build``):

* ``libstarpu_amd_kernels.so`` -- hand-written gfx950 HIP tile kernels
  (C ABI: include/starpu_amd_kernels.h), bound in ``starpu_amd.kernels``;
* ``libstarpu_amd.so`` -- the C runtime (StarPU API: include/starpu.h);
* ``libstarpu_amd_chol.so`` -- the tiled Cholesky application
  (include/starpu_amd_chol.h); both bound in ``starpu_amd.runtime``.
"""
__all__ = ["kernels", "runtime"]
