"""BF16 rounding and the split products of the mixed-precision updates,
restated in numpy.

TEST INFRASTRUCTURE ONLY (tests/): the checker for the mixed-precision
GEMM kernels (sa_sgemm_tn).  The reference has no mixed-precision path;
this follows the published bfloat16 format (1 sign, 8 exponent, 7 stored
mantissa bits: the upper half of an IEEE binary32) with round-to-nearest,
ties-to-even, as the gfx950 conversion instructions do for finite values.
"""
from __future__ import annotations

import numpy as np


def bf16_round(x) -> np.ndarray:
    """float32 values rounded to the nearest bfloat16 (ties to even),
    returned as float32."""
    f = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32).reshape(f.shape)
    return np.where(np.isnan(f), f, out)


def split_bf16(x):
    """x = hi + lo + (residual) with hi = bf16(x), lo = bf16(x - hi)."""
    f = np.asarray(x, dtype=np.float32)
    hi = bf16_round(f)
    lo = bf16_round(f - hi)
    return hi, lo


def gemm_tn_model(mode: int, A, B) -> np.ndarray:
    """A^T B as the kernel forms its products, summed in float64:
    mode 0 FP32 inputs, 1 bf16(A)^T bf16(B), 2 the three-term split."""
    A32 = np.asarray(A, dtype=np.float32)
    B32 = np.asarray(B, dtype=np.float32)
    if mode == 0:
        return A32.astype(np.float64).T @ B32.astype(np.float64)
    if mode == 1:
        return bf16_round(A32).astype(np.float64).T @ bf16_round(B32).astype(np.float64)
    ah, al = (t.astype(np.float64) for t in split_bf16(A32))
    bh, bl = (t.astype(np.float64) for t in split_bf16(B32))
    return ah.T @ bh + ah.T @ bl + al.T @ bh
