"""ORACLE — TEST INFRASTRUCTURE ONLY.

Row-scaled complex FP16/BF16 storage of an FP64 matrix, restated in numpy.

The reference's explicit-scale precedent is quantize_scaled_complex32
(crates/sonium-compute/src/scaled_storage.rs:105-127: one scale from the
largest |re| or |im|, interleaved re/im halves).  The device format keeps
that layout with one power-of-two scale per row, chosen from the FP64 row:

    e_i   = floor(log2(max_j max(|Re A_ij|, |Im A_ij|)))   (0 for a zero or
            non-finite maximum)
    stored = half(float32(A_ij * 2^(T - e_i))),  T = 14 (FP16) or 0 (BF16)
    row_inv[i] = 2^(e_i - T)

half() rounds to nearest, ties to even; the FP64 value is rounded to FP32
first (the scaling by a power of two is exact), so a value within 2^-24 of a
half-way point can round differently than a direct FP64 rounding would.
Rows are padded with zeros to ``ld`` complex entries.
"""

from __future__ import annotations

import numpy as np

TOP = {"fp16": 14, "bf16": 0}


def _bf16_bits(f32: np.ndarray) -> np.ndarray:
    b = f32.view(np.uint32).astype(np.uint64)
    return (((b + 0x7FFF + ((b >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def half_bits(values: np.ndarray, fmt: str) -> np.ndarray:
    """FP64 values -> FP32 (nearest even) -> FP16/BF16 bit patterns (nearest even)."""
    f32 = np.asarray(values, dtype=np.float64).astype(np.float32)
    if fmt == "fp16":
        with np.errstate(over="ignore"):
            return f32.astype(np.float16).view(np.uint16)
    return _bf16_bits(f32)


def row_exponents(A: np.ndarray) -> np.ndarray:
    m = np.maximum(np.abs(A.real), np.abs(A.imag)).max(axis=1)
    ok = (m > 0) & np.isfinite(m)
    _, ex = np.frexp(np.where(ok, m, 1.0))  # m = f 2^ex, f in [0.5, 1)
    return np.where(ok, ex - 1, 0).astype(np.int64)


def row_scaled_halves(A: np.ndarray, fmt: str, ld: int):
    """(halves uint16 n x 2*ld, row_inv float64 n) of the complex matrix A (n x n)."""
    A = np.asarray(A, dtype=np.complex128)
    n, cols = A.shape
    top = TOP[fmt]
    e = row_exponents(A)
    scale = np.ldexp(1.0, top - e)[:, None]
    out = np.zeros((n, 2 * ld), dtype=np.uint16)
    out[:, 0:2 * cols:2] = half_bits(A.real * scale, fmt)
    out[:, 1:2 * cols:2] = half_bits(A.imag * scale, fmt)
    return out, np.ldexp(1.0, e - top)


def quantized_columns(X: np.ndarray, fmt: str) -> np.ndarray:
    """The block operand as the device quantizes it: per column a power-of-two
    scale from the column's largest |re| or |im| (taken in FP32), entries
    rounded through FP32 to the format, returned dequantized in FP64."""
    X = np.asarray(X, dtype=np.complex128)
    X2 = X.reshape(X.shape[0], -1)
    m = np.maximum(np.abs(X2.real), np.abs(X2.imag)).astype(np.float32).max(axis=0).astype(np.float64)
    ok = (m > 0) & np.isfinite(m)
    _, ex = np.frexp(np.where(ok, m, 1.0))
    e = np.where(ok, ex - 1, 0)
    s = np.ldexp(1.0, TOP[fmt] - e)[None, :]

    def deq(bits):
        if fmt == "fp16":
            return bits.view(np.float16).astype(np.float64)
        return (bits.astype(np.uint32) << 16).view(np.float32).astype(np.float64)

    q = (deq(half_bits(X2.real * s, fmt)) + 1j * deq(half_bits(X2.imag * s, fmt))) / s
    return q.reshape(X.shape)

