"""Pins of the oracle's row-scaled FP16/BF16 storage (oracle/lowp.py) by
closed forms: exact bit patterns of powers of two, ties to even at the
storage precision, the FP32 intermediate rounding, the power-of-two row
exponent at its boundaries, zero rows and the zero padding."""

import numpy as np
import pytest

from oracle import lowp


def test_fp16_powers_of_two_and_scale():
    # row max 3 -> e = 1 -> scale 2^13; 1 -> 2^13 (0x7000), 0.5 -> 2^12 (0x6C00), -3 -> -3*2^13 (0xF600)
    A = np.array([[1.0 + 0j, 0.5j, -3.0 + 0j]])
    h, inv = lowp.row_scaled_halves(A, "fp16", 4)
    assert inv[0] == 2.0 ** -13
    assert list(h[0, :6]) == [0x7000, 0x0000, 0x0000, 0x6C00, 0xF600, 0x0000]
    assert np.all(h[0, 6:] == 0)  # padding


def test_fp16_ties_round_to_even_and_fp32_intermediate():
    # row max in [1, 2): scale 2^14, half spacing at 2^14 is 16
    vals = np.array([1.0, 1 + 2.0 ** -11, 1 + 3 * 2.0 ** -11, 1 + 2.0 ** -11 + 2.0 ** -40, 1 + 2.0 ** -11 + 2.0 ** -20])
    A = vals[None, :] + 0j
    h, inv = lowp.row_scaled_halves(A, "fp16", 8)
    assert inv[0] == 2.0 ** -14
    re = h[0, 0:10:2].view(np.float16).astype(np.float64)
    # 16384 + 8 is a tie -> 16384 (even); 16384 + 24 -> 16416 (even);
    # + 2^-40 vanishes in FP32, leaving the tie; + 2^-20 survives FP32 and rounds up
    assert list(re) == [16384.0, 16384.0, 16416.0, 16384.0, 16400.0]


def test_bf16_ties_and_exponent_boundaries():
    A = np.array([[1.0 + 0j, 1 + 2.0 ** -8, 1 + 3 * 2.0 ** -8, 1 + 2.0 ** -7]])
    h, inv = lowp.row_scaled_halves(A, "bf16", 4)
    assert inv[0] == 1.0
    assert list(h[0, 0:8:2]) == [0x3F80, 0x3F80, 0x3F82, 0x3F81]
    # e from the largest |re| or |im|: exactly 2^k gives e = k, just below gives k-1
    for m, e in ((4.0, 2), (np.nextafter(4.0, 0), 1), (2.0 ** -30, -30), (0.75, -1)):
        B = np.array([[0.1 * m + 0j, 1j * m]])
        assert lowp.row_exponents(B)[0] == e


def test_zero_and_nonfinite_rows_use_exponent_zero():
    A = np.zeros((2, 3), dtype=np.complex128)
    A[1, 0] = np.inf
    assert list(lowp.row_exponents(A)) == [0, 0]
    h, inv = lowp.row_scaled_halves(A[:1], "fp16", 4)
    assert inv[0] == 2.0 ** -14 and np.all(h == 0)


@pytest.mark.parametrize("fmt,u", [("fp16", 2.0 ** -11), ("bf16", 2.0 ** -8)])
def test_dequantized_error_is_one_rounding(fmt, u):
    rng = np.random.default_rng(3)
    A = (rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20))) * np.exp(rng.uniform(-20, 20, (20, 1)))
    h, inv = lowp.row_scaled_halves(A, fmt, 32)
    if fmt == "fp16":
        v = h.view(np.float16).astype(np.float64)
    else:
        v = (h.astype(np.uint32) << 16).view(np.float32).astype(np.float64)
    D = (v[:, 0:40:2] + 1j * v[:, 1:40:2]) * inv[:, None]
    floor = (2.0 ** -24 if fmt == "fp16" else 0.0) * inv[:, None]
    assert np.all(np.abs(D.real - A.real) <= 1.01 * u * np.abs(A.real) + floor)
    assert np.all(np.abs(D.imag - A.imag) <= 1.01 * u * np.abs(A.imag) + floor)
    # the largest scaled entry of each row stays in [2^T, 2^(T+1))
    top = lowp.TOP[fmt]
    rowmax = np.maximum(np.abs(A.real), np.abs(A.imag)).max(axis=1)
    s = rowmax / inv / 2.0 ** top
    assert np.all((s >= 1.0) & (s < 2.0))


def test_quantized_columns_scale_and_rounding():
    # column 0: max 3 -> scale 2^13 (FP16)
    X = np.array([[3.0 + 0j, 0.5j], [1 + 2.0 ** -12 + 0j, 0.25 + 0j]])
    q = lowp.quantized_columns(X, "fp16")
    assert q[0, 0] == 3.0 and q[1, 1] == 0.25 and q[0, 1] == 0.5j
    # (1 + 2^-12) * 2^13 = 8194, and the FP16 spacing at 2^13 is 8 -> 8192
    assert q[1, 0] == 1.0
    # a zero column stays zero; a 1-D vector keeps its shape
    assert np.all(lowp.quantized_columns(np.zeros(4, complex), "bf16") == 0)
    assert lowp.quantized_columns(np.ones(3, complex), "bf16").shape == (3,)
