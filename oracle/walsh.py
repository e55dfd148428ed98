"""Independent full-size oracle for binary Hamming cells (pure mathematics).

On X = Z_2^n the coverage count is a convolution over (Z_2^n, xor):
  cnt = 1_S * 1_B,  1_B(x) = [popcount(x) <= R],
and the integer Walsh-Hadamard transform diagonalises it:
  cnt = WHT( WHT(1_S) . WHT(1_B) ) / 2^n        (exact in int64).
The same identity gives the gain and loss fields from the role indicator.
It shares no code or structure with the per-axis DP, so agreement with the
engine at K_2(24,3) is a size-independent check of the whole transform.
"""
import numpy as np


def wht_inplace(a):
    """Unnormalised integer Walsh-Hadamard transform (length 2^n, int64)."""
    n = a.size
    h = 1
    while h < n:
        v = a.reshape(-1, 2, h)
        x = v[:, 0, :].copy()
        v[:, 0, :] += v[:, 1, :]
        x -= v[:, 1, :]
        v[:, 1, :] = x
        h *= 2
    return a


def ball_indicator(n, R):
    idx = np.arange(1 << n, dtype=np.int64)
    pc = np.zeros(1 << n, dtype=np.int64)
    x = idx.copy()
    for _ in range(n):
        pc += x & 1
        x >>= 1
    return (pc <= R).astype(np.int64)


def binary_field(indicator, n, R):
    """sum_{u : popcount(u ^ x) <= R} indicator[u] for every x, exact."""
    f = np.asarray(indicator, dtype=np.int64).copy()
    b = ball_indicator(n, R)
    wht_inplace(f)
    wht_inplace(b)
    f *= b
    wht_inplace(f)
    return f >> n
