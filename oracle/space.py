"""Mixed-radix indexing of X = Z_{a_0} x ... x Z_{a_{n-1}} (oracle copy).

Little-endian: index = sum_j digit_j * prod_{i<j} a_i
(reference src/coldcore/space.py:1-45).
"""
import numpy as np


class Space:
    def __init__(self, axes):
        self.axes = [int(a) for a in axes]
        self.n = len(self.axes)
        self.pow = np.ones(self.n, dtype=np.int64)
        for j in range(1, self.n):
            self.pow[j] = self.pow[j - 1] * self.axes[j - 1]
        self.size = int(np.prod(np.array(self.axes, dtype=object))) if self.n else 1

    def digits(self, idx):
        """int64 indices -> (k, n) int64 digit array."""
        x = np.asarray(idx, dtype=np.int64).ravel().copy()
        out = np.empty((len(x), self.n), dtype=np.int64)
        for j, a in enumerate(self.axes):
            out[:, j] = x % a
            x //= a
        return out

    def index(self, digs):
        return np.asarray(digs, dtype=np.int64) @ self.pow

    def all_digits(self):
        return self.digits(np.arange(self.size, dtype=np.int64))
