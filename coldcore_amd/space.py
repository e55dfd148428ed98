"""Finite product space X = prod_j Z_{ax_j}: index <-> digit conversions.

Mirrors reference src/coldcore/space.py:10-45 (same class name and
methods).  Little-endian: index = sum_j digit_j * prod_{i<j} ax_i.
"""
import numpy as np


class ProductSpace:

    def __init__(self, q_or_axes, n=None):
        if n is not None:
            axes = [int(q_or_axes)] * int(n)
        else:
            axes = [int(a) for a in q_or_axes]
        self.axes = axes
        self.n = len(axes)
        self.q = axes[0] if len(set(axes)) == 1 else None
        size = 1
        pows = []
        for a in axes:
            pows.append(size)
            size *= a
        self.size = size
        self.pow_q = np.array(pows, dtype=np.int64)
        self._ax = np.array(axes, dtype=np.int64)

    def word_index(self, words):
        """(k, n) digit array -> int64 indices."""
        return np.asarray(words, dtype=np.int64) @ self.pow_q

    def index_word(self, idx):
        """int64 indices -> (k, n) uint8 digit array."""
        x = np.asarray(idx, dtype=np.int64).ravel().copy()
        out = np.empty((len(x), self.n), dtype=np.uint8)
        for i in range(self.n):
            out[:, i] = x % self._ax[i]
            x //= self._ax[i]
        return out

    def all_words(self):
        """(size, n) digit array of the whole space (tiny spaces only)."""
        return self.index_word(np.arange(self.size, dtype=np.int64))
