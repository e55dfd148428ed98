"""Mixed-radix words of the cell: the integer index of a word and its digits.

A cell is the product of cyclic axes Z_{a_0} x ... x Z_{a_{n-1}}.  Words are
numbered little-endian (axis 0 varies fastest): the index of digits
(d_0, ..., d_{n-1}) is sum_j d_j * s_j with place values s_0 = 1 and
s_{j+1} = s_j * a_j.  This is the layout of every plane in HBM (the engine's
axis-j fibers are the words s_j apart), so host and device agree on indices.

Interface as reference src/coldcore/space.py:10-45 (``ProductSpace`` built
from (q, n) or an axis list; ``size``, ``n``, ``axes``, ``q``, ``pow_q``,
``word_index``, ``index_word``, ``all_words``).
"""
import numpy as np


class ProductSpace:
    """A mixed-radix product space: ``ProductSpace(q, n)`` is Z_q^n,
    ``ProductSpace([a_0, ..., a_{n-1}])`` the general product."""

    def __init__(self, q_or_axes, n=None):
        radices = ([int(q_or_axes)] * int(n) if n is not None
                   else [int(a) for a in q_or_axes])
        self.axes = radices
        self.n = len(radices)
        # the common radix of a homogeneous cell, None for mixed radix
        self.q = radices[0] if radices and radices.count(radices[0]) == self.n else None
        self._radix = np.asarray(radices, dtype=np.int64)
        # place values: an exclusive running product of the radices
        place = np.ones(self.n, dtype=np.int64)
        if self.n > 1:
            place[1:] = np.cumprod(self._radix[:-1])
        self.pow_q = place
        self.size = int(place[-1] * self._radix[-1]) if self.n else 1

    def word_index(self, words):
        """Digit rows (k, n) to their int64 indices (the dot product with the
        place values)."""
        return np.asarray(words, dtype=np.int64) @ self.pow_q

    def index_word(self, idx):
        """int64 indices to digit rows (k, n), uint8: digit j is
        floor(index / s_j) mod a_j."""
        flat = np.asarray(idx, dtype=np.int64).reshape(-1)
        digits = (flat[:, None] // self.pow_q[None, :]) % self._radix[None, :]
        return digits.astype(np.uint8)

    def all_words(self):
        """Every word of the cell in index order, (size, n) -- small cells
        only: it materialises size x n digits."""
        return self.index_word(np.arange(self.size, dtype=np.int64))
