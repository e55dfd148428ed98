"""Brute-force oracle: every field straight from the coverage predicate.

O(|X|^2 n), capped at 2^16 points like the reference
(reference src/coldcore/reference/base.py:29-43).  The four predicates:

* hamming   : #{j : a_j != b_j} <= R           (reference covering.py:16-17)
* lee       : sum_j min(|a_j-b_j|, ax_j-|a_j-b_j|) <= R   (lee.py:22-25)
* torus_linf: max_j min(|a_j-b_j|, ax_j-|a_j-b_j|) <= R   (torus.py:22-25)
* grid_linf : max_j |a_j-b_j| <= R              (grid.py:23-25)

Fields (reference base.py:65-106, protocol.py:163-219):
  cnt[u]   = #{c in code (multiset) : covers(u, c)}
  gain[w]  = #{u in ball(w) : cnt[u] < mu}
  loss[w]  = #{u in ball(w) : 1 <= cnt[u] <= mu}
"""
import numpy as np

from .protocol import DenseQueries
from .space import Space

MAX_SPACE = 1 << 16
PROBLEMS = ("hamming", "torus_linf", "lee", "grid_linf")


def covers(problem, axes, R, a, b):
    """Boolean coverage predicate for broadcastable digit arrays."""
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    if problem == "hamming":
        return (a != b).sum(axis=-1) <= R
    d = np.abs(a - b)
    if problem == "grid_linf":
        return d.max(axis=-1) <= R
    ax = np.asarray(axes, dtype=np.int64)
    d = np.minimum(d, ax - d)
    if problem == "lee":
        return d.sum(axis=-1) <= R
    if problem == "torus_linf":
        return d.max(axis=-1) <= R
    raise ValueError(problem)


class BruteForce(DenseQueries):
    def __init__(self, axes, R, problem="hamming", mu=1):
        if problem not in PROBLEMS:
            raise ValueError(problem)
        self.problem, self.R, self.mu = problem, int(R), int(mu)
        self.space = Space(axes)
        if self.space.size > MAX_SPACE:
            raise ValueError("brute force is capped at 2^16 points")
        if problem == "torus_linf" and any(2 * R + 1 > a for a in axes):
            raise ValueError("need 2R+1 <= ax_j on every axis")
        self._digs = self.space.all_digits()
        size = self.space.size
        self.cnt = np.zeros(size, dtype=np.int64)
        self.map = np.zeros(size, dtype=np.int64)

    def _cov(self, a, b):
        return covers(self.problem, self.space.axes, self.R, a, b)

    def ball_indices(self, w):
        return np.flatnonzero(self._cov(self._digs, self._digs[int(w)][None, :]))

    def recount(self, code):
        self.cnt[:] = 0
        for w in np.atleast_1d(np.asarray(code, dtype=np.int64)):
            self.cnt[self.ball_indices(w)] += 1

    def _field(self, mask):
        sel = np.flatnonzero(mask)
        self.map[:] = 0
        if not len(sel):
            return
        size = self.space.size
        chunk = max(1, (1 << 22) // len(sel))
        for lo in range(0, size, chunk):
            hi = min(size, lo + chunk)
            m = self._cov(self._digs[lo:hi, None, :], self._digs[None, sel, :])
            self.map[lo:hi] = m.sum(axis=1)

    def gain_map(self):
        self._field(self.role_mask(0))

    def loss_map(self):
        self._field(self.role_mask(1))

    def ball_update(self, idx, delta):
        for w in np.atleast_1d(np.asarray(idx, dtype=np.int64)):
            self.cnt[self.ball_indices(w)] += delta

    def ball_gather(self, idx, target):
        mask = self.role_mask(target)
        idx = np.atleast_1d(np.asarray(idx, dtype=np.int64))
        return np.array([int(mask[self.ball_indices(w)].sum()) for w in idx],
                        dtype=np.int32)


def distance(problem, axes, a, b):
    """Exact metric distance for the layered families (brute force)."""
    a = np.asarray(a, dtype=np.int64)
    b = np.asarray(b, dtype=np.int64)
    if problem == "hamming":
        return (a != b).sum(axis=-1)
    d = np.abs(a - b)
    d = np.minimum(d, np.asarray(axes, dtype=np.int64) - d)
    return d.sum(axis=-1)


def layers_brute(problem, axes, R, seed_words, idx):
    """layer d at x = #{s in seed (multiset) : distance(x, s) == d}."""
    from .space import Space
    sp = Space(axes)
    X = sp.digits(np.asarray(idx, dtype=np.int64))
    S = sp.digits(np.asarray(seed_words, dtype=np.int64))
    out = np.zeros((len(X), R + 1), dtype=np.int64)
    if len(S):
        d = distance(problem, axes, X[:, None, :], S[None, :, :])
        for r in range(R + 1):
            out[:, r] = (d == r).sum(axis=1)
    return out
