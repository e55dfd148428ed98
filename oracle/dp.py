"""numpy restatement of the reference ``cpu`` backend (the axis DP).

Follows reference src/coldcore/dp.py:
  * per-axis passes on a (layers, H, Q, S) view of the layer stack
    (dp.py:235-273);
  * transform = all n passes then the layer sum (dp.py:275-285);
  * recount / gain / loss seeding (dp.py:287-308);
  * walks as (centre + per-(position, delta) displacement) gathers
    (dp.py:115-220, 311-367);
and the dense-field query semantics of reference src/coldcore/protocol.py:
163-219 (mu predicates, deficit, hist binning, ascending extract).

Every array is int32 by default, like the reference's own int32 planes, so
wrap-around (if any cell ever reached it) is reproduced exactly.
"""
from itertools import combinations

import numpy as np

from .protocol import DenseQueries
from .space import Space

PROBLEMS = ("hamming", "torus_linf", "lee", "grid_linf")
_GATHER_ELEMS = 1 << 23


def delta_sets(axes, R, problem):
    """Admissible circular shifts per position (reference patterns.py:34-50)."""
    out = []
    for a in axes:
        if problem == "hamming":
            ds = list(range(1, a))
        elif problem == "lee":
            ds = [d for d in range(1, a) if min(d, a - d) <= R]
        elif problem in ("torus_linf", "grid_linf"):
            ds = [d for d in range(1, a) if d <= R or a - d <= R]
        else:
            raise ValueError(problem)
        out.append(ds)
    return out


def max_weight(n, R, problem):
    """Max changed positions per support pattern (reference patterns.py:53-59)."""
    if problem == "hamming":
        return min(R, n)
    if problem == "lee":
        return min(n, R)
    return n


class NumpyDP(DenseQueries):
    def __init__(self, axes, R, problem="hamming", mu=1, dtype=np.int32):
        if problem not in PROBLEMS:
            raise ValueError(f"unknown problem {problem!r}")
        self.problem, self.R, self.mu = problem, int(R), int(mu)
        if self.mu < 1:
            raise ValueError("mu must be >= 1")
        self.space = Space(axes)
        self.axes = self.space.axes
        self.n = self.space.n
        if problem == "torus_linf" and any(2 * R + 1 > a for a in self.axes):
            raise ValueError("need 2R+1 <= ax_j on every axis")
        size = self.space.size
        self.dtype = np.dtype(dtype)
        self.nlayers = 1 if problem in ("torus_linf", "grid_linf") else self.R + 1
        self.cnt = np.zeros(size, dtype=self.dtype)
        self.map = np.zeros(size, dtype=self.dtype)
        self._A = np.zeros((self.nlayers, size), dtype=self.dtype)
        self._geom = []
        for j in range(self.n):
            s = int(self.space.pow[j])
            q = self.axes[j]
            self._geom.append((size // (s * q), q, s))
        self._build_keys()

    # ------------------------------------------------------------ support set
    def _build_keys(self):
        """Ball as rows of (position*span + delta-1) keys, pad = zero slot."""
        n, R = self.n, self.R
        ds = delta_sets(self.axes, R, self.problem)
        mw = max_weight(n, R, self.problem)
        span = max(self.axes)
        pad = n * span
        rows = [np.full((1, max(mw, 1)), pad, dtype=np.int64)]
        for w in range(1, mw + 1):
            for subs in combinations(range(n), w):
                sets = [np.asarray(ds[p], dtype=np.int64) for p in subs]
                if any(len(s) == 0 for s in sets):
                    continue
                grids = np.meshgrid(*sets, indexing="ij")
                d = np.stack([g.ravel() for g in grids], axis=1)
                if self.problem == "lee":
                    wt = sum(np.minimum(d[:, k], self.axes[p] - d[:, k])
                             for k, p in enumerate(subs))
                    d = d[wt <= R]
                if not len(d):
                    continue
                blk = np.full((len(d), max(mw, 1)), pad, dtype=np.int64)
                for k, p in enumerate(subs):
                    blk[:, k] = p * span + d[:, k] - 1
                rows.append(blk)
        self._span = span
        self._keys = np.concatenate(rows)
        self.ball_len = len(self._keys)

    def _balls(self, idx):
        """(m, |B|) covered indices per centre and a validity mask (grid)."""
        idx = np.asarray(idx, dtype=np.int64).ravel()
        digs = self.space.digits(idx)
        m, span = len(idx), self._span
        disp = np.zeros((m, self.n * span + 1), dtype=np.int64)
        ok = np.ones((m, self.n * span + 1), dtype=bool)
        for p, a in enumerate(self.axes):
            d = np.arange(1, a, dtype=np.int64)
            base = digs[:, p:p + 1]
            sh = (base + d[None, :]) % a - base
            disp[:, p * span:p * span + a - 1] = sh * int(self.space.pow[p])
            ok[:, p * span:p * span + a - 1] = np.abs(sh) <= self.R
        off = idx[:, None] + disp[:, self._keys].sum(axis=2)
        if self.problem != "grid_linf":
            return off, None
        return off, ok[:, self._keys].all(axis=2)

    def ball_indices(self, w):
        off, good = self._balls([int(w)])
        row = off[0] if good is None else off[0][good[0]]
        return np.sort(row)

    def _chunk(self):
        return max(1, _GATHER_ELEMS // max(self.ball_len, self.n * self._span))

    # ------------------------------------------------------------ axis passes
    def _view(self, j):
        h, q, s = self._geom[j]
        return self._A.reshape(self.nlayers, h, q, s)

    def _pass_hamming(self, j):
        v = self._view(j)
        t = v.sum(axis=2, dtype=self.dtype)
        for d in range(self.nlayers - 1, 0, -1):
            v[d] += t[d - 1][:, None, :] - v[d - 1]

    def _pass_lee(self, j):
        v = self._view(j)
        q = self._geom[j][1]
        half = q // 2
        for d in range(self.nlayers - 1, 0, -1):
            for w in range(1, min(d, half) + 1):
                src = v[d - w]
                v[d] += np.roll(src, w, axis=1)
                if not (q % 2 == 0 and w == half):
                    v[d] += np.roll(src, -w, axis=1)

    def _pass_torus(self, j):
        v = self._view(j)
        acc = np.zeros_like(v[0])
        for e in range(-self.R, self.R + 1):
            acc += np.roll(v[0], -e, axis=1)
        v[0] = acc

    def _pass_grid(self, j):
        v = self._view(j)
        h, q, s = self._geom[j]
        cs = np.zeros((h, q + 1, s), dtype=self.dtype)
        np.cumsum(v[0], axis=1, out=cs[:, 1:, :])
        a = np.arange(q)
        lo = np.maximum(a - self.R, 0)
        hi = np.minimum(a + self.R, q - 1) + 1
        v[0] = cs[:, hi, :] - cs[:, lo, :]

    def _transform(self, out):
        step = {"hamming": self._pass_hamming, "lee": self._pass_lee,
                "torus_linf": self._pass_torus,
                "grid_linf": self._pass_grid}[self.problem]
        for j in range(self.n):
            step(j)
        if self.nlayers == 1:
            out[:] = self._A[0]
        else:
            np.sum(self._A, axis=0, out=out, dtype=self.dtype)

    def _seed(self, a0):
        self._A[0] = a0
        if self.nlayers > 1:
            self._A[1:] = 0

    # ------------------------------------------------------------ fields
    def recount(self, code):
        code = np.atleast_1d(np.asarray(code, dtype=np.int64))
        self._A[0] = 0
        if code.size:
            np.add.at(self._A[0], code, 1)
        if self.nlayers > 1:
            self._A[1:] = 0
        self._transform(self.cnt)

    def gain_map(self):
        self._seed(self.role_mask(0))
        self._transform(self.map)

    def loss_map(self):
        self._seed(self.role_mask(1))
        self._transform(self.map)

    def layers_after_transform(self):
        """The exact distance-count layers A_0..A_R left by the last transform."""
        return self._A.copy()

    # ------------------------------------------------------------ walks
    def ball_update(self, idx, delta):
        idx = np.atleast_1d(np.asarray(idx, dtype=np.int64))
        step = self._chunk()
        for lo in range(0, len(idx), step):
            off, good = self._balls(idx[lo:lo + step])
            for k in range(len(off)):
                row = off[k] if good is None else off[k][good[k]]
                self.cnt[row] += delta

    def ball_gather(self, idx, target):
        idx = np.atleast_1d(np.asarray(idx, dtype=np.int64))
        out = np.empty(len(idx), dtype=np.int32)
        if not len(idx):
            return out
        mask = self.role_mask(target)
        step = self._chunk()
        for lo in range(0, len(idx), step):
            off, good = self._balls(idx[lo:lo + step])
            hit = mask[off]
            if good is not None:
                hit &= good
            out[lo:lo + step] = hit.sum(axis=1)
        return out

    def layers_at(self, init_mode, idx, code=None):
        """(k, R+1) exact distance layers at idx, seeded with the codewords
        (init 0, ``code``), the deficient words (1) or the critical words
        (2); cnt and the map are unchanged."""
        if self.nlayers < 2:
            raise ValueError("layers exist for the layered families only")
        saved = self._A.copy()
        if init_mode == 0:
            self._A[0] = 0
            if code is not None and len(code):
                np.add.at(self._A[0], np.asarray(code, dtype=np.int64), 1)
        else:
            self._A[0] = self.role_mask(1 if init_mode == 2 else 0)
        self._A[1:] = 0
        step = {"hamming": self._pass_hamming, "lee": self._pass_lee}[self.problem]
        for j in range(self.n):
            step(j)
        out = self._A[:, np.asarray(idx, dtype=np.int64)].T.astype(np.int32)
        self._A[:] = saved
        return out
