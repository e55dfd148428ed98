"""Dense-field queries shared by the oracle backends (test infrastructure).

Restates the problem-blind half of reference src/coldcore/protocol.py:
143-219 (DenseFieldBackend): the mu-fold role predicates and every
count / map query over the dense ``cnt`` and ``map`` arrays, plus the
Backend attributes the search passes read (protocol.py:60-80).
"""
import numpy as np


class DenseQueries:
    extract_cap = 1 << 22
    supports_hist = True

    # ---- Backend protocol attributes
    @property
    def indep_dist(self):
        return 2 * self.R if self.problem in ("hamming", "lee") else None

    def word_index(self, words):
        return self.space.index(words)

    def index_word(self, idx):
        return self.space.digits(idx).astype(np.uint8)

    # ---- roles and queries
    def role_mask(self, target):
        if target == 0:
            return self.cnt < self.mu
        if target == 1:
            return (self.cnt >= 1) & (self.cnt <= self.mu)
        return self.cnt == target

    def count_eq(self, target):
        return int((self.cnt == target).sum())

    def uncovered(self):
        return int(np.maximum(self.mu - self.cnt.astype(np.int64), 0).sum())

    def map_max(self):
        return int(self.map.max(initial=0))

    def map_hist(self, nbins, vmax):
        out = np.zeros(nbins, dtype=np.int32)
        v = self.map[self.map > 0].astype(np.int64)
        if len(v) and vmax > 0:
            np.add.at(out, np.minimum((v - 1) * nbins // vmax, nbins - 1), 1)
        return out

    def map_extract(self, thr, cap=None):
        cap = min(cap or self.extract_cap, self.extract_cap)
        idx = np.flatnonzero(self.map >= thr).astype(np.int64)
        found = len(idx)
        idx = idx[:cap]
        return idx, self.map[idx].astype(np.int32), found

    def map_read_at(self, idx):
        return self.map[np.atleast_1d(np.asarray(idx, dtype=np.int64))].astype(np.int32)

    def read_cnt(self, off=0, length=None):
        length = self.space.size - off if length is None else length
        return self.cnt[off:off + length].astype(np.uint16)
