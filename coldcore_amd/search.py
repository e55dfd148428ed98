"""The search passes that call the hot path: lazy greedy, ruin, peel, LNS.

The caller side of the engine (reference src/coldcore/search.py): a
problem-blind search over any backend exposing the coldcore Backend
protocol (``GpuBackend`` here, or a host backend).  Only the passes the
benchmark workloads use are restated:

* ``greedy_fill``  exact lazy greedy (search.py:260-360)
* ``ruin``         low / hole / cluster / random / deficit (search.py:536-618)
* ``peel``         batched zero-loss removal (search.py:621-669)
* ``lns``          ruin-and-recreate rounds at fixed M, fixed mix or the
                   adaptive ``OperatorBandit`` (search.py:96-167, 672-784),
                   with an optional round cap so a run can be clock-free
* ``uncovered_words`` deficit-word scan over ``read_cnt`` (search.py:462-529)

Decisions consume the random stream in the reference's order, draw for
draw, so a run is a function of the seed and the exact fields alone: the
same seed gives the same move sequence on every backend whose fields and
extraction order agree (the engine returns extraction hits in ascending
order, as the host backends do).  The reference's frozen trace digests
(tests/test_search_determinism.py:110-140) pin this restatement.
"""
import heapq
import time

import numpy as np

RUIN_MODES = ("low", "hole", "cluster", "random")
RUIN_MIX = {"low": 1.0, "hole": 2.0, "cluster": 1.0, "random": 1.0}
RUIN_MODES_EXT = RUIN_MODES + ("deficit",)
RUIN_MIX_EXT = dict(RUIN_MIX, deficit=2.0)
FIXED_DRAW = ("low", "hole", "hole", "cluster", "random")
FIXED_DRAW_DEFICIT = FIXED_DRAW + ("deficit", "deficit")
DEFICIT_WORDS = 4
SCAN_CHUNK = 1 << 24


class OperatorBandit:
    """Roulette-wheel ruin-operator selection over an EMA of tiered
    rewards, floor-mixed so no operator starves (search.py:96-167)."""

    REWARD_BEST, REWARD_BETTER, REWARD_EQUAL, REWARD_WORSE = 3.0, 1.0, 0.3, 0.0

    def __init__(self, rng, ops=RUIN_MODES, floor=0.15, decay=0.9,
                 prior=RUIN_MIX):
        self.ops = list(ops)
        self.rng = rng
        self.floor, self.decay = float(floor), float(decay)
        w = np.array([float(prior.get(o, 1.0)) for o in self.ops])
        self.score = w * (len(w) / w.sum())
        self.uses = np.zeros(len(self.ops), dtype=np.int64)
        self.reward = np.zeros(len(self.ops), dtype=np.float64)

    def probabilities(self, boost=None):
        w = np.maximum(self.score, 0.0)
        k = len(self.ops)
        if boost:
            w = w.copy()
            for op, f in boost.items():
                if op in self.ops:
                    i = self.ops.index(op)
                    w[i] = max(w[i], self.floor) * float(f)
        if w.sum() <= 0:
            return np.full(k, 1.0 / k)
        return self.floor / k + (1.0 - self.floor) * w / w.sum()

    def select(self, boost=None):
        i = int(self.rng.choice(len(self.ops), p=self.probabilities(boost)))
        self.uses[i] += 1
        return self.ops[i]

    def update(self, op, reward):
        i = self.ops.index(op)
        self.reward[i] += reward
        self.score[i] = self.decay * self.score[i] + (1.0 - self.decay) * reward

    def summary(self):
        p = self.probabilities()
        return {o: {"uses": int(self.uses[i]),
                    "reward": float(self.reward[i] / max(1, self.uses[i])),
                    "p": float(p[i])} for i, o in enumerate(self.ops)}


def _digit_distance(digs, centre):
    return (digs != centre[None, :]).sum(axis=1)


class _StaleQueue:
    """The lazy greedy's priority queue of (-bound, word) pairs: the refresh's
    candidates as one array sorted by (-bound, word) plus a binary heap of
    the pairs pushed back with exact gains.  pop() takes the smaller head, so
    the pop order is that of one heap holding every pair, without building a
    heap of the whole (up to 2^18-entry) refresh."""

    __slots__ = ("k", "i", "p", "n", "h")

    def __init__(self, keys, ids):
        self.k, self.i = keys, ids
        self.p, self.n = 0, len(keys)
        self.h = []

    @classmethod
    def empty(cls):
        return cls(np.zeros(0, dtype=np.int64), np.zeros(0, dtype=np.int64))

    def __bool__(self):
        return self.p < self.n or bool(self.h)

    def _head(self):
        return (int(self.k[self.p]), int(self.i[self.p]))

    def pop(self):
        if self.p < self.n and (not self.h or self._head() < self.h[0]):
            t = self._head()
            self.p += 1
            return t
        return heapq.heappop(self.h)

    def push(self, t):
        heapq.heappush(self.h, t)

    def top_key(self):
        if self.p < self.n and (not self.h or self._head() < self.h[0]):
            return int(self.k[self.p])
        return self.h[0][0]


class Searcher:
    """Ruin-and-recreate search over one backend (search.py:170-257)."""

    def __init__(self, backend, seed=1, deadline=None, on_log=None,
                 on_round=None, on_solved=None):
        self.b = backend
        # every recount here is followed by uncovered() and most gain maps
        # by map_max(): let the transforms fold those censuses
        if hasattr(backend, "set_fold"):
            backend.set_fold(True)
        self.rng = np.random.default_rng(seed * 1000003 + 12345)
        self.deadline = deadline
        self.on_log = on_log or (lambda s: None)
        #: called after every LNS round with (searcher, record) -- the move
        #: log the learned proposal trains on
        self.on_round = on_round
        #: called with (M, code) whenever a pass establishes a feasible code
        #: of a new size (the checkpoint hook, reference search.py:35, 188)
        self.on_solved = on_solved or (lambda m, code: None)
        self.code = np.zeros(0, dtype=np.int64)
        self.uncov = None
        self.best_code = None
        self.best_uncov = None
        self.stats = {"transforms": 0, "placements": 0, "removals": 0,
                      "rounds": 0, "recounts": 0, "restarts": 0}
        self._scan_off = 0
        self._deficit_cache = (None, None)
        self.loop_reserve_s = 10.0
        self.build_reserve_s = 5.0

    # ------------------------------------------------ state
    def time_left(self):
        return np.inf if self.deadline is None else self.deadline - time.time()

    def load(self, code):
        arr = np.asarray(code)
        if arr.ndim == 2:
            arr = self.b.word_index(arr.astype(np.int64))
        self.code = (np.unique(arr.astype(np.int64)) if arr.size
                     else np.zeros(0, dtype=np.int64))
        self.b.recount(self.code)
        self.stats["transforms"] += 1
        self.uncov = self.b.uncovered()
        self._snapshot()

    def _snapshot(self):
        if self.best_uncov is None or self.uncov < self.best_uncov or (
                self.uncov == self.best_uncov and
                (self.best_code is None or len(self.code) < len(self.best_code))):
            self.best_uncov = self.uncov
            self.best_code = self.code.copy()

    def verify_full(self):
        self.b.recount(self.code)
        self.stats["transforms"] += 1
        self.stats["recounts"] += 1
        self.uncov = self.b.uncovered()
        return self.uncov

    # ------------------------------------------------ lazy greedy
    def _refresh_heap(self, codeset, cap):
        """Gain map -> threshold slice -> stale-bound heap."""
        b = self.b
        b.gain_map()
        self.stats["transforms"] += 1
        mx = b.map_max()
        if mx <= 0:
            return None, 1
        thr = 1
        if mx > 1 and getattr(b, "supports_hist", True):
            hist = b.map_hist(1024, mx)
            tail = np.cumsum(hist[::-1])[::-1]
            ok = np.flatnonzero(tail <= cap * 0.9)
            bin0 = int(ok[0]) if len(ok) else 1023
            thr = max(1, bin0 * mx // 1024 + 1)
        idx, val, found = b.map_extract(thr, cap=cap)
        while found > cap and thr < mx:
            thr = min(mx, thr + max(1, (mx - thr) // 2))
            idx, val, found = b.map_extract(thr, cap=cap)
        idx = np.asarray(idx, dtype=np.int64)
        keys = -np.asarray(val, dtype=np.int64)
        keep = ~np.isin(idx, np.fromiter(codeset, dtype=np.int64, count=len(codeset)))
        idx, keys = idx[keep], keys[keep]
        order = np.lexsort((idx, keys))
        return _StaleQueue(keys[order], idx[order]), thr

    def greedy_fill(self, target_m, batch=64, refresh_every=512):
        """Exact lazy greedy until covered / target_m / deadline: a
        placement is taken only when its exact gain (ball_gather role 0)
        is at least the next stale bound.  Returns the uncovered count."""
        b = self.b
        codeset = set(self.code.tolist())
        since = refresh_every
        heap, thr = _StaleQueue.empty(), 1
        cap = min(b.extract_cap, 1 << 18)
        while self.uncov > 0 and len(codeset) < target_m:
            if self.time_left() < self.build_reserve_s:
                break
            if since >= refresh_every or not heap:
                heap, thr = self._refresh_heap(codeset, cap)
                if heap is None:
                    break
                since = 0
                if not heap:
                    if thr > 1:
                        since, thr = refresh_every, 1
                        continue
                    break
            cand = []
            while heap and len(cand) < batch:
                v, i = heap.pop()
                if i not in codeset:
                    cand.append((-v, i))
            if not cand:
                since = refresh_every
                continue
            widx = np.fromiter((c[1] for c in cand), dtype=np.int64,
                               count=len(cand))
            exact = b.ball_gather(widx, 0)
            bi = int(np.argmax(exact))
            bg = int(exact[bi])
            ex = exact.tolist()
            bound = -heap.top_key() if heap else 0
            if bg < 1:
                since = refresh_every
                continue
            if bg >= bound and bg >= thr:
                w = int(widx[bi])
                b.ball_update(np.array([w], dtype=np.int64), +1)
                codeset.add(w)
                self.uncov -= bg
                self.stats["placements"] += 1
                since += 1
                for k, (v, i) in enumerate(cand):
                    if i != w and ex[k] > 0:
                        heap.push((-ex[k], i))
            else:
                for k, (v, i) in enumerate(cand):
                    if ex[k] > 0:
                        heap.push((-ex[k], i))
                if bg < thr:
                    since = refresh_every
        self.code = np.array(sorted(codeset), dtype=np.int64)
        return self.uncov

    # ------------------------------------------------ deficit words
    def uncovered_words(self, limit=DEFICIT_WORDS, chunk=SCAN_CHUNK):
        """Up to ``limit`` deficient word indices, scanned from a rotating
        offset and cached against the solution (search.py:462-529)."""
        if not hasattr(self.b, "read_cnt") or self.uncov == 0:
            return np.zeros(0, dtype=np.int64)
        key = (len(self.code), self.uncov,
               hash(np.ascontiguousarray(self.code).tobytes()))
        ckey, cval = self._deficit_cache
        if ckey == key and cval is not None and len(cval) >= min(limit,
                                                                 self.uncov):
            return cval[:limit]
        size = self.b.space.size
        mu = getattr(self.b, "mu", 1)
        found = []
        off = int(self._scan_off) % size
        read = 0
        while read < size and len(found) < limit:
            take = min(chunk, size - off, size - read)
            hit = np.flatnonzero(np.asarray(self.b.read_cnt(off, take)) < mu)
            if len(hit):
                got = hit[:limit - len(found)] + off
                found.extend(int(v) for v in got)
                self._scan_off = (int(got[-1]) + 1) % size
            read += take
            off = (off + take) % size
        if len(found) < limit:
            self._scan_off = 0
        out = np.array(sorted(found), dtype=np.int64)
        if len(out):
            self._deficit_cache = (key, out)
        return out

    # ------------------------------------------------ ruin
    def losses(self):
        self.b.loss_map()
        self.stats["transforms"] += 1
        return self.b.map_read_at(self.code)

    def _near(self, centre_idx, m):
        digs = self.b.index_word(self.code).astype(np.int16)
        c = self.b.index_word(np.array([centre_idx]))[0].astype(np.int16)
        return _digit_distance(digs, c)

    def ruin_pick(self, k, mode):
        """Positions (into self.code) of the k incumbents a ruin removes."""
        m = len(self.code)
        if mode == "deficit":
            tgt = self.uncovered_words()
            if not len(tgt):
                return self.ruin_pick(k, "hole")
            self.stats["deficit_ruins"] = self.stats.get("deficit_ruins", 0) + 1
            digs = self.b.index_word(self.code).astype(np.int16)
            d = np.full(m, np.iinfo(np.int32).max, dtype=np.int64)
            for t in self.b.index_word(tgt).astype(np.int16):
                np.minimum(d, _digit_distance(digs, t), out=d)
            order = np.argsort(d + self.rng.random(m))
            near = order[:min(m, max(3 * k, k + 8))]
            return self.rng.choice(near, size=k, replace=False)
        if mode == "random":
            return self.rng.choice(m, size=k, replace=False)
        if mode == "cluster":
            digs = self.b.index_word(self.code).astype(np.int16)
            c = digs[self.rng.integers(m)]
            return np.argsort(_digit_distance(digs, c) + self.rng.random(m))[:k]
        if mode == "hole":
            self.b.gain_map()
            self.stats["transforms"] += 1
            mx = self.b.map_max()
            if mx <= 0:
                return self.ruin_pick(k, "low")
            oi, _ov, _found = self.b.map_extract(max(1, mx // 2), cap=4096)
            if not len(oi):
                return self.ruin_pick(k, "low")
            anchor = int(oi[self.rng.integers(len(oi))])
            return np.argsort(self._near(anchor, m) + self.rng.random(m))[:k]
        lo = self.losses().astype(np.float64)          # "low"
        jit = self.rng.exponential(scale=max(1.0, lo.mean() * 0.15), size=m)
        return np.argsort(lo + jit)[:k]

    def remove_at(self, pick):
        removed = self.code[pick]
        self.b.ball_update(removed, -1)
        self.code = np.delete(self.code, pick)
        self.uncov = self.b.uncovered()
        self.stats["removals"] += int(len(pick))
        return removed

    def ruin(self, k, mode="low"):
        """Remove k incumbents chosen by ``mode``; returns their indices."""
        k = min(k, len(self.code) - 1)
        if k <= 0:
            return np.zeros(0, dtype=np.int64)
        return self.remove_at(self.ruin_pick(k, mode))

    # ------------------------------------------------ peel
    def peel(self):
        """Drop zero-loss incumbents to a fixpoint, in pairwise independent
        batches (digit distance > indep_dist)."""
        b = self.b
        total = 0
        while self.time_left() >= self.build_reserve_s:
            lo = self.losses()
            zeros = np.flatnonzero(lo == 0)
            if not len(zeros):
                break
            exact = b.ball_gather(self.code[zeros], 1)
            zeros = zeros[exact == 0]
            if not len(zeros):
                break
            self.rng.shuffle(zeros)
            sep = b.indep_dist
            if sep is not None and sep < b.space.n:
                digs = b.index_word(self.code[zeros]).astype(np.int16)
                sel = [0]
                for i in range(1, len(zeros)):
                    d = (digs[np.array(sel)] != digs[i][None, :]).sum(axis=1)
                    if (d > sep).all():
                        sel.append(i)
                pick = zeros[np.array(sel)]
            else:
                pick = zeros[:1]
            b.ball_update(self.code[pick], -1)
            keep = np.ones(len(self.code), dtype=bool)
            keep[pick] = False
            self.code = self.code[keep]
            total += len(pick)
            self.stats["removals"] += len(pick)
        if total:
            self.uncov = self.b.uncovered()
        return total

    # ------------------------------------------------ LNS at fixed M
    def lns(self, M, round_budget_s=np.inf, kmin=8, kmax=None,
            select="fixed", max_rounds=None, deficit=False, deficit_hint=3,
            deficit_boost=4.0, proposer=None, operator=None):
        """Ruin/recreate rounds at fixed M until feasible, the budget or
        ``max_rounds`` (search.py:702-784).  Rejects a worse round with
        probability 0.9, restoring the best snapshot; re-verifies the field
        every 50 rounds.  Returns the uncovered count.

        ``proposer`` (callable(searcher, k, mode) -> positions in the code)
        adds a "learned" operator: weighted like ``hole`` in the fixed mix,
        one more arm of the bandit under ``select="adaptive"``.
        ``operator`` replaces the ruin's region choice of the drawn mode
        (the move logger uses it); without either the pass draws exactly
        as the reference does."""
        if select not in ("fixed", "adaptive"):
            raise ValueError(select)
        t0 = time.time()
        kmax = kmax or max(kmin + 1, M // 20)
        self.best_uncov = None
        self.best_code = None
        if len(self.code) < M:
            self.greedy_fill(M)
        self._snapshot()
        draw = FIXED_DRAW_DEFICIT if deficit else FIXED_DRAW
        ops, prior = (RUIN_MODES_EXT, RUIN_MIX_EXT) if deficit else (RUIN_MODES, RUIN_MIX)
        if proposer is not None:
            draw = draw + ("learned", "learned")
            ops = ops + ("learned",)
            prior = dict(prior, learned=float(getattr(proposer, "prior", 2.0)))
        bandit = None
        if select == "adaptive":
            bandit = (OperatorBandit(self.rng, ops=ops, prior=prior)
                      if (deficit or proposer is not None)
                      else OperatorBandit(self.rng))
        hint = {"deficit": deficit_boost} if deficit else None
        r = 0
        while self.uncov > 0:
            if time.time() - t0 > round_budget_s or \
                    self.time_left() < self.loop_reserve_s:
                break
            if max_rounds is not None and r >= max_rounds:
                break
            pre, pre_best = self.uncov, self.best_uncov
            k = int(self.rng.integers(kmin, kmax + 1))
            near = deficit and self.uncov <= deficit_hint
            if bandit is None:
                if near and self.rng.random() < 0.5:
                    mode = "deficit"
                else:
                    mode = draw[int(self.rng.integers(len(draw)))]
            else:
                mode = bandit.select(hint if near else None)
            kk = min(k, len(self.code) - 1)
            if mode == "learned" or operator is not None:
                pick = np.zeros(0, dtype=np.int64)
                if kk > 0:
                    fn = proposer if mode == "learned" else operator
                    pick = np.asarray(fn(self, kk, mode), dtype=np.int64)
                    self.remove_at(pick)
            else:
                before = self.code
                self.ruin(k, mode)
                pick = np.flatnonzero(~np.isin(before, self.code))
            self.greedy_fill(M)
            self.stats["rounds"] += 1
            if mode == "learned":
                self.stats["learned_rounds"] = self.stats.get("learned_rounds", 0) + 1
            r += 1
            accepted = True
            if self.uncov > pre and self.rng.random() > 0.10:
                accepted = False
                self.code = self.best_code.copy()
                self.b.recount(self.code)
                self.stats["transforms"] += 1
                self.uncov = self.b.uncovered()
            else:
                self._snapshot()
            if self.on_round is not None:
                self.on_round(self, dict(mode=mode, k=int(len(pick)),
                                         pre=int(pre), post=int(self.uncov),
                                         accepted=accepted, pick=pick))
            if bandit is not None:
                if self.best_uncov < pre_best:
                    reward = bandit.REWARD_BEST
                elif self.uncov < pre:
                    reward = bandit.REWARD_BETTER
                elif self.uncov == pre:
                    reward = bandit.REWARD_EQUAL
                else:
                    reward = bandit.REWARD_WORSE
                bandit.update(mode, reward)
                self.stats["ops"] = bandit.summary()
            if r % 50 == 0:
                u_inc = self.uncov
                if self.verify_full() != u_inc:
                    self.on_log(f"drift: incremental {u_inc} != recount "
                                f"{self.uncov}")
        if self.uncov == 0:
            self._snapshot()
        elif self.best_code is not None and self.best_uncov < self.uncov:
            self.code = self.best_code.copy()
            self.b.recount(self.code)
            self.uncov = self.b.uncovered()
        return self.uncov

    def portfolio(self, M, budget_s=np.inf, streams=4, slice_s=None, kmin=8,
                  kmax=None, ruin_frac=0.15, select="fixed", slice_rounds=None,
                  max_restarts=None, proposer=None):
        """Time-sliced restart portfolio at fixed M with a shared incumbent
        (reference search.py:1536-1620): ``streams`` solutions -- the current
        one and perturbations of it (a kick of ruin_frac * M incumbents by
        each ruin operator in turn, then the exact greedy refill) -- each
        given an LNS slice per round, round-robin, on the one backend (a
        stream is resumed by a recount of its code).  The best state seen is
        the incumbent; after every round the worst stream is replaced by a
        fresh perturbation of it.  Stops when a stream is feasible, at the
        budget, or -- clock-free runs -- after ``max_restarts`` restarts,
        with slices of ``slice_rounds`` LNS rounds instead of ``slice_s``
        seconds.  ``proposer``: a learned ruin proposal, one more operator of
        every stream's LNS (as in :meth:`lns`).  Leaves the incumbent
        loaded; returns its uncovered count."""
        t0 = time.time()
        streams = max(1, int(streams))
        slice_s = slice_s or max(1.0, budget_s / (3.0 * streams))
        if len(self.code) > M:
            self.ruin(len(self.code) - M, "low")
        if len(self.code) < M and self.uncov > 0:
            self.greedy_fill(M)
        kick = max(1, int(ruin_frac * M))
        base = self.code.copy()

        def fresh(code, mode):
            self.load(code)
            self.ruin(kick, mode)
            self.greedy_fill(M)
            return self.code.copy(), self.uncov

        pop = [(base, self.uncov)]
        if self.uncov > 0:
            pop += [fresh(base, RUIN_MODES[i % len(RUIN_MODES)])
                    for i in range(1, streams)]
        inc, inc_u = min(pop, key=lambda cu: cu[1])

        def out_of_time():
            return (time.time() - t0 > budget_s or
                    self.time_left() < self.loop_reserve_s)

        restarts = 0
        while inc_u > 0 and not out_of_time():
            for i in range(len(pop)):
                left = budget_s - (time.time() - t0)
                if left <= 0 or self.time_left() < self.loop_reserve_s:
                    break
                self.load(pop[i][0])
                u = self.lns(M, round_budget_s=min(slice_s, left), kmin=kmin,
                             kmax=kmax, select=select, max_rounds=slice_rounds,
                             proposer=proposer)
                pop[i] = (self.code.copy(), u)
                if u < inc_u:
                    inc, inc_u = pop[i]
                if u == 0:
                    break
            if inc_u == 0 or time.time() - t0 > budget_s:
                break
            if max_restarts is not None and restarts >= max_restarts:
                break
            worst = int(np.argmax([u for _, u in pop]))
            self.stats["restarts"] += 1
            restarts += 1
            mode = RUIN_MODES[self.stats["restarts"] % len(RUIN_MODES)]
            pop[worst] = fresh(inc, mode)
            if pop[worst][1] < inc_u:
                inc, inc_u = pop[worst]
            self.on_log(f"portfolio M={M} streams={len(pop)} best={inc_u} "
                        f"restart=s{worst}/{mode}")
        self.best_uncov = None
        self.best_code = None
        self.load(inc)
        return self.uncov


def descend_rounds(searcher, rounds, kmin=8, select="fixed"):
    """Clock-free notch descent: from a feasible code of size M, drop one
    incumbent (exact-loss ruin) and run LNS at M - 1 until it is feasible
    again, then notch further; stops after ``rounds`` LNS rounds in total.
    Returns (rounds used, [(rounds so far, feasible size), ...])."""
    s = searcher
    used = 0
    sizes = [(0, len(s.code))]
    while used < rounds and s.uncov == 0 and len(s.code) > 1:
        M = len(s.code) - 1
        before = s.stats["rounds"]
        s.ruin(1, "low")
        s.lns(M, kmin=kmin, select=select, max_rounds=rounds - used)
        used += s.stats["rounds"] - before
        if s.uncov == 0:
            sizes.append((used, len(s.code)))
            s.on_solved(len(s.code), s.code.copy())
    return used, sizes


def config1_run(backend, start, rounds=1000, seed=1):
    """K_3(8,2)-style run: load the start code, read the three fields, lazy
    greedy to a cover, peel, then ``rounds`` LNS rounds of notch descent.
    Returns the move trace (per-round mode / uncovered / size) and the final
    best code."""
    trace = []
    s = Searcher(backend, seed=seed,
                 on_round=lambda s, r: trace.append((r["mode"], r["post"],
                                                     len(s.code))))
    s.load(np.asarray(start, dtype=np.int64))
    backend.gain_map()
    g_max = backend.map_max()
    backend.loss_map()
    l_max = backend.map_max()
    s.greedy_fill(10 ** 9)
    greedy_size = len(s.code)
    s.peel()
    used, sizes = descend_rounds(s, rounds)
    best = s.best_code if s.uncov != 0 and s.best_code is not None else s.code
    return dict(fields=(int(g_max), int(l_max)), greedy=greedy_size,
                peeled=sizes[0][1], sizes=sizes, rounds=used, trace=trace,
                final=np.asarray(sorted(int(x) for x in s.code)),
                uncov=int(s.uncov))
