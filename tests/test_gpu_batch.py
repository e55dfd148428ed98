"""Batched multi-cell engine: every cell's fields bit-exact against the CPU
oracle after one batched transform (a persistent launch per sweep level),
device greedy and LNS trajectories
identical to a CPU restatement of the same decisions, and the captured HIP
graph identical to direct launches (replayed many times in a row).

The batch has no counterpart in the reference (one cell per process,
reference docs/PLUGIN.md:73-76); each cell's fields are the reference's
dct_transform fields (reference src/include/coldcore/plugin.h:90-104) and
are checked against the same oracle as the singleton engine.
"""
import numpy as np
import pytest

from oracle.dp import NumpyDP

pytestmark = pytest.mark.gpu

# uneven cells: all four families, mixed radix, multi-sweep cells (2^14 and
# 2^16 words), and king tori / grid boxes with axes beyond 16 (the runtime
# window path)
BATCH_CELLS = [
    ("hamming", [2] * 10, 2),
    ("hamming", [3] * 6, 3),
    ("hamming", [2, 3, 4, 5], 2),
    ("hamming", [2] * 14, 1),
    ("hamming", [2] * 16, 3),
    ("hamming", [10, 9, 8], 3),
    ("lee", [5, 5, 5], 2),
    ("lee", [7, 7], 3),
    ("lee", [4, 5, 6], 3),
    ("lee", [16, 16], 3),
    ("torus_linf", [9, 9], 1),
    ("torus_linf", [7, 7, 7], 1),
    ("torus_linf", [20, 21], 3),
    ("torus_linf", [31, 31], 1),
    ("grid_linf", [8, 8], 1),
    ("grid_linf", [5, 6, 4], 2),
    ("grid_linf", [17, 5], 3),
    ("grid_linf", [2, 3, 4, 6, 8], 1),
    # axes far beyond 32: the runtime-length window in the sweeps
    ("torus_linf", [100, 7], 3),
    ("grid_linf", [200, 3], 3),
]


def _batch(cells, cap=4096):
    from coldcore_amd.batch import GpuBatch
    return GpuBatch(cells, cap=cap)


def _codes(cells, frac=0.02, seed=0):
    out = []
    for i, (_, axes, _) in enumerate(cells):
        size = int(np.prod(axes))
        rng = np.random.default_rng(seed * 1000 + i)
        m = max(2, int(size * frac))
        out.append(rng.integers(0, size, m).astype(np.int64))   # a multiset
    return out


@pytest.mark.parametrize("threads", ["256", "512", "1024"])
def test_batched_transform_matches_oracle(monkeypatch, threads):
    """Recount, gain and loss of every cell from one launch each (level
    launches of 256, 512 and 1,024 threads per tile)."""
    monkeypatch.setenv("CCB_THREADS", threads)
    codes = _codes(BATCH_CELLS)
    with _batch(BATCH_CELLS) as b:
        info = b.info()
        assert info["threads"] == int(threads)
        assert info["cells"] == len(BATCH_CELLS)
        assert info["sweeps"] > len(BATCH_CELLS)     # multi-sweep cells
        assert info["levels"] >= 2
        for i, c in enumerate(codes):
            b.set_code(i, c)
        b.recount()
        oracles = []
        for i, (p, axes, R) in enumerate(BATCH_CELLS):
            o = NumpyDP(axes, R, problem=p)
            o.recount(codes[i])
            oracles.append(o)
            np.testing.assert_array_equal(b.read_cnt(i), o.cnt.astype(np.uint16),
                                          err_msg=f"cnt cell {i} {p} {axes} R{R}")
        b.gain_maps()
        for i, o in enumerate(oracles):
            o.gain_map()
            np.testing.assert_array_equal(b.read_map(i), o.map,
                                          err_msg=f"gain cell {i}")
        b.loss_maps()
        for i, o in enumerate(oracles):
            o.loss_map()
            np.testing.assert_array_equal(b.read_map(i), o.map,
                                          err_msg=f"loss cell {i}")


def test_batch_autotune_keeps_fields_exact():
    """The schedule autotuner only changes the plan: after it, every cell's
    recount, gain and loss still equal the oracle."""
    cells = BATCH_CELLS[::2]
    codes = _codes(cells, seed=5)
    with _batch(cells) as b:
        for i, c in enumerate(codes):
            b.set_code(i, c)
        b.recount()
        budget, us = b.autotune(reps=2)
        assert budget in (8192, 16384, 32768, 65536, 131072) and us > 0
        assert b.info()["lds_budget"] == budget
        assert b.info()["threads"] in (256, 512, 1024)
        b.recount()
        for which in ("recount", "gain", "loss"):
            if which == "gain":
                b.gain_maps()
            elif which == "loss":
                b.loss_maps()
            for i, (p, axes, R) in enumerate(cells):
                o = NumpyDP(axes, R, problem=p)
                o.recount(codes[i])
                if which == "recount":
                    np.testing.assert_array_equal(b.read_cnt(i),
                                                  o.cnt.astype(np.uint16))
                else:
                    getattr(o, which + "_map")()
                    np.testing.assert_array_equal(b.read_map(i), o.map,
                                                  err_msg=f"{which} cell {i}")


def test_passenger_rows_and_round_tiles_exact(monkeypatch):
    """A 64 KiB budget gives contiguous first-sweep tiles of several groups
    (passenger rows: a round tile of 2 x 4,096 binary words, generic tiles
    of 6 x 840 and 2 x 3,024 words): recount, gain and loss still equal the
    oracle."""
    monkeypatch.setenv("CCB_LDS", "65536")
    cells = [("hamming", [2] * 14, 1), ("hamming", [4, 5, 6, 7, 8, 9], 2),
             ("grid_linf", [6, 7, 8, 9, 10], 1), ("hamming", [3] * 9, 3)]
    codes = _codes(cells, frac=0.01, seed=9)
    with _batch(cells) as b:
        assert b.info()["lds_budget"] == 65536
        for i, c in enumerate(codes):
            b.set_code(i, c)
        b.recount()
        oracles = []
        for i, (p, axes, R) in enumerate(cells):
            o = NumpyDP(axes, R, problem=p)
            o.recount(codes[i])
            oracles.append(o)
            np.testing.assert_array_equal(b.read_cnt(i), o.cnt.astype(np.uint16))
        for which in ("gain", "loss"):
            getattr(b, which + "_maps")()
            for i, o in enumerate(oracles):
                getattr(o, which + "_map")()
                np.testing.assert_array_equal(b.read_map(i), o.map,
                                              err_msg=f"{which} cell {i}")


def test_batched_transform_repeats_exactly():
    """The persistent work list resets itself: repeated launches agree."""
    cells = BATCH_CELLS[:8]
    codes = _codes(cells, seed=3)
    with _batch(cells) as b:
        for i, c in enumerate(codes):
            b.set_code(i, c)
        b.recount()
        first = [b.read_cnt(i) for i in range(len(cells))]
        for _ in range(5):
            b.recount()
        for i in range(len(cells)):
            np.testing.assert_array_equal(b.read_cnt(i), first[i])


# ------------------------------------------------------------------ search
M64 = (1 << 64) - 1


def _mix(x):
    """splitmix64 finaliser, as the device LNS control kernels compute it."""
    x = (x + 0x9E3779B97F4A7C15) & M64
    x = ((x ^ (x >> 30)) * 0xBF58476D1CE4E5B9) & M64
    x = ((x ^ (x >> 27)) * 0x94D049BB133111EB) & M64
    return x ^ (x >> 31)


class CpuCell:
    """CPU restatement of one cell's device greedy / LNS decisions, on the
    oracle's exact fields."""

    def __init__(self, p, axes, R, cap):
        self.o = NumpyDP(axes, R, problem=p)
        self.code = []
        self.cap = cap
        self.o.recount(np.zeros(0, dtype=np.int64))

    def uncov(self):
        return int((self.o.cnt == 0).sum())

    def argmax_gain(self):
        self.o.gain_map()
        g = self.o.map
        i = int(np.argmax(g))
        return i if g[i] > 0 else None

    def add(self, w):
        self.code.append(w)
        self.o.ball_update([w], +1)

    def remove(self, w):
        self.o.ball_update([w], -1)

    def greedy_step(self):
        if self.uncov() and len(self.code) < self.cap:
            w = self.argmax_gain()
            if w is not None:
                self.add(w)

    def lns_init(self, seed, ci):
        self.un = self.uncov()
        self.rng = _mix((seed * 0x100000001B3 + ci) & M64)
        self.best = list(self.code) if self.un == 0 else None
        self.rounds = self.accepted = self.reverted = 0

    def _next(self):
        self.rng = _mix(self.rng)
        return self.rng

    def lns_round(self, k):
        pre = self.un
        M = len(self.code)
        kk = min(k, M - 1)
        nrem = max(kk, 0)
        budget = 0 if kk <= 0 else (kk - 1 if self.un == 0 else kk)
        rem = []
        for j in range(nrem):
            top = M - 1 - j
            r = self._next() % (top + 1)
            self.code[r], self.code[top] = self.code[top], self.code[r]
            rem.append(self.code[top])
        del self.code[M - nrem:]
        for w in rem:
            self.remove(w)
        added = []
        for j in range(budget):
            w = self.argmax_gain()
            if w is None or len(self.code) >= self.cap:
                continue
            self.add(w)
            added.append(w)
        now = self.uncov()
        revert = False
        if nrem > 0 and now > pre:
            u = (self._next() >> 11) * (1.0 / 9007199254740992.0)
            revert = u > 0.10
        if revert:
            del self.code[len(self.code) - len(added):]
            for w in added:
                self.remove(w)
            for w in rem:
                self.add(w)
            self.un = pre
            self.reverted += 1
        else:
            self.un = now
            if nrem > 0:
                self.accepted += 1
            if now == 0 and (self.best is None or len(self.code) < len(self.best)):
                self.best = list(self.code)
        self.rounds += 1


SEARCH_CELLS = [
    ("hamming", [3] * 5, 1),
    ("hamming", [2] * 8, 2),
    ("lee", [5, 5], 1),
    ("lee", [6, 7], 2),
    ("torus_linf", [9, 9], 1),
    ("torus_linf", [19, 17], 1),
    ("grid_linf", [6, 6], 1),
    ("grid_linf", [4, 3, 5], 1),
]


def _greedy_all(b, cpus, steps):
    b.greedy(steps)
    for c in cpus:
        for _ in range(steps):
            c.greedy_step()


def test_greedy_matches_cpu_restatement():
    cap = 512
    cpus = [CpuCell(p, a, r, cap) for p, a, r in SEARCH_CELLS]
    with _batch(SEARCH_CELLS, cap=cap) as b:
        for i in range(len(SEARCH_CELLS)):
            b.set_code(i, np.zeros(0, dtype=np.int64))
        b.recount()
        for chunk in (5, 20, 200):
            _greedy_all(b, cpus, chunk)
            for i, c in enumerate(cpus):
                assert b.code(i).tolist() == c.code, f"cell {i} after greedy"
                np.testing.assert_array_equal(b.read_cnt(i),
                                              c.o.cnt.astype(np.uint16))
        for i, c in enumerate(cpus):
            assert c.uncov() == 0, f"greedy left cell {i} uncovered"


@pytest.mark.parametrize("graph", [True, False], ids=["graph", "direct"])
def test_lns_matches_cpu_restatement(graph):
    """Rounds of ruin k / recreate / keep-or-revert: codes, cnt, counters
    and best codes equal the CPU restatement after every batch of rounds;
    the graph replays 3 x 10 rounds with host reads in between."""
    cap, k, seed = 512, 3, 7
    cpus = [CpuCell(p, a, r, cap) for p, a, r in SEARCH_CELLS]
    with _batch(SEARCH_CELLS, cap=cap) as b:
        for i in range(len(SEARCH_CELLS)):
            b.set_code(i, np.zeros(0, dtype=np.int64))
        b.recount()
        _greedy_all(b, cpus, 300)
        b.lns_init(seed)
        for i, c in enumerate(cpus):
            c.lns_init(seed, i)
        for _ in range(3):
            b.lns_run(10, k=k, graph=graph)
            for c in cpus:
                for _ in range(10):
                    c.lns_round(k)
            for i, c in enumerate(cpus):
                st = b.stats(i)
                assert b.code(i).tolist() == c.code, f"cell {i} code"
                np.testing.assert_array_equal(b.read_cnt(i),
                                              c.o.cnt.astype(np.uint16))
                assert st["uncovered"] == c.un
                assert (st["rounds"], st["accepted"], st["reverted"]) == \
                    (c.rounds, c.accepted, c.reverted)
                best = b.code(i, best=True)
                if c.best is None:
                    assert st["best"] == -1
                else:
                    assert st["best"] == len(c.best)
                    assert best.tolist() == c.best


def test_lns_best_codes_cover():
    """Invariants on a bigger batch: cnt is the recount of the current code,
    uncovered matches cnt, every best code covers its space and is no
    larger than the greedy start."""
    cells = [("hamming", [3] * n, R) for n in (6, 7) for R in (1, 2)] + \
        [("lee", [9, 9, 9], 2), ("torus_linf", [23, 23], 1),
         ("grid_linf", [16, 12, 9], 1), ("grid_linf", [2] * 8, 1)]
    with _batch(cells, cap=4096) as b:
        for i in range(len(cells)):
            b.set_code(i, np.zeros(0, dtype=np.int64))
        b.recount()
        b.greedy(2000)
        start = [b.stats(i)["M"] for i in range(len(cells))]
        b.lns_init(11)
        b.lns_run(40, k=4, graph=True)
        for i, (p, axes, R) in enumerate(cells):
            st = b.stats(i)
            o = NumpyDP(axes, R, problem=p)
            o.recount(b.code(i))
            cnt = b.read_cnt(i)
            np.testing.assert_array_equal(cnt, o.cnt.astype(np.uint16))
            assert st["uncovered"] == int((cnt == 0).sum())
            assert st["rounds"] == 40
            best = b.code(i, best=True)
            assert 0 < st["best"] <= start[i]
            o.recount(best)
            assert int((o.cnt == 0).sum()) == 0, f"best code of cell {i}"


def test_batch64_workload_matches_oracle():
    """The 64-cell batched workload exactly as benchmarked (Hamming K_3(n,R)
    for n = 8..12 and R = 1..3, Lee 9x9x9, king tori 9..31, 25 grid boxes up
    to 2^20 words; the batch's autotuned plan; seeded random starts at 1.5x
    the device greedy size): recount, gain and loss of all 64 cells equal
    the oracle, then 20 LNS rounds equal the CPU restatement word for word
    on every cell of at most 2^15 words."""
    import bench
    from coldcore_amd.batch import GpuBatch
    cells = bench.batch_cells64()
    assert len(cells) == 64
    with GpuBatch(cells, cap=65535) as b:
        b.autotune(reps=2)
        empty = np.zeros(0, dtype=np.int64)
        for i in range(len(cells)):
            b.set_code(i, empty)
        b.recount()
        b.greedy_cover(256)
        starts = []
        for i, (_, a, _) in enumerate(cells):
            size = int(np.prod(a))
            greedy = b.stats(i)["M"]
            assert b.stats(i)["uncovered"] == 0
            m = min(size, 65535, int(round(1.5 * greedy)))
            starts.append(np.random.default_rng(i).choice(size, m, replace=False)
                          .astype(np.int64))
        for i, c in enumerate(starts):
            b.set_code(i, c)
        b.recount()
        oracles = []
        for i, (p, axes, R) in enumerate(cells):
            o = NumpyDP(axes, R, problem=p)
            o.recount(starts[i])
            oracles.append(o)
            np.testing.assert_array_equal(b.read_cnt(i), o.cnt.astype(np.uint16),
                                          err_msg=f"cnt cell {i} {p} {axes} R{R}")
        for which in ("gain", "loss"):
            getattr(b, which + "_maps")()
            for i, o in enumerate(oracles):
                getattr(o, which + "_map")()
                np.testing.assert_array_equal(b.read_map(i), o.map,
                                              err_msg=f"{which} cell {i} {cells[i]}")
        # LNS from the same starts: the device runs all 64 cells, the CPU
        # restatement follows the small ones (cells are independent)
        small = [i for i, (_, a, _) in enumerate(cells) if int(np.prod(a)) <= 1 << 15]
        assert len(small) >= 30
        cpus = {}
        for i in small:
            p, a, R = cells[i]
            c = CpuCell(p, a, R, 65535)
            c.code = starts[i].tolist()
            c.o.recount(starts[i])
            cpus[i] = c
        seed, k = 0, 2
        b.lns_init(seed)
        for i, c in cpus.items():
            c.lns_init(seed, i)
        b.lns_run(20, k=k, graph=True)
        for i, c in cpus.items():
            for _ in range(20):
                c.lns_round(k)
            st = b.stats(i)
            assert b.code(i).tolist() == c.code, f"cell {i} {cells[i]} code"
            np.testing.assert_array_equal(b.read_cnt(i), c.o.cnt.astype(np.uint16))
            assert st["uncovered"] == c.un
            assert (st["rounds"], st["accepted"], st["reverted"]) == \
                (c.rounds, c.accepted, c.reverted)
