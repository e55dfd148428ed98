"""Full-size parity on the benchmark cells, by size-independent checks.

* K_2(24,3) (2^24 words): cnt and the gain field against the integer
  Walsh-Hadamard convolution (an oracle that shares nothing with the axis
  DP), plus the translation-invariance sums sum cnt = M |B| and
  sum gain = |B| #{cnt < mu}.
* K_5(10,3) and the mixed-radix cell [4,5,6,7,8,4,5,6,7] R=3 against the
  numpy axis-DP oracle (a few seconds each on one core).
* walks at full size: gain_map == ball_gather(., 0) on sampled words and
  update-then-undo is the identity.
"""
import numpy as np
import pytest

from oracle.dp import NumpyDP
from oracle.walsh import binary_field
from coldcore_amd.patterns import ball_size

pytestmark = pytest.mark.gpu


def _code(size, p, seed=0):
    return np.flatnonzero(np.random.default_rng(seed).random(size) < p).astype(np.int64)


def test_k2_24_3_against_walsh_hadamard(gpu_cell):
    n, R = 24, 3
    size = 1 << n
    code = _code(size, 1e-4)
    g = gpu_cell([2] * n, R)
    g.recount(code)
    cnt = g.read_cnt()
    ind = np.zeros(size, dtype=np.int64)
    np.add.at(ind, code, 1)
    want = binary_field(ind, n, R)
    assert (cnt.astype(np.int64) == want).all()
    B = ball_size(2, n, R)
    assert int(cnt.sum(dtype=np.int64)) == len(code) * B
    g.gain_map()
    gain = g.map_all()
    want_gain = binary_field((want < 1).astype(np.int64), n, R)
    assert (gain.astype(np.int64) == want_gain).all()
    assert int(gain.sum(dtype=np.int64)) == B * int((want < 1).sum())
    g.loss_map()
    loss = g.map_all()
    want_loss = binary_field((want == 1).astype(np.int64), n, R)
    assert (loss.astype(np.int64) == want_loss).all()
    # walks against the field at sampled words
    rng = np.random.default_rng(2)
    probe = rng.integers(0, size, 4096).astype(np.int64)
    g.gain_map()
    assert (g.ball_gather(probe, 0) == g.map_read_at(probe)).all()
    w = probe[:64]
    g.ball_update(w, +1)
    g.ball_update(w, -1)
    assert (g.read_cnt() == cnt).all()


@pytest.mark.parametrize("axes,R", [([5] * 10, 3),
                                    ([4, 5, 6, 7, 8, 4, 5, 6, 7], 3)],
                         ids=["K5(10,3)", "mixed9-R3"])
def test_large_cells_against_numpy_oracle(gpu_cell, axes, R):
    c = NumpyDP(axes, R)
    size = c.space.size
    code = _code(size, 1e-4)
    g = gpu_cell(axes, R)
    g.recount(code)
    c.recount(code)
    assert (g.read_cnt() == c.read_cnt()).all()
    g.gain_map()
    c.gain_map()
    assert (g.map_all() == c.map).all()
    g.loss_map()
    c.loss_map()
    assert (g.map_all() == c.map).all()
    mx = c.map_max()
    assert g.map_max() == mx
    gi, gv, gf = g.map_extract(max(1, mx // 2), cap=1 << 18)
    ci, cv, cf = c.map_extract(max(1, mx // 2), cap=1 << 18)
    assert gf == cf and (gi == ci).all() and (gv == cv).all()
    assert (g.map_hist(1024, mx) == c.map_hist(1024, mx)).all()


def _ball_masks(n, R):
    """xor masks of the radius-R Hamming ball in Z_2^n (every popcount <= R)."""
    from itertools import combinations
    out = [0]
    for d in range(1, R + 1):
        for c in combinations(range(n), d):
            out.append(sum(1 << i for i in c))
    return np.array(out, dtype=np.int64)


@pytest.mark.parametrize("tuning", ["autotuned", "planner"])
def test_k2_30_4_full_size_invariants(gpu_cell, monkeypatch, tuning):
    """K_2(30,4), 2^30 words, the 40,000-word seeded code of the benchmark,
    under the autotuner's schedule and under the planner's own: the
    translation-invariance sums, a brute-force count at sampled words,
    brute-force gain / loss at a few words from the downloaded cnt, and the
    walks equal to the fields there."""
    if tuning == "planner":
        monkeypatch.setenv("CC_AUTOTUNE", "0")
    n, R = 30, 4
    size = 1 << n
    code = np.random.default_rng(0).choice(size, 40000, replace=False).astype(np.int64)
    masks = _ball_masks(n, R)
    B = ball_size(2, n, R)
    assert len(masks) == B == 31931
    g = gpu_cell([2] * n, R)
    g.recount(code)
    cnt = g.read_cnt()
    assert int(cnt.sum(dtype=np.int64)) == len(code) * B
    rng = np.random.default_rng(5)
    probe = np.concatenate([code[:64], rng.integers(0, size, 4096)]).astype(np.int64)
    want = np.empty(len(probe), dtype=np.int64)
    for i in range(0, len(probe), 128):
        x = (probe[i:i + 128, None] ^ code[None, :]).astype(np.uint64)
        want[i:i + 128] = (np.bitwise_count(x) <= R).sum(axis=1)
    assert (cnt[probe].astype(np.int64) == want).all()
    zeros = int((cnt == 0).sum())
    ones = int((cnt == 1).sum())
    assert g.count_eq(0) == zeros and g.count_eq(1) == ones
    few = probe[::256]
    g.gain_map()
    plan = g.sweep_plan2(0)
    print(f"K_2(30,4) {tuning} schedule:",
          [(s["first_axis"], s["axes"], s["kernel"], s["RR"], s["W"], s["ND"],
            s["bytes_in"], s["bytes_out"]) for s in plan], flush=True)
    tot = 0
    for off in range(0, size, 1 << 26):
        tot += int(g.map_all(off, 1 << 26).sum(dtype=np.int64))
    assert tot == B * zeros
    gain = g.map_read_at(probe)
    assert (g.ball_gather(probe, 0) == gain).all()
    for x, v in zip(few, gain[::256]):
        assert int((cnt[x ^ masks] == 0).sum()) == v
    g.loss_map()
    assert int(g.map_read_at(code).sum(dtype=np.int64)) == ones
    loss = g.map_read_at(probe)
    assert (g.ball_gather(probe, 1) == loss).all()
    for x, v in zip(few, loss[::256]):
        assert int((cnt[x ^ masks] == 1).sum()) == v


def test_k2_30_4_restart_portfolio(gpu_cell):
    """Config 5 as a workload: the 40,000-word start on K_2(30,4), then a
    short restart portfolio (8 streams time-sharing the one HBM-resident
    cell): the incumbent never has more words than M, its uncovered count
    is a from-scratch recount's and no worse than the start's."""
    from coldcore_amd.search import Searcher
    start = np.random.default_rng(0).choice(1 << 30, 40000, replace=False).astype(np.int64)
    g = gpu_cell([2] * 30, 4)
    s = Searcher(g, seed=1)
    s.load(start)
    u0 = s.uncov
    assert u0 == g.count_eq(0) > 0
    u = s.portfolio(40000, streams=8, slice_rounds=1, max_restarts=2, kmin=8,
                    kmax=64, ruin_frac=0.01)
    print(f"K_2(30,4) portfolio: uncovered {u0} -> {u}, rounds "
          f"{s.stats['rounds']}, restarts {s.stats['restarts']}", flush=True)
    assert len(s.code) <= 40000 and u <= u0
    assert s.verify_full() == u
