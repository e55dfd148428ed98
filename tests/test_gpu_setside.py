"""Set-side gathers and the swap-evaluation fast path: bit-exact against the
oracle and against the support walk, near a cover and on random states,
with the predicate list invalidated by every cnt change and both overflow
fallbacks taken (reference tests/test_gpu_fast_paths.py:70-111, 225-292
test the same behaviour of the reference's set side and swap_eval).
"""
import numpy as np
import pytest

from oracle.dp import NumpyDP

pytestmark = pytest.mark.gpu

CELLS = [
    ("hamming", [3] * 6, 2),
    ("hamming", [2] * 12, 3),
    ("hamming", [2, 3, 4, 5, 6], 2),
    ("hamming", [9, 10, 7], 2),
    ("lee", [5, 6, 7], 2),
    ("lee", [8, 8], 3),
    ("torus_linf", [7, 9, 11], 2),
    ("grid_linf", [4, 5, 6], 3),
    ("grid_linf", [9, 9], 1),
]
IDS = [f"{p}-{'x'.join(map(str, a))}-R{r}" for p, a, r in CELLS]


def _greedy_cover(c):
    """A covering code by exact greedy on the oracle (mu-fold)."""
    code = []
    c.recount(np.zeros(0, dtype=np.int64))
    while c.uncovered():
        c.gain_map()
        w = int(np.argmax(c.map))
        code.append(w)
        c.ball_update([w], +1)
    return np.array(code, dtype=np.int64)


def _cell(gpu_cell, monkeypatch, problem, axes, R, mu, policy, ratio=None):
    monkeypatch.setenv("CC_SET_POLICY", str(policy))
    if ratio is not None:
        monkeypatch.setenv("CC_SET_RATIO", str(ratio))
    else:
        monkeypatch.delenv("CC_SET_RATIO", raising=False)
    return gpu_cell(axes, R, problem=problem, mu=mu)


def _run_gathers(g, c, keep, cands):
    """Recount, gathers for every role, then updates with gathers after
    each, all checked against the oracle c (which follows the same steps);
    returns the engine's counters."""
    c.recount(keep)
    g.recount(keep)
    for tgt in (0, 1):
        np.testing.assert_array_equal(g.ball_gather(cands, tgt),
                                      c.ball_gather(cands, tgt))
    # every cnt change invalidates the list: add two words, remove one
    for w, d in ((cands[:2], +1), (cands[:1], -1)):
        c.ball_update(w, d)
        g.ball_update(w, d)
        np.testing.assert_array_equal(g.ball_gather(cands, 0),
                                      c.ball_gather(cands, 0))
    return g.gpu_stats()


@pytest.mark.parametrize("mu", [1, 2])
@pytest.mark.parametrize("problem,axes,R", CELLS, ids=IDS)
def test_set_side_gather_near_cover(gpu_cell, monkeypatch, problem, axes, R, mu):
    """A cover with a few words removed: the set side (forced) and the walk
    give the oracle's counts for every role, and the list follows updates."""
    c = NumpyDP(axes, R, problem=problem, mu=mu)
    cover = _greedy_cover(c)
    rng = np.random.default_rng(len(axes) * 10 + R + mu)
    keep = cover[rng.permutation(len(cover))[: max(1, len(cover) - 3)]]
    cands = rng.integers(0, c.space.size, 500).astype(np.int64)
    # ratio 1000: a few removed words uncover up to 3 |B| words, more than
    # the default 2 |B| list limit on the small balls here
    g = _cell(gpu_cell, monkeypatch, problem, axes, R, mu, policy=2, ratio=1000)
    st = _run_gathers(g, c, keep, cands)
    assert st["gather_set_side"] >= 2 and st["set_scans"] >= 2
    # the cached list is reused while cnt is unchanged
    g.ball_gather(cands[:10], 0)
    assert g.gpu_stats()["set_scans"] == st["set_scans"]
    g.close()
    c = NumpyDP(axes, R, problem=problem, mu=mu)
    g = _cell(gpu_cell, monkeypatch, problem, axes, R, mu, policy=0)
    st = _run_gathers(g, c, keep, cands)
    assert st["gather_set_side"] == 0 and st["set_scans"] == 0


@pytest.mark.parametrize("problem,axes,R", CELLS[::2], ids=IDS[::2])
def test_set_side_gather_large_lists(gpu_cell, monkeypatch, problem, axes, R):
    """Random states (most words uncovered): with a large ratio the whole Z
    is listed and compared (thousands of entries); with ratio 1 the listing
    overflows and the walk answers."""
    from coldcore_amd.patterns import support_patterns
    c = NumpyDP(axes, R, problem=problem)
    size = c.space.size
    rng = np.random.default_rng(5)
    code = rng.choice(size, max(2, size // 200), replace=False).astype(np.int64)
    cands = rng.integers(0, size, 2000).astype(np.int64)
    c.recount(code)
    want = c.ball_gather(cands, 0)
    nz = int((c.cnt == 0).sum())
    g = _cell(gpu_cell, monkeypatch, problem, axes, R, 1, policy=2, ratio=100000)
    g.recount(code)
    np.testing.assert_array_equal(g.ball_gather(cands, 0), want)
    assert g.gpu_stats()["gather_set_side"] >= 1
    g.close()
    g = _cell(gpu_cell, monkeypatch, problem, axes, R, 1, policy=2, ratio=1)
    g.recount(code)
    np.testing.assert_array_equal(g.ball_gather(cands, 0), want)
    st = g.gpu_stats()
    blen = len(support_patterns(axes, R, problem))
    if nz > blen:                     # ratio 1: at most |B| listed words
        assert st["set_overflows"] >= 1 and st["gather_walk"] >= 1
    else:
        assert st["gather_set_side"] >= 1


@pytest.mark.parametrize("policy", [1, 2])
@pytest.mark.parametrize("mu", [1, 2])
@pytest.mark.parametrize("problem,axes,R", CELLS, ids=IDS)
def test_swap_eval_fast_path(gpu_cell, monkeypatch, problem, axes, R, mu, policy):
    """swap_eval by drop lists and corrections (candidate gains by the walk,
    policy 1, or the set side, policy 2) equals the oracle's drop / gather /
    undo sequence, with a repeated drop, and leaves cnt untouched."""
    c = NumpyDP(axes, R, problem=problem, mu=mu)
    cover = _greedy_cover(c)
    size = c.space.size
    rng = np.random.default_rng(3 + mu)
    g = _cell(gpu_cell, monkeypatch, problem, axes, R, mu, policy=policy)
    c.recount(cover)
    g.recount(cover)
    k = min(6, len(cover))
    drops = np.concatenate([rng.choice(cover, size=k, replace=False),
                            cover[:1]]).astype(np.int64)     # a repeated drop
    cands = [rng.integers(0, size, size=int(rng.integers(0, 80))).astype(np.int64)
             for _ in range(len(drops))]
    cnt0 = g.read_cnt().copy()
    loss, gains = g.swap_eval(drops, cands)
    assert (g.read_cnt() == cnt0).all()
    st = g.gpu_stats()
    assert st["swap_eval_fast"] == 1 and st["swap_eval_fallback"] == 0
    for i, w in enumerate(drops):
        wa = np.array([w], dtype=np.int64)
        assert loss[i] == c.ball_gather(wa, 1)[0]
        c.ball_update(wa, -1)
        if len(cands[i]):
            np.testing.assert_array_equal(gains[i], c.ball_gather(cands[i], 0))
        c.ball_update(wa, +1)


def test_swap_eval_overflow_falls_back(gpu_cell, monkeypatch):
    """Many repeated drops overflow the drop-list buffer: the exact
    mutate / gather / undo path answers instead."""
    from coldcore_amd.patterns import support_patterns
    problem, axes, R = "hamming", [3] * 6, 2
    c = NumpyDP(axes, R, problem=problem)
    cover = _greedy_cover(c)
    g = _cell(gpu_cell, monkeypatch, problem, axes, R, 1, policy=1)
    c.recount(cover)
    g.recount(cover)
    drops = np.repeat(cover[:4], 100).astype(np.int64)   # 400 drops
    rng = np.random.default_rng(9)
    cands = [rng.integers(0, c.space.size, 3).astype(np.int64) for _ in drops]
    # the drop lists' true total: words at cnt == mu in each drop's ball
    listed = sum(int((c.cnt[c.ball_indices(w)] == 1).sum()) for w in drops)
    lcap = max(4096, 64 * len(support_patterns(axes, R, problem)))
    loss, gains = g.swap_eval(drops, cands)
    st = g.gpu_stats()
    assert st["swap_eval_fallback"] == (1 if listed > lcap else 0)
    assert st["swap_eval_fast"] == (0 if listed > lcap else 1)
    for i in range(0, len(drops), 100):
        wa = drops[i:i + 1]
        assert loss[i] == c.ball_gather(wa, 1)[0]
        c.ball_update(wa, -1)
        np.testing.assert_array_equal(gains[i], c.ball_gather(cands[i], 0))
        c.ball_update(wa, +1)


def test_set_side_off_for_tables_that_are_not_the_ball(gpu_cell, monkeypatch):
    """A radius-R table with one pattern missing is walked as given; the set
    side (which counts the true ball) must stay off, so gathers follow the
    table."""
    import ctypes
    from coldcore_amd.gpu import load_plugin, _ptr, _U64P, _LLP, _I32P
    from coldcore_amd.patterns import support_patterns
    axes, R = [3] * 5, 1
    pats = np.ascontiguousarray(support_patterns(axes, R, "hamming")[:-1])
    lib = load_plugin()
    monkeypatch.setenv("CC_SET_POLICY", "2")
    ax = (ctypes.c_int * len(axes))(*axes)
    by = ctypes.c_longlong(0)
    assert lib.dct_init3(0, len(axes), ax, R, _ptr(pats, _U64P), len(pats),
                         1 << 10, ctypes.byref(by), 0, 0, 0) == 0
    try:
        code = np.array([0, 100, 200], dtype=np.int64)
        assert lib.dct_set_code(_ptr(code, _LLP), 3) == 0
        assert lib.dct_transform(0, 2) == 0
        cand = np.arange(243, dtype=np.int64)
        out = np.zeros(243, dtype=np.int32)
        assert lib.dct_ball_gather(_ptr(cand, _LLP), 243, 0, _ptr(out, _I32P)) == 0
        cnt = np.zeros(243, dtype=np.uint16)
        assert lib.dct_read_cnt(0, 243, cnt.ctypes.data_as(
            ctypes.POINTER(ctypes.c_uint16))) == 0
        # the table's own neighbourhoods, walked on the host
        from coldcore_amd.space import ProductSpace
        sp = ProductSpace(axes)
        digs = sp.index_word(cand).astype(np.int64)   # (243, 5)
        want = np.zeros(243, dtype=np.int32)
        for p in pats:
            d = digs.copy()
            for k in range(8):
                b = (int(p) >> (8 * k)) & 255
                if not b:
                    break
                pos, dl = b >> 4, b & 15
                d[:, pos] = (d[:, pos] + dl) % axes[pos]
            want += (cnt[sp.word_index(d)] == 0)
        np.testing.assert_array_equal(out, want)
        buf = (ctypes.c_longlong * 8)()
        lib.dct_stats(buf, 8)
        assert buf[0] == 0          # no set-side gather
    finally:
        lib.dct_free()
