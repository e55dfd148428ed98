"""GPU parity: every field and query of the engine, bit-exact against the
CPU oracle, through the C ABI (ctypes).

Follows the reference's GPU test matrix (reference tests/test_gpu_fast_paths.py:
181-222 for the transform, tests/test_dp_backend.py:21-94 for the
per-family fields and queries) with the oracle as the reference side.
"""
import numpy as np
import pytest

from cells import FAMILY_CELLS, FAMILY_IDS, TRANSFORM_CELLS, TRANSFORM_IDS
from oracle.dp import NumpyDP

pytestmark = pytest.mark.gpu


# schedule variants: the default plan (packed-layer sweeps on homogeneous
# Hamming cells with q <= 5, the register-round kernel elsewhere), the
# unpacked register-round kernels, the per-axis kernel only, and a tiny LDS
# budget that forces short sweeps and multi-round strided tiles (the
# multi-sweep paths on small cells)
# (with small tiles the unpacked cells run several sweeps, which exchange
# compact bit-field words by default; "planes" keeps the layer planes)
VARIANTS = {"default": {}, "unpacked": {"CC_PACKED": "0"},
            "generic": {"CC_SWEEP_KERNEL": "generic"},
            "small-tiles": {"CC_SWEEP_LDS": "4096"},
            "planes": {"CC_SWEEP_LDS": "4096", "CC_COMPACT": "0"}}


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("axes,R,mu", TRANSFORM_CELLS, ids=TRANSFORM_IDS)
def test_hamming_transform_matches_oracle(gpu_cell, monkeypatch, axes, R, mu,
                                          variant):
    """cnt from a multiset code, the full gain map and the full loss map,
    including R >= n cells, odd and even n and mixed radix, under every
    schedule variant (reference tests/test_gpu_fast_paths.py:190-222)."""
    for k, v in VARIANTS[variant].items():
        monkeypatch.setenv(k, v)
    c = NumpyDP(axes, R, mu=mu)
    rng = np.random.default_rng(sum(axes) + R)
    size = c.space.size
    every = np.arange(size, dtype=np.int64)
    g = gpu_cell(axes, R, mu=mu)
    for m in (1, max(2, size // 50), max(3, size // 7)):
        code = rng.integers(0, size, size=m).astype(np.int64)  # multiset
        g.recount(code)
        c.recount(code)
        assert (g.read_cnt() == c.read_cnt()).all()
        g.gain_map()
        c.gain_map()
        assert (g.map_read_at(every) == c.map_read_at(every)).all()
        g.loss_map()
        c.loss_map()
        assert (g.map_read_at(every) == c.map_read_at(every)).all()


def _fields_match(g, c):
    assert (g.read_cnt() == c.read_cnt()).all()
    assert g.uncovered() == c.uncovered()
    for t in range(4):
        assert g.count_eq(t) == c.count_eq(t)
    every = np.arange(c.space.size, dtype=np.int64)
    for which in ("gain_map", "loss_map"):
        getattr(g, which)()
        getattr(c, which)()
        assert (g.map_read_at(every) == c.map_read_at(every)).all(), which
        assert g.map_max() == c.map_max()
        mx = c.map_max()
        if mx > 0:
            assert (g.map_hist(16, mx) == c.map_hist(16, mx)).all()
            for thr in (1, max(1, mx // 2), mx):
                gi, gv, gf = g.map_extract(thr)
                ci, cv, cf = c.map_extract(thr)
                assert gf == cf
                # ascending order: identical arrays, not just sets
                assert (gi == ci).all() and (gv == cv).all()


@pytest.mark.parametrize("problem,axes,R", FAMILY_CELLS, ids=FAMILY_IDS)
@pytest.mark.parametrize("mu", [1, 2])
def test_family_fields_and_queries(gpu_cell, problem, axes, R, mu):
    c = NumpyDP(axes, R, problem=problem, mu=mu)
    g = gpu_cell(axes, R, problem=problem, mu=mu)
    rng = np.random.default_rng(hash((problem, tuple(axes), R, mu)) % 2 ** 32)
    size = c.space.size
    for trial in range(3):
        m = int(rng.integers(1, max(2, size // 4)))
        code = np.unique(rng.integers(0, size, size=m).astype(np.int64))
        g.recount(code)
        c.recount(code)
        _fields_match(g, c)
        add = rng.integers(0, size, size=3).astype(np.int64)
        g.ball_update(add, +1)
        c.ball_update(add, +1)
        assert (g.read_cnt() == c.read_cnt()).all()
        probe = rng.integers(0, size, size=min(size, 32)).astype(np.int64)
        for target in (0, 1, 2):
            assert (g.ball_gather(probe, target)
                    == c.ball_gather(probe, target)).all()
        g.ball_update(add, -1)
        c.ball_update(add, -1)
        assert (g.read_cnt() == c.read_cnt()).all()
        off = int(rng.integers(0, size))
        assert (g.read_cnt(off, size - off) == c.read_cnt(off, size - off)).all()


def test_async_updates_then_reads(gpu_cell):
    """400 back-to-back asynchronous updates, then exact reads
    (reference tests/test_gpu_fast_paths.py:138-166)."""
    q, n, R = 3, 6, 1
    c = NumpyDP([q] * n, R)
    rng = np.random.default_rng(5)
    g = gpu_cell([q] * n, R)
    code = np.sort(rng.choice(c.space.size, 30, replace=False))
    g.recount(code)
    c.recount(code)
    live = list(code)
    for i in range(400):
        if i % 3 == 2 and live:
            w = np.array([live.pop(int(rng.integers(len(live))))])
            g.ball_update(w, -1)
            c.ball_update(w, -1)
        else:
            w = rng.integers(0, c.space.size, size=1 + i % 70)
            live.extend(int(x) for x in w)
            g.ball_update(w.astype(np.int64), +1)
            c.ball_update(w.astype(np.int64), +1)
        if i % 50 == 0:
            assert g.count_eq(0) == c.count_eq(0)
    assert (g.read_cnt() == c.read_cnt()).all()
    g.gain_map()
    c.gain_map()
    every = np.arange(c.space.size, dtype=np.int64)
    assert (g.map_read_at(every) == c.map_read_at(every)).all()


def test_large_batches_stage_and_match(gpu_cell):
    """Update batches above the by-value limit (staged, synchronous) and a
    gather batch above the 2^16 staging chunk."""
    axes, R = [4] * 8, 2
    c = NumpyDP(axes, R)
    g = gpu_cell(axes, R)
    rng = np.random.default_rng(17)
    size = c.space.size
    words = rng.integers(0, size, size=3000).astype(np.int64)
    g.recount(np.zeros(0, dtype=np.int64))
    c.recount(np.zeros(0, dtype=np.int64))
    g.ball_update(words, +1)
    c.ball_update(words, +1)
    assert (g.read_cnt() == c.read_cnt()).all()
    probe = rng.integers(0, size, size=70000).astype(np.int64)
    assert (g.ball_gather(probe, 0) == c.ball_gather(probe, 0)).all()
    g.ball_update(words[:2000], -1)
    c.ball_update(words[:2000], -1)
    assert (g.read_cnt() == c.read_cnt()).all()


@pytest.mark.parametrize("cap", [1, 7, 100, 4096])
def test_extract_overflow_counts_exactly(gpu_cell, cap):
    """found is exact past the cap; stored entries are the first cap hits in
    ascending order (reference tests/test_gpu_fast_paths.py:335-358)."""
    c = NumpyDP([5] * 6, 2)
    rng = np.random.default_rng(cap)
    code = np.sort(rng.choice(c.space.size, 30, replace=False))
    g = gpu_cell([5] * 6, 2)
    g.recount(code)
    c.recount(code)
    g.gain_map()
    c.gain_map()
    for thr in (1, 2, 5):
        gi, gv, gf = g.map_extract(thr, cap=cap)
        ci, cv, cf = c.map_extract(thr, cap=cap)
        assert gf == cf
        assert len(gi) == min(cap, cf)
        assert (gi == ci).all() and (gv == cv).all()


def test_counter_limit_and_rejection_preserves_state(gpu_cell):
    """The 65,535-occurrence boundary; every rejected operation leaves cnt
    byte-identical (reference tests/test_gpu_counter_limits.py:12-41)."""
    code = np.zeros(65535, dtype=np.int64)
    g = gpu_cell([3, 3], 1, mu=65535)
    g.recount(code)
    before = g.read_cnt().copy()
    assert before[0] == 65535
    for operation in (
        lambda: g.recount(np.zeros(65536, dtype=np.int64)),
        lambda: g.ball_update([0], +1),
        lambda: g.ball_update([1], -1),
        lambda: g.ball_update([0], 2),
        lambda: g.recount([-1]),
        lambda: g.recount([9]),
    ):
        with pytest.raises(RuntimeError):
            operation()
        np.testing.assert_array_equal(g.read_cnt(), before)
    g.ball_update([0], -1)
    assert g.read_cnt(0, 1)[0] == 65534
    g.ball_update([0], +1)
    np.testing.assert_array_equal(g.read_cnt(), before)
    g.recount([8])
    before = g.read_cnt().copy()
    with pytest.raises(RuntimeError):
        g.ball_update([8, 8], -1)
    np.testing.assert_array_equal(g.read_cnt(), before)
    g.ball_update([8], -1)
    assert not g.read_cnt().any()


@pytest.mark.parametrize("problem,axes,R", [
    ("hamming", [3] * 5, 1), ("lee", [5, 6, 7], 2),
    ("torus_linf", [7, 9, 11], 2), ("grid_linf", [4, 5, 6], 3)])
@pytest.mark.parametrize("mu", [1, 2])
def test_update_delta_and_swap_eval(gpu_cell, problem, axes, R, mu):
    """update_delta reports the exact change of uncovered(); swap_eval equals
    the drop / gather / undo sequence and leaves cnt untouched
    (reference tests/test_gpu_fast_paths.py:225-292)."""
    c = NumpyDP(axes, R, problem=problem, mu=mu)
    g = gpu_cell(axes, R, problem=problem, mu=mu)
    rng = np.random.default_rng(77)
    size = c.space.size
    code = np.sort(rng.choice(size, size=max(4, size // 40),
                              replace=False)).astype(np.int64)
    g.recount(code)
    c.recount(code)
    for w in (code[:1], code[1:4], rng.integers(0, size, 3)):
        w = np.asarray(w, dtype=np.int64)
        sign = -1 if np.isin(w, code).all() else +1
        u0 = c.uncovered()
        c.ball_update(w, sign)
        dd = g.ball_update_delta(w, sign)
        assert dd == c.uncovered() - u0
        assert g.uncovered() == c.uncovered()
        c.ball_update(w, -sign)
        assert g.ball_update_delta(w, -sign) == -dd
    k = min(5, len(code))
    drops = rng.choice(code, size=k, replace=False).astype(np.int64)
    cands = [rng.integers(0, size, size=int(rng.integers(0, 60))).astype(np.int64)
             for _ in range(k)]
    cnt0 = g.read_cnt().copy()
    loss, gains = g.swap_eval(drops, cands)
    assert (g.read_cnt() == cnt0).all()
    for i, w in enumerate(drops):
        wa = np.array([w], dtype=np.int64)
        assert loss[i] == c.ball_gather(wa, 1)[0]
        c.ball_update(wa, -1)
        if len(cands[i]):
            assert (gains[i] == c.ball_gather(cands[i], 0)).all()
        c.ball_update(wa, +1)


@pytest.mark.parametrize("problem,axes,R", [
    ("hamming", [3, 3], 1), ("lee", [5, 5], 1),
    ("torus_linf", [5, 5], 1), ("grid_linf", [5, 5], 1)])
@pytest.mark.parametrize("mu", [1, 2, 3])
def test_owner_loss_is_the_marginal_deficit(gpu_cell, problem, axes, R, mu):
    """dct_loss_owner equals removing each centre and recounting
    (reference tests/test_gpu_counter_limits.py:44-72)."""
    import ctypes
    code = np.array([0, 1, 2], dtype=np.int64)
    c = NumpyDP(axes, R, problem=problem, mu=mu)
    c.recount(code)
    base = c.uncovered()
    expected = []
    for i in range(len(code)):
        c.recount(np.delete(code, i))
        expected.append(c.uncovered() - base)
    g = gpu_cell(axes, R, problem=problem, mu=mu, use_owner=True)
    out = np.empty(len(code), dtype=np.int32)
    rc = g._lib.dct_loss_owner(
        code.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)), len(code),
        out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)))
    assert rc == 0
    np.testing.assert_array_equal(out, expected)
    g.ball_update(code, -1)
    assert not g.read_cnt().any()


# multi-sweep unpacked cells: q > 5 Hamming, mixed radix, Lee; the sweeps
# exchange compact words (every layer a bit field sized by its proven bound)
COMPACT_CELLS = [("hamming", [8] * 6, 4, 1), ("hamming", [10] * 5, 5, 1),
                 ("hamming", [4, 5, 6, 7, 8, 4], 3, 1), ("hamming", [7] * 6, 3, 2),
                 ("lee", [9] * 5, 4, 1), ("lee", [6, 7, 8, 9], 6, 2),
                 ("hamming", [6] * 7, 6, 1)]


@pytest.mark.parametrize("problem,axes,R,mu", COMPACT_CELLS,
                         ids=[f"{p}-{'x'.join(map(str, a))}-R{r}-mu{m}"
                              for p, a, r, m in COMPACT_CELLS])
@pytest.mark.parametrize("lds", ["16384", "65536"])
def test_compact_words_multi_sweep(gpu_cell, monkeypatch, problem, axes, R, mu, lds):
    """Recount of multiset codes (seed bound = the largest multiplicity,
    capped by the code size), gain and loss on cells that run several
    unpacked sweeps, with compact words and with layer planes between the
    sweeps: both bit-exact against the oracle, and the compact plan moves
    fewer bytes per word."""
    monkeypatch.setenv("CC_SWEEP_LDS", lds)
    c = NumpyDP(axes, R, problem=problem, mu=mu)
    size = c.space.size
    every = np.arange(size, dtype=np.int64)
    bpw = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("CC_COMPACT", mode)
        g = gpu_cell(axes, R, problem=problem, mu=mu)
        plan = g.sweep_plan2(0)
        assert len(plan) >= 2 and all(s["W"] == 32 for s in plan)
        bpw[mode] = sum(s["bytes_in"] + s["bytes_out"] for s in plan)
        rng = np.random.default_rng(size + R)
        for m, mult in ((max(2, size // 3000), 1), (max(3, size // 200), 5),
                        (20000, 3)):
            base = rng.integers(0, size, size=m).astype(np.int64)
            code = np.repeat(base, rng.integers(1, mult + 1, size=m))[:65535]
            g.recount(code)
            c.recount(code)
            assert (g.read_cnt() == c.read_cnt()).all(), (mode, m, mult)
            for which in ("gain_map", "loss_map"):
                getattr(g, which)()
                getattr(c, which)()
                assert (g.map_all() == c.map_read_at(every)).all(), (mode, which, m)
        g.close()
    assert bpw["1"] < bpw["0"], bpw
