"""L-inf cells with axes longer than 16 through the long-axis table format
(``dct_init3x``): every field and query bit-exact against the oracle.

The reference's v1 / ox tables hold a delta in a nibble, so its ABI stops
at axes of 16 (reference docs/PLUGIN.md:142-152); the batch workload's
k x k king tori run to k = 31.  The cells here cover both L-inf families,
R = 1..3, one to three axes above 16, mixed with short axes, and tiles that
cut a long axis into several sweeps (a 4 KiB LDS budget).
"""
import numpy as np
import pytest

from oracle.dp import NumpyDP
from test_gpu_parity import _fields_match

pytestmark = pytest.mark.gpu

LONG_CELLS = [
    ("torus_linf", [17, 17], 1),
    ("torus_linf", [31, 31], 1),
    ("torus_linf", [20, 24], 2),
    ("torus_linf", [17, 18, 19], 1),
    ("torus_linf", [100, 7], 3),
    ("grid_linf", [20, 24], 2),
    ("grid_linf", [64, 9], 1),
    ("grid_linf", [200, 3], 3),
    ("grid_linf", [33, 2, 17], 1),
]
LONG_IDS = [f"{p}-{'x'.join(map(str, a))}-R{r}" for p, a, r in LONG_CELLS]
VARIANTS = {"default": {}, "small-tiles": {"CC_SWEEP_LDS": "4096"}}


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("problem,axes,R", LONG_CELLS, ids=LONG_IDS)
def test_long_axis_fields_and_walks(gpu_cell, monkeypatch, problem, axes, R,
                                    variant):
    for k, v in VARIANTS[variant].items():
        monkeypatch.setenv(k, v)
    mu = 1 + (len(axes) % 2)
    c = NumpyDP(axes, R, problem=problem, mu=mu)
    g = gpu_cell(axes, R, problem=problem, mu=mu)
    assert g.longax
    rng = np.random.default_rng(sum(axes) * 7 + R)
    size = c.space.size
    for m in (1, max(2, size // 30), max(3, size // 5)):
        code = np.unique(rng.integers(0, size, size=m).astype(np.int64))
        g.recount(code)
        c.recount(code)
        _fields_match(g, c)
        add = rng.integers(0, size, size=5).astype(np.int64)
        g.ball_update(add, +1)
        c.ball_update(add, +1)
        assert (g.read_cnt() == c.read_cnt()).all()
        probe = rng.integers(0, size, size=min(size, 64)).astype(np.int64)
        for target in (0, 1, 2):
            assert (g.ball_gather(probe, target) == c.ball_gather(probe, target)).all()
        g.ball_update(add, -1)
        c.ball_update(add, -1)
        assert (g.read_cnt() == c.read_cnt()).all()


@pytest.mark.parametrize("policy", [0, 2])
@pytest.mark.parametrize("problem,axes,R", LONG_CELLS[::2], ids=LONG_IDS[::2])
def test_long_axis_gathers_swaps_and_deltas(gpu_cell, monkeypatch, problem,
                                            axes, R, policy):
    """Set-side gathers (policy 2 forces them where the table is the ball)
    and the walk (policy 0), update_delta and swap_eval against the
    oracle's drop / gather / undo sequence."""
    monkeypatch.setenv("CC_SET_POLICY", str(policy))
    monkeypatch.setenv("CC_SET_RATIO", "100000")
    c = NumpyDP(axes, R, problem=problem)
    g = gpu_cell(axes, R, problem=problem)
    rng = np.random.default_rng(len(axes) + R + policy)
    size = c.space.size
    code = np.sort(rng.choice(size, size=max(4, size // 25),
                              replace=False)).astype(np.int64)
    g.recount(code)
    c.recount(code)
    cands = rng.integers(0, size, 300).astype(np.int64)
    for target in (0, 1):
        assert (g.ball_gather(cands, target) == c.ball_gather(cands, target)).all()
    if policy == 2:
        assert g.gpu_stats()["gather_set_side"] >= 1
    for w in (code[:1], code[1:3], rng.integers(0, size, 2)):
        w = np.asarray(w, dtype=np.int64)
        sign = -1 if np.isin(w, code).all() else +1
        u0 = c.uncovered()
        c.ball_update(w, sign)
        assert g.ball_update_delta(w, sign) == c.uncovered() - u0
        c.ball_update(w, -sign)
        g.ball_update_delta(w, -sign)
    k = min(4, len(code))
    drops = rng.choice(code, size=k, replace=False).astype(np.int64)
    cl = [rng.integers(0, size, size=int(rng.integers(1, 50))).astype(np.int64)
          for _ in range(k)]
    cnt0 = g.read_cnt().copy()
    loss, gains = g.swap_eval(drops, cl)
    assert (g.read_cnt() == cnt0).all()
    for i, w in enumerate(drops):
        wa = np.array([w], dtype=np.int64)
        assert loss[i] == c.ball_gather(wa, 1)[0]
        c.ball_update(wa, -1)
        assert (gains[i] == c.ball_gather(cl[i], 0)).all()
        c.ball_update(wa, +1)


def test_long_axis_limits_are_enforced(gpu_cell):
    """R > 3 with an axis above 16, an axis above 1024, and a long-axis
    Hamming cell are rejected at init; a rejected init leaves no cell."""
    for problem, axes, R in (("torus_linf", [20, 20], 4),
                             ("grid_linf", [1100, 2], 1),
                             ("hamming", [17, 2], 1)):
        with pytest.raises((RuntimeError, ValueError)):
            gpu_cell(axes, R, problem=problem)


def test_malformed_long_tables_are_rejected():
    """dct_init3x proves the table walk-safe first: a delta equal to its
    axis, a position past n, a repeated position or a field after the
    terminating zero returns -1 and leaves no cell."""
    import ctypes
    from coldcore_amd.gpu import load_plugin, _ptr, _U64P
    from coldcore_amd.patterns import support_patterns_long
    lib = load_plugin()
    axes, R = [20, 24], 1
    good = support_patterns_long(axes, R, "torus_linf")
    ax = (ctypes.c_int * 2)(*axes)

    def init(tbl):
        tbl = np.ascontiguousarray(tbl, dtype=np.uint64)
        by = ctypes.c_longlong(0)
        rc = lib.dct_init3x(1, 2, ax, R, _ptr(tbl, _U64P), len(tbl), 1 << 10,
                            ctypes.byref(by), 0, 0, 0)
        if rc == 0:
            lib.dct_free()
        return rc

    assert init(good) == 0
    for bad in ((0 << 10) | 20,                 # delta == axis size
                (2 << 10) | 1,                  # position past n
                ((0 << 10) | 1) | (((0 << 10) | 2) << 16),   # repeated position
                ((1 << 10) | 1) << 16):         # a field after the zero field
        tbl = np.concatenate([good, np.array([bad], dtype=np.uint64)])
        assert tbl.dtype == np.uint64
        assert init(tbl) == -1, hex(bad)

