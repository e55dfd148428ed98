"""Host-side logic (no GPU): support tables, the binding's table order, and
the C-ABI library's exported surface."""
import ctypes
import os
import re

import numpy as np
import pytest

from coldcore_amd.patterns import (ball_patterns, ball_size, decode_fields,
                                   decode_v1, lee_ball_size, support_patterns,
                                   support_patterns_long, support_patterns_ox)
from coldcore_amd.space import ProductSpace
from oracle.dp import NumpyDP

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "coldcore_amd", "plugin.h")
LIB = os.path.join(ROOT, "coldcore_amd", "libcoldcore_covering.so")


def test_index_roundtrip():
    sp = ProductSpace(5, 6)
    idx = np.array([0, 1, 5, 5 ** 6 - 1, 12345], dtype=np.int64)
    assert (sp.word_index(sp.index_word(idx)) == idx).all()
    mixed = ProductSpace([2, 3, 4])
    assert mixed.size == 24 and list(mixed.pow_q) == [1, 2, 6]


@pytest.mark.parametrize("q,n,R", [(3, 4, 2), (2, 8, 3), (4, 5, 2), (5, 3, 1)])
def test_hamming_table_is_the_ball(q, n, R):
    pats = ball_patterns(q, n, R)
    assert len(pats) == ball_size(q, n, R) and pats[0] == 0
    sp = ProductSpace(q, n)
    centre = np.arange(n) % q
    seen = set()
    for p in pats:
        d = decode_v1(p)
        assert len(d) <= R
        w = centre.copy()
        for pos, delta in d.items():
            w[pos] = (w[pos] + delta) % q
        seen.add(int(sp.word_index(w[None, :])[0]))
    assert len(seen) == ball_size(q, n, R)


def test_radius_above_length_is_whole_space():
    for q, n, R in [(3, 2, 3), (2, 3, 5)]:
        assert len(ball_patterns(q, n, R)) == q ** n == ball_size(q, n, R)


def test_ox_matches_v1():
    pats = ball_patterns(4, 5, 2)
    masks, offs = support_patterns_ox([4] * 5, 2)
    assert len(masks) == len(pats)
    for k in range(len(pats)):
        d = decode_v1(pats[k])
        assert sum(1 << p for p in d) == int(masks[k])
        for p, dl in d.items():
            assert (int(offs[k]) >> (4 * p)) & 0xF == dl


@pytest.mark.parametrize("axes,R", [([5, 5], 1), ([7, 7], 2), ([4, 5, 6], 3),
                                    ([8, 8], 4), ([9, 9, 9], 2)])
def test_lee_table_size(axes, R):
    assert len(support_patterns(axes, R, "lee")) == lee_ball_size(axes, R)


@pytest.mark.parametrize("problem,axes,R", [
    ("hamming", [2, 3, 4, 5], 2), ("lee", [4, 5, 6], 3),
    ("torus_linf", [7, 7, 7], 1), ("grid_linf", [5, 6, 4], 2),
    ("grid_linf", [7, 7], 5), ("hamming", [3] * 5, 2)])
def test_table_walks_the_oracle_ball(problem, axes, R):
    """Decoding the packed bytes and walking them from a centre gives the
    oracle's ball (the same check the reference runs,
    tests/test_dp_backend.py:163-194)."""
    pats = support_patterns(axes, R, problem)
    dp = NumpyDP(axes, R, problem)
    sp = ProductSpace(axes)
    for w in (0, sp.size // 3, sp.size - 1):
        dig = sp.index_word([w])[0].astype(np.int64)
        got = set()
        for v in pats.tolist():
            x, ok = w, True
            for p, d in decode_v1(v).items():
                nd = (int(dig[p]) + d) % axes[p]
                sh = nd - int(dig[p])
                if problem == "grid_linf" and abs(sh) > R:
                    ok = False
                    break
                x += sh * int(sp.pow_q[p])
            if ok:
                got.add(x)
        assert got == set(dp.ball_indices(w).tolist())


@pytest.mark.parametrize("problem,axes,R", [
    ("torus_linf", [20, 24], 2), ("torus_linf", [17, 18, 19], 1),
    ("grid_linf", [64, 9], 1), ("grid_linf", [200, 3], 3),
    ("grid_linf", [33, 2, 17], 1), ("torus_linf", [100, 7], 3)])
def test_long_axis_table_walks_the_oracle_ball(problem, axes, R):
    """The long-axis table (16-bit pos<<10|delta fields) decodes and walks
    to the oracle's ball, with the open grid's off-board moves dropped as
    the kernels drop them; its order is a permutation of the same set."""
    from coldcore_amd.gpu import locality_order
    pats = support_patterns_long(axes, R, problem)
    assert pats[0] == 0 and len(set(pats.tolist())) == len(pats)
    dp = NumpyDP(axes, R, problem)
    sp = ProductSpace(axes)
    for w in (0, 1, sp.size // 3, sp.size - 1):
        dig = sp.index_word([w])[0].astype(np.int64)
        got = []
        for v in pats.tolist():
            x, ok = w, True
            for p, d in decode_fields(v, 16).items():
                a = axes[p]
                sh = d if int(dig[p]) + d < a else d - a
                if problem == "grid_linf" and abs(sh) > R:
                    ok = False
                    break
                x += sh * int(sp.pow_q[p])
            if ok:
                got.append(x)
        assert sorted(got) == sorted(dp.ball_indices(w).tolist())
    out = locality_order(pats, axes, bits=16)
    assert out[0] == 0 and (np.sort(out) == np.sort(pats)).all()


def test_long_axis_table_limits():
    with pytest.raises(ValueError):
        support_patterns_long([1100, 2], 1, "grid_linf")      # delta field
    with pytest.raises(ValueError):
        support_patterns_long([17] * 5, 1, "torus_linf")      # 5 moved axes
    # on short axes it is the v1 table in wider fields
    a = support_patterns([9, 9], 2, "torus_linf")
    b = support_patterns_long([9, 9], 2, "torus_linf")
    assert [decode_v1(x) for x in a] == [decode_fields(y, 16) for y in b]


def test_locality_order_is_permutation():
    from coldcore_amd.gpu import locality_order
    for problem, axes, R in [("hamming", [5] * 6, 3), ("lee", [7, 7, 7], 3),
                             ("grid_linf", [4, 5, 6], 3)]:
        pats = support_patterns(axes, R, problem)
        out = locality_order(pats, axes)
        assert out[0] == 0 and (np.sort(out) == np.sort(pats)).all()


BATCH_HEADER = os.path.join(ROOT, "include", "coldcore_amd", "batch.h")


def _declared_symbols(header=HEADER, prefix="dct_"):
    with open(header) as f:
        text = f.read()
    return sorted(set(re.findall(r"\b(" + prefix + r"[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_v1_surface():
    from coldcore_amd.gpu import ADDITIVE, OPTIONAL, REQUIRED
    declared = _declared_symbols()
    assert set(REQUIRED) | set(OPTIONAL) | set(ADDITIVE) <= set(declared)


def test_library_exports_every_declared_symbol():
    """Load the built library (no device call) and resolve every symbol the
    header declares."""
    if not os.path.exists(LIB):
        pytest.fail("libcoldcore_covering.so is not built "
                    "(python -m coldcore_amd.build)")
    lib = ctypes.CDLL(LIB)
    for name in _declared_symbols() + _declared_symbols(BATCH_HEADER, "ccb_"):
        assert hasattr(lib, name), name


def test_batch_header_matches_binding():
    from coldcore_amd.batch import BATCH_SYMBOLS
    assert set(_declared_symbols(BATCH_HEADER, "ccb_")) == set(BATCH_SYMBOLS)


# ---------------------------------------------------------------- planner
def test_plan_probe_k2_24_3_schedule(monkeypatch):
    """K_2(24,3): axes 0-11 in one persistent LDS-tiled sweep with 8-bit
    fields, then two six-axis column sweeps with 16-bit fields; 4-byte
    compact words between sweeps: 2+4 + 4+4 + 4+4 = 22 bytes per word.  With
    narrow words on, the layers' 1+4+7+8 = 20 bits (after 12 axes) and
    1+5+8+10 = 24 bits (after 18) travel as 3-byte words: 18 bytes."""
    from coldcore_amd.gpu import plan_probe
    plan = plan_probe([2] * 24, 3)
    assert [(s["first_axis"], s["axes"]) for s in plan] == [(0, 12), (12, 6), (18, 6)]
    assert [s["W"] for s in plan] == [8, 16, 16]
    assert sum(s["bytes_in"] + s["bytes_out"] for s in plan) == 22
    monkeypatch.setenv("CC_PK_NARROW", "1")
    plan = plan_probe([2] * 24, 3)
    assert [(s["bytes_in"], s["bytes_out"]) for s in plan] == [(2, 3), (3, 3), (3, 4)]


@pytest.mark.parametrize("q,ns,Rs", [(2, range(4, 41, 3), range(1, 7)),
                                     (3, range(3, 26, 2), range(1, 5)),
                                     (4, range(3, 21, 2), range(1, 5)),
                                     (5, range(2, 18, 2), range(1, 4))])
def test_plan_probe_schedules_are_consistent(q, ns, Rs):
    """Every packed schedule the planner makes covers the axes once, in
    order; each sweep reads what the previous one wrote; seeds are read
    only by the first sweep and the map / cnt written only by the last; and
    (checked inside the planner) every sweep's kernel instantiation exists."""
    from coldcore_amd.gpu import plan_probe
    for n in ns:
        if q ** n > 2 ** 40:
            continue
        for R in Rs:
            for kind, s0, M in ((0, 1, 1), (1, 1, 500), (1, 3, 65535)):
                plan = plan_probe([q] * n, R, kind, s0, M)
                if not plan:
                    continue
                j = 0
                for i, sw in enumerate(plan):
                    assert sw["first_axis"] == j and sw["axes"] >= 1
                    j += sw["axes"]
                    assert sw["W"] in (8, 16) and 1 <= sw["ND"] <= 4
                    assert sw["kernel"] >= 1
                    if i == 0 and kind == 1 and sw["kernel"] == 5:
                        # the persistent first sweep seeds a code from
                        # its sorted word list: no plane is read
                        assert sw["bytes_in"] == 0
                    elif i == 0:
                        assert sw["bytes_in"] == (2 if kind == 0 else 4)
                    else:
                        assert sw["bytes_in"] == plan[i - 1]["bytes_out"]
                    if i == len(plan) - 1:
                        assert sw["bytes_out"] == (4 if kind == 0 else 2)
                    else:
                        # dword planes (4 / 8 bytes) or, with R >= 3, a
                        # narrow 16 / 24-bit word
                        assert sw["bytes_out"] in ((2, 3, 4, 8) if R >= 3 else (4, 8))
                assert j == n, (q, n, R, kind)
