"""Pins of the CPU oracle (no GPU).

The oracle is only trusted after these checks, each chosen so a plausible
mistake (dropped term, wrong sign or index, transposed operand) fails one:
* the axis DP equals brute force on the reference's 16-cell matrix, mu 1/2
  (reference tests/test_dp_backend.py:21-94);
* brute force equals closed-form ball sizes (empty code: gain = |B|) and
  the perfect codes of tests/golden/known_answers.json (cnt == 1
  everywhere), which fix every field value independently of any code here;
* the Walsh-Hadamard oracle equals brute force on binary cells;
* the reference's identities (tests/test_reference_backend.py:23-81).
"""
import json
import os

import numpy as np
import pytest

from cells import FAMILY_CELLS, FAMILY_IDS, ORACLE_ONLY_CELLS
from oracle.dp import NumpyDP
from oracle.ref import BruteForce
from oracle.walsh import binary_field

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "known_answers.json")
ALL = FAMILY_CELLS + ORACLE_ONLY_CELLS


@pytest.mark.parametrize("problem,axes,R", ALL,
                         ids=FAMILY_IDS + ["hamming-5x2x3-R0"])
@pytest.mark.parametrize("mu", [1, 2])
def test_dp_equals_brute_force(problem, axes, R, mu):
    rng = np.random.default_rng(hash((problem, tuple(axes), R, mu)) % 2 ** 32)
    ref = BruteForce(axes, R, problem, mu)
    dp = NumpyDP(axes, R, problem, mu)
    size = ref.space.size
    for trial in range(3):
        m = int(rng.integers(1, max(2, size // 4)))
        code = np.unique(rng.integers(0, size, size=m).astype(np.int64))
        ref.recount(code)
        dp.recount(code)
        assert (ref.cnt == dp.cnt).all()
        for which in ("gain_map", "loss_map"):
            getattr(ref, which)()
            getattr(dp, which)()
            assert (ref.map == dp.map).all(), which
        add = rng.integers(0, size, size=3).astype(np.int64)
        ref.ball_update(add, +1)
        dp.ball_update(add, +1)
        assert (ref.cnt == dp.cnt).all()
        probe = rng.integers(0, size, size=min(size, 32)).astype(np.int64)
        for t in (0, 1, 2):
            assert (ref.ball_gather(probe, t) == dp.ball_gather(probe, t)).all()
        for w in probe[:6]:
            assert (ref.ball_indices(w) == dp.ball_indices(w)).all()


def _hamming_ball(q, n, R):
    from math import comb
    return sum(comb(n, r) * (q - 1) ** r for r in range(min(R, n) + 1))


@pytest.mark.parametrize("problem,axes,R", ALL,
                         ids=FAMILY_IDS + ["hamming-5x2x3-R0"])
def test_empty_code_gain_is_ball_size(problem, axes, R):
    """With nothing covered every word's gain is its ball size; for the
    shift-invariant metrics that is one closed-form number everywhere."""
    ref = BruteForce(axes, R, problem)
    ref.recount(np.zeros(0, dtype=np.int64))
    ref.gain_map()
    if problem == "hamming":
        # mixed radix: coefficients 0..R of prod_j (1 + (a_j - 1) z)
        poly = [1] + [0] * R
        for a in axes:
            poly = [poly[d] + (a - 1) * (poly[d - 1] if d else 0)
                    for d in range(R + 1)]
        assert (ref.map == sum(poly)).all()
    elif problem == "torus_linf":
        assert (ref.map == int(np.prod([2 * R + 1 for _ in axes]))).all()
    elif problem == "lee":
        poly = [1] + [0] * R
        for a in axes:
            k = [0] * (R + 1)
            for w in range(0, min(a // 2, R) + 1):
                k[w] = 1 if (w == 0 or 2 * w == a) else 2
            poly = [sum(poly[i] * k[d - i] for i in range(d + 1))
                    for d in range(R + 1)]
        assert (ref.map == sum(poly)).all()
    else:  # open grid: product of clamped per-axis window lengths
        digs = ref.space.all_digits()
        want = np.ones(ref.space.size, dtype=np.int64)
        for j, a in enumerate(axes):
            lo = np.maximum(digs[:, j] - R, 0)
            hi = np.minimum(digs[:, j] + R, a - 1)
            want *= hi - lo + 1
        assert (ref.map == want).all()


def test_perfect_codes_cover_exactly_once():
    """Known answers: perfect codes tile the space, so cnt == 1 everywhere,
    gain == 0 everywhere, loss(c) == |B| at every codeword."""
    with open(GOLDEN) as f:
        cases = json.load(f)["perfect_codes"]
    assert len(cases) >= 4
    for case in cases:
        axes, R, prob = case["axes"], case["R"], case["problem"]
        code = np.array(case["code"], dtype=np.int64)
        for B in (BruteForce, NumpyDP):
            if B is BruteForce and int(np.prod(axes)) > (1 << 16):
                continue
            b = B(axes, R, prob)
            b.recount(code)
            assert (b.cnt == 1).all(), case["name"]
            b.gain_map()
            assert (b.map == 0).all()
            b.loss_map()
            assert (b.map[code] == case["ball"]).all()


def test_walsh_equals_brute_force():
    for n, R, m in [(8, 1, 9), (10, 3, 40), (12, 2, 100)]:
        rng = np.random.default_rng(n)
        code = rng.integers(0, 1 << n, m).astype(np.int64)  # multiset
        ref = BruteForce([2] * n, R)
        ref.recount(code)
        ind = np.zeros(1 << n, dtype=np.int64)
        np.add.at(ind, code, 1)
        assert (binary_field(ind, n, R) == ref.cnt).all()
        ref.gain_map()
        assert (binary_field((ref.cnt < 1).astype(np.int64), n, R) == ref.map).all()


def test_reference_identities():
    """gain == gather(., 0) everywhere, loss[code] == gather(code, 1),
    update == recount (reference tests/test_reference_backend.py:23-81)."""
    b = NumpyDP([3] * 4, 1)
    rng = np.random.default_rng(0)
    code = np.sort(rng.choice(81, 7, replace=False))
    b.recount(code)
    b.gain_map()
    every = np.arange(81, dtype=np.int64)
    assert (b.map_read_at(every) == b.ball_gather(every, 0)).all()
    b.loss_map()
    assert (b.map_read_at(code) == b.ball_gather(code, 1)).all()
    b.recount(code[:4])
    b.ball_update(code[4:], +1)
    inc = b.cnt.copy()
    b.recount(code)
    assert (inc == b.cnt).all()
    mx = b.map_max()
    hist = b.map_hist(16, max(mx, 1))
    assert hist.sum() == int((b.map > 0).sum())


def test_hist_and_extract_semantics():
    b = NumpyDP([4] * 4, 1)
    b.recount(np.array([0, 77, 200], dtype=np.int64))
    b.gain_map()
    mx = b.map_max()
    v = b.map[b.map > 0].astype(np.int64)
    want = np.bincount(np.minimum((v - 1) * 8 // mx, 7), minlength=8)
    assert (b.map_hist(8, mx) == want).all()
    idx, val, found = b.map_extract(mx // 2, cap=5)
    full = np.flatnonzero(b.map >= mx // 2)
    assert found == len(full) and (idx == full[:5]).all()
    assert (val == b.map[full[:5]]).all()
