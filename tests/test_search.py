"""The search counterpart (caller side of the hot path).

Pinned by the reference's frozen move-trace digests (reference
tests/test_search_determinism.py:24-50, 110-118): greedy build, peel and a
fixed ruin/recreate sequence over every ruin operator, on the brute-force
backend.  Then the same traces on the numpy axis-DP oracle (cpu == ref,
reference tests/test_dp_backend.py:124-160) and, on the GPU, through the
engine's C ABI: identical trajectories across all three.
"""
import hashlib

import numpy as np
import pytest

from coldcore_amd.search import OperatorBandit, Searcher
from oracle.dp import NumpyDP
from oracle.ref import BruteForce


def digest(trace):
    h = hashlib.sha256()
    for u, code in trace:
        h.update(f"{u}|{','.join(map(str, code))};".encode())
    return h.hexdigest()[:16]


def greedy_ruin_trace(backend, seed, rounds=16, notch=2):
    s = Searcher(backend, seed=seed)
    s.load(np.zeros(0, dtype=np.int64))
    s.greedy_fill(10 ** 9)
    trace = [(int(s.uncov), s.code.tolist())]
    s.peel()
    trace.append((int(s.uncov), s.code.tolist()))
    target = max(1, len(s.code) - notch)
    modes = ["low", "hole", "cluster", "random", "deficit"]
    for i in range(rounds):
        s.ruin(1 + i % 4, modes[i % len(modes)])
        s.greedy_fill(target)
        trace.append((int(s.uncov), s.code.tolist()))
    return trace


#: (id, (axes, R, problem, mu), seed, frozen digest of the reference)
PINS = [
    ("greedy-K3(4,1)", ([3] * 4, 1, "hamming", 1), 4, "7fe8ede9952ee50e"),
    ("greedy-K2(6,1)", ([2] * 6, 1, "hamming", 1), 11, "57c1e4d063f33278"),
    ("greedy-lee5x5", ([5, 5], 1, "lee", 1), 3, "3f8e8af3d3acc2ac"),
    ("greedy-mu2", ([2] * 6, 1, "hamming", 2), 5, "3f4b42018740787c"),
]


@pytest.mark.parametrize("label,cell,seed,pin", PINS, ids=[p[0] for p in PINS])
def test_frozen_digest_on_brute_force(label, cell, seed, pin):
    axes, R, prob, mu = cell
    assert digest(greedy_ruin_trace(BruteForce(axes, R, prob, mu), seed)) == pin


@pytest.mark.parametrize("label,cell,seed,pin", PINS, ids=[p[0] for p in PINS])
def test_frozen_digest_on_numpy_dp(label, cell, seed, pin):
    axes, R, prob, mu = cell
    assert digest(greedy_ruin_trace(NumpyDP(axes, R, prob, mu), seed)) == pin


def lns_trace(backend, seed, rounds, select="fixed"):
    log = []
    s = Searcher(backend, seed=seed,
                 on_round=lambda s, r: log.append((r["mode"], r["post"],
                                                   len(s.code))))
    s.load(np.zeros(0, dtype=np.int64))
    s.greedy_fill(10 ** 9)
    s.peel()
    M = len(s.code) - 1
    s.ruin(1, "low")                 # one notch below the feasible size
    s.lns(M, kmin=1, kmax=3, select=select, max_rounds=rounds)
    return log, s.code.tolist(), int(s.uncov)


@pytest.mark.parametrize("select", ["fixed", "adaptive"])
def test_lns_rounds_identical_cpu_and_ref(select):
    a = lns_trace(BruteForce([3] * 5, 1), 7, 60, select)
    b = lns_trace(NumpyDP([3] * 5, 1), 7, 60, select)
    assert a == b and len(a[0]) > 0


def test_bandit_probabilities():
    rng = np.random.default_rng(0)
    b = OperatorBandit(rng)
    p = b.probabilities()
    assert abs(p.sum() - 1) < 1e-12 and p[1] > p[0]       # hole prior x2
    for _ in range(30):
        b.update("random", b.REWARD_BEST)
    assert b.probabilities()[3] > p[3]
    assert min(b.probabilities()) >= b.floor / len(b.ops) - 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("label,cell,seed,pin", PINS, ids=[p[0] for p in PINS])
def test_frozen_digest_on_gpu_engine(gpu_cell, label, cell, seed, pin):
    """The engine through its C ABI reproduces the reference's frozen
    search trace: every field, gather, update and extraction order agree."""
    axes, R, prob, mu = cell
    g = gpu_cell(axes, R, problem=prob, mu=mu)
    assert digest(greedy_ruin_trace(g, seed)) == pin


@pytest.mark.gpu
@pytest.mark.parametrize("select", ["fixed", "adaptive"])
def test_lns_rounds_identical_gpu_and_cpu(gpu_cell, select):
    g = gpu_cell([3] * 6, 1)
    assert lns_trace(g, 7, 80, select) == lns_trace(NumpyDP([3] * 6, 1), 7,
                                                    80, select)


@pytest.mark.gpu
def test_config1_k3_8_2_trajectory_identical_gpu_and_cpu(gpu_cell):
    """K_3(8,2): 40-word seeded start, fields, greedy, peel and 1,000 LNS
    rounds of notch descent give the same per-round trace and final code
    on the engine and on the numpy oracle."""
    from coldcore_amd.search import config1_run
    start = np.random.default_rng(0).choice(3 ** 8, size=40, replace=False)
    g = gpu_cell([3] * 8, 2)
    a = config1_run(g, start, rounds=1000)
    b = config1_run(NumpyDP([3] * 8, 2), start, rounds=1000)
    assert a["rounds"] == b["rounds"] == 1000
    assert a["fields"] == b["fields"] and a["greedy"] == b["greedy"]
    assert a["sizes"] == b["sizes"] and a["trace"] == b["trace"]
    assert (a["final"] == b["final"]).all() and a["uncov"] == b["uncov"]


def test_config1_prefix_identical_cpu_and_ref():
    """The first rounds of the same run on brute force and the axis DP."""
    from coldcore_amd.search import config1_run
    start = np.random.default_rng(0).choice(3 ** 6, size=12, replace=False)
    a = config1_run(BruteForce([3] * 6, 2), start, rounds=40)
    b = config1_run(NumpyDP([3] * 6, 2), start, rounds=40)
    assert a["trace"] == b["trace"] and (a["final"] == b["final"]).all()


# ---- the reference search core's own config-1 trajectory -----------------
# tests/golden/lns_k3_8_2.json was written by tests/golden/make_lns_trace.py
# from the reference's Searcher on its numpy backend (K_3(8,2), 40-word start,
# greedy, peel, 1,000 LNS rounds of notch descent): per round the uncovered
# count after accept / reject and the code size, and the final code.
def _golden_lns():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "lns_k3_8_2.json")
    with open(path) as f:
        return json.load(f)


def _check_against_reference_run(backend):
    from coldcore_amd.search import config1_run
    gold = _golden_lns()
    start = np.random.default_rng(0).choice(3 ** 8, size=40, replace=False)
    r = config1_run(backend, start, rounds=gold["rounds"])
    assert r["greedy"] == gold["greedy"] and r["peeled"] == gold["peeled"]
    got = [[int(post), int(size)] for _mode, post, size in r["trace"]]
    want = [[u, size] for _M, u, size in gold["trace_M_uncov_size"]]
    assert len(got) == len(want) == gold["rounds"]
    first_bad = next((i for i, (a, b) in enumerate(zip(got, want)) if a != b), None)
    assert first_bad is None, (first_bad, got[first_bad], want[first_bad])
    assert [list(x) for x in r["sizes"]] == gold["sizes"]
    assert r["uncov"] == gold["final_uncovered"]
    final = np.asarray(r["final"], dtype=np.int64)
    assert hashlib.sha256(final.tobytes()).hexdigest() == gold["final_sha256"]


def test_config1_lns_equals_reference_trajectory_on_oracle():
    """The restated greedy / peel / LNS, driven by the numpy oracle, follows
    the reference's own 1,000-round trajectory round for round."""
    _check_against_reference_run(NumpyDP([3] * 8, 2))


@pytest.mark.gpu
def test_config1_lns_equals_reference_trajectory_on_gpu(gpu_cell):
    """The same 1,000 rounds through the engine's C ABI."""
    _check_against_reference_run(gpu_cell([3] * 8, 2))


# ---- restart portfolio (reference search.py:1536-1620) --------------------
def _portfolio_run(backend, seed=3):
    s = Searcher(backend, seed=seed)
    s.load(np.zeros(0, dtype=np.int64))
    s.greedy_fill(10 ** 9)
    s.peel()
    M = len(s.code) - 3                      # a hard notch: several rounds
    u = s.portfolio(M, streams=4, slice_rounds=4, max_restarts=3, kmin=2)
    return s, M, u


def test_portfolio_invariants_on_oracle():
    """The incumbent is loaded at the end, its uncovered count is a
    from-scratch recount's, and no stream exceeds M words."""
    c = NumpyDP([3] * 6, 1)
    s, M, u = _portfolio_run(c)
    assert len(s.code) <= M
    assert u == s.uncov == s.verify_full()
    assert s.stats["restarts"] <= 3
    assert (np.diff(s.code) > 0).all()


@pytest.mark.gpu
def test_portfolio_gpu_equals_oracle(gpu_cell):
    """Clock-free, the portfolio is deterministic: the engine's run equals the
    oracle's restart for restart."""
    a, _, ua = _portfolio_run(NumpyDP([3] * 6, 1))
    b, _, ub = _portfolio_run(gpu_cell([3] * 6, 1))
    assert ua == ub
    assert a.code.tolist() == b.code.tolist()
    assert a.stats == b.stats
