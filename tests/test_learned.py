"""The learned ruin proposal: exact layer features, the network's kernels
against the CPU oracle, training, and what the trained network gains.

CPU part: the oracle network's backward equals finite differences; the
oracle's distance layers equal brute force; features and logging run on
the oracle backend.  GPU part: the engine's layers equal the oracle's, the
MFMA forward and gradient equal the oracle network, training lowers the
loss, and a network trained on plain-LNS logs ranks held-out regions better
than chance.
"""
import numpy as np
import pytest

from oracle import mlp as omlp
from oracle.dp import NumpyDP
from oracle.ref import layers_brute


def test_oracle_backward_equals_finite_differences():
    rng = np.random.default_rng(0)
    p = rng.normal(0, 0.3, omlp.P)
    X = rng.random((24, 16))
    t = rng.choice([0.0, 0.5, 1.0], 24)
    _, g = omlp.loss_and_grad(p, X, t, rounding=False)
    idx = np.concatenate([rng.choice(omlp.H * omlp.IN, 20, replace=False),
                          omlp.H * omlp.IN + rng.choice(2 * omlp.H, 20, replace=False),
                          [omlp.P - 1]])
    for i in idx:
        e = np.zeros(omlp.P)
        e[i] = 1e-6
        num = (omlp.loss_and_grad(p + e, X, t, False)[0]
               - omlp.loss_and_grad(p - e, X, t, False)[0]) / 2e-6
        assert abs(num - g[i]) <= 1e-6 + 1e-4 * abs(num), i


def test_bf16_rounding_is_nearest_even():
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 3.14159],
                 dtype=np.float32)
    r = omlp.bf16(v)
    assert r[0] == 1.0 and r[1] == 1.0            # tie -> even
    assert r[2] == 1.0 + 2 ** -6                  # tie -> even (up)
    assert r[3] == -2.5 and abs(r[4] - 3.140625) < 1e-12


def test_band_features():
    """Radius <= 2 windows pass unchanged; for R > 2 each window's layers
    become (layer 0, mean of layers 1..R-1, layer R): 15 inputs."""
    from coldcore_amd.learned import band_features
    rng = np.random.default_rng(0)
    F2 = rng.random((7, 15)).astype(np.float32)
    assert (band_features(F2, 2) == F2).all()
    F4 = rng.random((7, 25)).astype(np.float32)
    B = band_features(F4, 4)
    assert B.shape == (7, 15)
    for w in range(5):
        x = F4[:, 5 * w:5 * w + 5]
        assert np.allclose(B[:, 3 * w], x[:, 0])
        assert np.allclose(B[:, 3 * w + 1], x[:, 1:4].mean(axis=1))
        assert np.allclose(B[:, 3 * w + 2], x[:, 4])


@pytest.mark.parametrize("problem,axes,R", [
    ("hamming", [3] * 5, 2), ("hamming", [2, 3, 4], 3), ("lee", [5, 6, 7], 3)])
@pytest.mark.parametrize("init", [0, 1, 2])
def test_oracle_layers_equal_brute_force(problem, axes, R, init):
    rng = np.random.default_rng(3)
    c = NumpyDP(axes, R, problem)
    size = c.space.size
    code = rng.integers(0, size, max(2, size // 9))
    c.recount(code)
    idx = rng.integers(0, size, 50)
    got = c.layers_at(init, idx, code=code)
    seed = (code if init == 0 else
            np.flatnonzero(c.role_mask(0 if init == 1 else 1)))
    assert (got == layers_brute(problem, axes, R, seed, idx)).all()
    before = c.cnt.copy()
    c.layers_at(init, idx, code=code)
    assert (c.cnt == before).all()


def test_logging_and_features_on_oracle():
    from coldcore_amd.learned import LoggingOperator, RegionScorer
    from coldcore_amd.search import Searcher
    c = NumpyDP([3] * 6, 2)
    op = LoggingOperator(RegionScorer(c))
    s = Searcher(c, seed=2, on_round=op.outcome)
    s.load(np.zeros(0, dtype=np.int64))
    s.greedy_fill(10 ** 9)
    s.peel()
    s.ruin(1, "low")
    s.lns(len(s.code), kmin=2, kmax=4, max_rounds=30, operator=op)
    X = np.array(op.X)
    assert len(X) == s.stats["rounds"] > 0 and X.shape[1] == 15
    assert (X >= 0).all() and (X <= 1).all()
    assert set(op.t) <= {0.0, 0.5, 1.0}


# ------------------------------------------------------------------ GPU
@pytest.mark.gpu
@pytest.mark.parametrize("problem,axes,R,env", [
    ("hamming", [3] * 10, 2, {}), ("hamming", [4] * 8, 2, {}),
    ("hamming", [2] * 14, 3, {}),
    ("hamming", [4, 5, 4, 5, 3], 2, {}), ("lee", [5, 6, 7, 5], 3, {}),
    ("hamming", [3] * 10, 2, {"CC_LAYER16": "0"}),
    ("hamming", [2] * 14, 3, {"CC_SWEEP_LDS": "4096"})])
def test_engine_layers_equal_oracle(gpu_cell, monkeypatch, problem, axes, R, env):
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    rng = np.random.default_rng(5)
    c = NumpyDP(axes, R, problem)
    g = gpu_cell(axes, R, problem=problem)
    size = c.space.size
    code = np.unique(rng.integers(0, size, max(2, size // 30)))
    g.recount(code)
    c.recount(code)
    cnt0 = g.read_cnt().copy()
    idx = rng.integers(0, size, 500)
    for init in (0, 1, 2):
        assert (g.layers_at(init, idx, code=code)
                == c.layers_at(init, idx, code=code)).all(), init
    assert (g.read_cnt() == cnt0).all()
    g.gain_map()                           # a layers transform leaves no state
    c.gain_map()
    assert (g.map_read_at(idx) == c.map_read_at(idx)).all()


@pytest.mark.gpu
def test_network_forward_and_gradient_equal_oracle():
    from coldcore_amd.learned import Net, init_params
    net = Net(max_batch=512, seed=1)
    try:
        rng = np.random.default_rng(1)
        p = init_params(1, 15).astype(np.float32)
        p[omlp.H * omlp.IN:omlp.H * omlp.IN + omlp.H] = rng.normal(0, 0.1, omlp.H)
        p[-1] = 0.3
        net.set_params(p)
        for B in (1, 16, 37, 300):
            X = np.zeros((B, 16), dtype=np.float32)
            X[:, :15] = rng.random((B, 15))
            t = rng.choice([0.0, 0.5, 1.0], B).astype(np.float32)
            y = net.forward(X)
            _, yo = omlp.forward(p.astype(np.float64), X)
            assert np.allclose(y, yo, rtol=1e-4, atol=1e-4), B
            loss, g = net.grad(X, t)
            lo, go = omlp.loss_and_grad(p.astype(np.float64), X, t)
            assert abs(loss - lo) <= 1e-5 + 1e-4 * abs(lo)
            # fp32 accumulation of bf16 products against exact sums
            scale = np.abs(go).max()
            assert np.allclose(g, go, rtol=2e-3, atol=2e-5 * scale), B
        before = net.get_params()
        net.train_step(X, t, lr=1e-2)
        assert not np.array_equal(before, net.get_params())
    finally:
        net.close()


@pytest.mark.gpu
def test_training_fits_a_separable_rule():
    from coldcore_amd.learned import Net, train
    net = Net(max_batch=1024, seed=2)
    try:
        rng = np.random.default_rng(2)
        X = rng.random((4096, 15)).astype(np.float32)
        t = (X[:, 1] + X[:, 4] > X[:, 9] + 0.5).astype(np.float32)
        losses = train(net, X, t, epochs=15, batch=256, lr=3e-3)
        assert losses[-1] < 0.5 * losses[0]
        acc = ((net.forward(X) > 0) == (t > 0.5)).mean()
        assert acc > 0.9
    finally:
        net.close()


@pytest.mark.gpu
def test_trained_network_ranks_held_out_regions():
    """Trained on plain-LNS move logs of K_3(8,2), the network orders
    held-out regions by outcome better than chance (AUC of improving vs
    worse rounds) -- what the learned proposal adds over the blind mix."""
    from coldcore_amd.gpu import GpuBackend
    from coldcore_amd.learned import Net, collect_records, train
    X, t = collect_records(lambda: GpuBackend([3] * 8, 2), seeds=[1, 2, 3],
                           rounds=500)
    assert len(X) > 1000
    rng = np.random.default_rng(0)
    perm = rng.permutation(len(X))
    cut = int(0.75 * len(X))
    tr, te = perm[:cut], perm[cut:]
    net = Net(max_batch=2048, seed=3)
    try:
        train(net, X[tr], t[tr], epochs=30, batch=256, lr=2e-3)
        y = net.forward(X[te])
        pos, neg = y[t[te] == 1.0], y[t[te] == 0.0]
        assert len(pos) > 10 and len(neg) > 10
        auc = (pos[:, None] > neg[None, :]).mean()
        assert auc > 0.55, auc
    finally:
        net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("axes,R", [([3] * 8, 2), ([4] * 6, 2), ([2] * 12, 2),
                                    ([3, 4, 5, 3, 4], 2), ([2] * 10, 1),
                                    ([2] * 12, 4), ([3] * 7, 3)])
def test_device_region_features_equal_host_features(gpu_cell, axes, R):
    """dct_region_features (three layer transforms and the windows on the
    device, one download) equals the windows the host computes from the
    oracle's exact layers."""
    from coldcore_amd.learned import RegionScorer
    c = NumpyDP(axes, R)
    g = gpu_cell(axes, R)
    size = c.space.size
    rng = np.random.default_rng(size)
    code = np.unique(rng.integers(0, size, max(4, size // 40)))
    c.recount(code)
    g.recount(code)
    cnt0 = g.read_cnt().copy()
    for nreg, k in ((1, 1), (7, 3), (32, 5), (64, 8)):
        k = min(k, len(code))
        regions = np.stack([rng.choice(len(code), k, replace=False)
                            for _ in range(nreg)])
        want = RegionScorer(c).features(RegionScorer(c).layers(code), regions)
        got = RegionScorer(g).region_features(code, regions)
        assert got.shape == want.shape
        assert np.allclose(got, want, rtol=1e-6, atol=1e-7), (nreg, k)
    assert (g.read_cnt() == cnt0).all()          # cnt unchanged


@pytest.mark.gpu
@pytest.mark.parametrize("axes,R", [([3] * 10, 2), ([4] * 8, 2), ([2, 3, 4, 5, 3, 2], 2),
                                    ([2] * 14, 4)])
def test_device_region_proposal_equals_host(gpu_cell, axes, R):
    """dct_region_propose picks the same regions (nearest code words by digit
    distance, the same hashed tie-break) as the host restatement, and their
    features equal the host windows of the oracle's layers."""
    from coldcore_amd.learned import RegionScorer, cluster_regions
    c = NumpyDP(axes, R)
    g = gpu_cell(axes, R)
    size = c.space.size
    rng = np.random.default_rng(size + 1)
    code = np.unique(rng.integers(0, size, max(8, size // 60)))
    c.recount(code)
    g.recount(code)
    for na, k, seed in ((1, 1, 5), (32, 4, 2 ** 62 + 17), (64, 9, 123456789)):
        anchors = rng.choice(len(code), size=min(na, len(code)), replace=False)
        regions, feats = g.region_propose(code, anchors, k, seed)
        want = cluster_regions(c.space, code, k, anchors, seed)
        assert (regions == want).all(), (na, k)
        sc = RegionScorer(c)
        assert np.allclose(feats, sc.windows(sc.layers(code), want),
                           rtol=1e-6, atol=1e-7)


@pytest.mark.gpu
def test_portfolio_with_learned_proposal_radius4(gpu_cell):
    """The restart portfolio with the learned proposal as one more operator
    on a radius-4 binary cell (banded windows, 15 inputs): every stream's
    moves are judged by the exact fields, so the incumbent it leaves is a
    code of size M whose uncovered count a fresh recount confirms."""
    from coldcore_amd.learned import LearnedProposer, Net, RegionScorer
    from coldcore_amd.search import Searcher
    g = gpu_cell([2] * 14, 4)
    start = np.random.default_rng(4).choice(1 << 14, 8, replace=False).astype(np.int64)
    net = Net(max_batch=1024, seed=3)
    try:
        s = Searcher(g, seed=2)
        s.load(start)
        prop = LearnedProposer(net, RegionScorer(g), ncand=16)
        u = s.portfolio(len(start), streams=4, slice_rounds=2, max_restarts=3,
                        kmin=2, kmax=5, ruin_frac=0.2, select="adaptive",
                        proposer=prop)
        assert len(s.code) == len(start)
        assert s.verify_full() == u
        assert s.stats.get("learned_rounds", 0) > 0
    finally:
        net.close()


@pytest.mark.gpu
@pytest.mark.parametrize("axes,R", [([3] * 8, 2), ([2] * 12, 3), ([2] * 18, 4),
                                    ([4] * 6, 2)])
def test_region_features_by_walks_equal_transforms(gpu_cell, monkeypatch, axes, R):
    """Region windows from per-word ball walks and code comparisons (the
    path large cells take) equal the windows from the three whole-space
    layer transforms, exactly, for dct_region_features and
    dct_region_propose; both equal the host windows of the oracle."""
    from coldcore_amd.learned import RegionScorer
    size = int(np.prod(axes))
    rng = np.random.default_rng(size + R)
    code = np.unique(rng.integers(0, size, max(8, size // 300)))
    got = {}
    for mode in ("0", "2"):
        monkeypatch.setenv("CC_REGION_WALK", mode)
        g = gpu_cell(axes, R)
        g.recount(code)
        regions = np.stack([np.random.default_rng(7).choice(len(code), 5, replace=False)
                            for _ in range(24)])
        anchors = np.arange(min(16, len(code)))
        got[mode] = (RegionScorer(g).region_features(code, regions),
                     g.region_propose(code, anchors, 4, 99))
        g.close()
    assert (got["0"][0] == got["2"][0]).all()
    assert (got["0"][1][0] == got["2"][1][0]).all()
    assert (got["0"][1][1] == got["2"][1][1]).all()
    if size <= 1 << 16:
        c = NumpyDP(axes, R)
        c.recount(code)
        sc = RegionScorer(c)
        assert np.allclose(got["2"][0], sc.features(sc.layers(code), regions),
                           rtol=1e-6, atol=1e-7)
