"""Learned ruin proposal: an MLP that scores which region LNS should ruin.

A ruin region is an anchor incumbent and its k-1 nearest incumbents (digit
distance, random tie-break -- the "cluster" shape of the reference's ruin
operators, reference src/coldcore/search.py:576-584).  Its features are five
windows of exact distance-layer counts, each R+1 wide (5 (R+1) <= 16):

  1. mean over the region of the critical layers L_d(w)
  2. mean over the region of the deficient layers G_d(w)
  3. max  over the region of G_d(w)
  4. mean over the region of the codeword layers C_d(w)
  5. min  over the region of L_d(w)

where X_d(w) is the exact number of words of that kind at distance exactly d
from w (dct_transform_layers: the transform's per-distance layers), divided
by the distance-d shell size so every feature lies in [0, 1].

The network (coldcore_amd/csrc/mlp.hip) is trained on the GPU by its own
forward and backward kernels on records the plain LNS logs (region features,
outcome: 1 improved, 0.5 sideways, 0 worse).  At search time it is one more
operator under the adaptive operator bandit, which stays the fallback: a
region the network proposes is still repaired by the exact lazy greedy and
judged by the exact fields, so no count is ever approximated.
"""
import ctypes
import time
from math import comb

import numpy as np

from .gpu import load_plugin
from .search import Searcher

IN, HID = 16, 256
NPARAM = HID * IN + HID + HID + 1
_FP = ctypes.POINTER(ctypes.c_float)


class Net:
    """The GPU network (one per process)."""

    def __init__(self, max_batch=4096, seed=0, in_dim=IN):
        self.lib = lib = load_plugin()
        for f in ("cc_mlp_init", "cc_mlp_free", "cc_mlp_set_params",
                  "cc_mlp_get_params", "cc_mlp_forward", "cc_mlp_grad",
                  "cc_mlp_adam", "cc_mlp_param_count"):
            getattr(lib, f).restype = ctypes.c_int
        lib.cc_mlp_adam.argtypes = [ctypes.c_float] * 4
        assert lib.cc_mlp_param_count() == NPARAM
        self.max_batch = max_batch
        rc = lib.cc_mlp_init(max_batch)
        if rc:
            raise RuntimeError(f"cc_mlp_init rc={rc}")
        self.set_params(init_params(seed, in_dim))

    def close(self):
        self.lib.cc_mlp_free()

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == NPARAM
        rc = self.lib.cc_mlp_set_params(p.ctypes.data_as(_FP))
        if rc:
            raise RuntimeError(f"cc_mlp_set_params rc={rc}")

    def get_params(self):
        p = np.empty(NPARAM, dtype=np.float32)
        rc = self.lib.cc_mlp_get_params(p.ctypes.data_as(_FP))
        if rc:
            raise RuntimeError(f"cc_mlp_get_params rc={rc}")
        return p

    @staticmethod
    def _x(X):
        X = np.asarray(X, dtype=np.float32)
        out = np.zeros((len(X), IN), dtype=np.float32)
        out[:, :X.shape[1]] = X
        return out

    def forward(self, X):
        X = self._x(X)
        y = np.empty(len(X), dtype=np.float32)
        for lo in range(0, len(X), self.max_batch):
            xb = np.ascontiguousarray(X[lo:lo + self.max_batch])
            yb = np.empty(len(xb), dtype=np.float32)
            rc = self.lib.cc_mlp_forward(xb.ctypes.data_as(_FP), len(xb),
                                         yb.ctypes.data_as(_FP))
            if rc:
                raise RuntimeError(f"cc_mlp_forward rc={rc}")
            y[lo:lo + len(xb)] = yb
        return y

    def grad(self, X, t):
        X = self._x(X)
        t = np.ascontiguousarray(t, dtype=np.float32)
        g = np.empty(NPARAM, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        rc = self.lib.cc_mlp_grad(X.ctypes.data_as(_FP), t.ctypes.data_as(_FP),
                                  len(X), g.ctypes.data_as(_FP),
                                  loss.ctypes.data_as(_FP))
        if rc:
            raise RuntimeError(f"cc_mlp_grad rc={rc}")
        return float(loss[0]), g

    def train_step(self, X, t, lr=1e-3):
        X = self._x(X)
        t = np.ascontiguousarray(t, dtype=np.float32)
        loss = np.zeros(1, dtype=np.float32)
        rc = self.lib.cc_mlp_grad(X.ctypes.data_as(_FP), t.ctypes.data_as(_FP),
                                  len(X), None, loss.ctypes.data_as(_FP))
        if rc == 0:
            rc = self.lib.cc_mlp_adam(lr, 0.9, 0.999, 1e-8)
        if rc:
            raise RuntimeError(f"mlp train step rc={rc}")
        return float(loss[0])


def init_params(seed, in_dim=IN):
    rng = np.random.default_rng(seed)
    p = np.zeros(NPARAM, dtype=np.float32)
    W1 = np.zeros((HID, IN), dtype=np.float32)
    W1[:, :in_dim] = rng.normal(0, np.sqrt(2.0 / in_dim), (HID, in_dim))
    p[:HID * IN] = W1.ravel()
    p[HID * IN + HID:HID * IN + 2 * HID] = rng.normal(0, np.sqrt(1.0 / HID), HID)
    return p


def band_features(F, R):
    """Five (R+1)-layer windows as the network's inputs: for R > 2 each
    window's layers become three bands -- layer 0, the mean of layers
    1..R-1, layer R -- so every radius gives 15 inputs; R <= 2 is returned
    unchanged (5 (R+1) <= 15).  Applied alike to the device and the host
    windows, so both paths feed the network the same numbers."""
    F = np.asarray(F, dtype=np.float32)
    if R <= 2:
        return F
    W = F.reshape(len(F), 5, R + 1)
    out = np.stack([W[:, :, 0], W[:, :, 1:R].mean(axis=2), W[:, :, R]], axis=2)
    return out.reshape(len(F), 15).astype(np.float32)


def shell_sizes(axes, R, problem="hamming"):
    """Words at exact distance d from any word, d = 0..R (Hamming: the
    elementary symmetric polynomials of a_j - 1)."""
    if problem != "hamming":
        raise ValueError("region features are defined for Hamming cells")
    poly = [1] + [0] * R
    for a in axes:
        poly = [poly[d] + (a - 1) * (poly[d - 1] if d else 0) for d in range(R + 1)]
    return np.maximum(np.array(poly, dtype=np.float64), 1.0)


class RegionScorer:
    """Exact-layer features of candidate ruin regions of one backend."""

    def __init__(self, backend):
        b = backend
        self.b = b
        self.R = int(b.R)
        self.shell = shell_sizes(b.space.axes, self.R)

    def layers(self, code):
        """(m, R+1) deficient, critical and codeword layers at the code."""
        b = self.b
        G = b.layers_at(1, code) / self.shell
        L = b.layers_at(2, code) / self.shell
        C = b.layers_at(0, code, code=code) / self.shell
        return G, L, C

    def region_features(self, code, regions):
        """(regions, inputs) features of equal-size regions (positions into
        ``code``): on the engine one device call (dct_region_features: the
        three layer transforms and the windows, one download); on a host
        backend the layers at the code and the windows in numpy; then
        :func:`band_features`."""
        W = np.asarray(regions, dtype=np.int64)
        if hasattr(self.b, "region_features"):
            F = self.b.region_features(code, np.asarray(code, dtype=np.int64)[W])
            return band_features(F, self.R)
        return self.features(self.layers(code), W)

    def windows(self, lay, regions):
        """(regions, 5 (R+1)) windows of equal-size regions (positions into
        the code), all regions at once."""
        G, L, C = lay
        W = np.asarray(regions, dtype=np.int64)          # (regions, k)
        Lw, Gw, Cw = L[W], G[W], C[W]                    # (regions, k, R+1)
        return np.concatenate([Lw.mean(1), Gw.mean(1), Gw.max(1), Cw.mean(1),
                               Lw.min(1)], axis=1).astype(np.float32)

    def features(self, lay, regions):
        """The network's inputs of equal-size regions: the windows, banded."""
        return band_features(self.windows(lay, regions), self.R)


_M64 = (1 << 64) - 1


def _mix64(x):
    """splitmix64's finaliser on uint64 arrays (the tie-break hash the device
    region picker uses)."""
    x = (x + np.uint64(0x9E3779B97F4A7C15)) & np.uint64(_M64)
    x = ((x ^ (x >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & np.uint64(_M64)
    x = ((x ^ (x >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & np.uint64(_M64)
    return x ^ (x >> np.uint64(31))


def cluster_regions(space, code, k, anchors, seed):
    """For every anchor position, the k code words nearest to it in digit
    distance, ties by a 24-bit hash of (seed, anchor, position) and then by
    position -- the device picker's order (dct_region_propose), restated on
    the host for the CPU backends.  A (anchors, k) array of positions."""
    code = np.asarray(code, dtype=np.int64)
    anchors = np.asarray(anchors, dtype=np.int64)
    radix = np.asarray(space.axes, dtype=np.int64)
    place = np.concatenate(([1], np.cumprod(radix[:-1]))).astype(np.int64)
    digs = ((code[:, None] // place[None, :]) % radix[None, :]).astype(np.int16)
    d = (digs[None, :, :] != digs[anchors][:, None, :]).sum(axis=2).astype(np.uint64)
    j = np.arange(len(code), dtype=np.uint64)
    with np.errstate(over="ignore"):
        h = _mix64(np.uint64(seed & _M64) ^ (anchors.astype(np.uint64)[:, None] << np.uint64(32))
                   ^ j[None, :]) & np.uint64(0xFFFFFF)
    key = (d << np.uint64(40)) | (h << np.uint64(16)) | j[None, :]
    return np.argsort(key, axis=1, kind="stable")[:, :k].astype(np.int64)


class LoggingOperator:
    """Wraps the plain operator mix to log (features, outcome) records: the
    region the plain operator ruins is featurised before the ruin, the
    round's outcome is attached after it."""

    def __init__(self, scorer):
        self.scorer = scorer
        self.X, self.t = [], []
        self._pending = None

    def __call__(self, searcher, k, mode):
        pick = searcher.ruin_pick(k, mode)
        self._pending = self.scorer.region_features(searcher.code, [pick])[0]
        return pick

    def outcome(self, searcher, rec):
        if self._pending is None:
            return
        pre, post = rec["pre"], rec["post"]
        self.X.append(self._pending)
        self.t.append(1.0 if post < pre else (0.5 if post == pre else 0.0))
        self._pending = None


class LearnedProposer:
    """Scores ``ncand`` cluster regions with the network and ruins the best
    (ties and exploration by a softmax at temperature ``tau``)."""

    def __init__(self, net, scorer, ncand=32, tau=0.25, norm=None, prior=2.0):
        self.net, self.scorer = net, scorer
        self.ncand, self.tau = ncand, tau
        self.norm = norm          # (mean, std) the network was trained with
        self.prior = prior        # the operator bandit's prior weight for it

    def __call__(self, searcher, k, mode=None):
        s = searcher
        m = len(s.code)
        k = min(k, m, 64)
        anchors = s.rng.choice(m, size=min(self.ncand, m), replace=False)
        seed = int(s.rng.integers(0, 2 ** 63))
        if hasattr(s.b, "region_propose"):
            # regions picked and featurised on the device, one call
            regions, F = s.b.region_propose(s.code, anchors, k, seed)
            F = band_features(F, self.scorer.R)
        else:
            regions = cluster_regions(s.b.space, s.code, k, anchors, seed)
            F = self.scorer.region_features(s.code, regions)
        if self.norm is not None:
            F = (F - self.norm[0]) / self.norm[1]
        y = self.net.forward(F)
        z = (y - y.max()) / self.tau
        p = np.exp(z) / np.exp(z).sum()
        return regions[int(s.rng.choice(len(regions), p=p))]


def collect_records(make_backend, seeds, rounds, kmin=2, kmax=6, log=None):
    """Plain-LNS move logs: for every seed a notch descent whose every round
    is logged by :class:`LoggingOperator`."""
    X, t = [], []
    for seed in seeds:
        b = make_backend()
        try:
            sc = RegionScorer(b)
            op = LoggingOperator(sc)
            s = Searcher(b, seed=seed, on_round=op.outcome)
            s.load(np.zeros(0, dtype=np.int64))
            s.greedy_fill(10 ** 9)
            s.peel()
            used = 0
            while used < rounds and s.uncov == 0 and len(s.code) > 1:
                M = len(s.code) - 1
                r0 = s.stats["rounds"]
                s.ruin(1, "low")
                s.lns(M, kmin=kmin, kmax=kmax, max_rounds=rounds - used,
                      operator=op)
                used += s.stats["rounds"] - r0
            X.extend(op.X)
            t.extend(op.t)
            if log:
                log(f"seed {seed}: {len(op.X)} records, size {len(s.code)}")
        finally:
            b.close()
    return np.array(X, dtype=np.float32), np.array(t, dtype=np.float32)


def _collect_worker(job):
    """One logging process: its own engine cell (the ABI holds one cell per
    process), plain-LNS notch descent on (axes, R) for every seed."""
    axes, R, seeds, rounds = job
    from .gpu import GpuBackend
    return collect_records(lambda: GpuBackend(axes, R), seeds=seeds, rounds=rounds)


def _collect_job(item):
    i, job = item
    return i, _collect_worker(job)


def collect_records_parallel(cells, seeds_per_cell, rounds, workers=8, log=None):
    """Move logs from ``workers`` processes at once (each with its own cell
    on the one GPU): cell i of ``cells`` ((axes, R) pairs) runs seeds
    100 i + j, j < seeds_per_cell, ``rounds`` logged LNS rounds each.
    Returns (X, t) concatenated in job order (deterministic for a given
    split)."""
    import multiprocessing as mp
    jobs = [(list(axes), int(R), [100 * i + j], int(rounds))
            for i, (axes, R) in enumerate(cells) for j in range(seeds_per_cell)]
    ctx = mp.get_context("spawn")
    parts = [None] * len(jobs)
    t0 = time.perf_counter()
    with ctx.Pool(max(1, min(workers, len(jobs)))) as pool:
        for done, (i, part) in enumerate(pool.imap_unordered(
                _collect_job, list(enumerate(jobs)), chunksize=1)):
            parts[i] = part
            if log:
                log(f"log job {done + 1}/{len(jobs)} done, "
                    f"{time.perf_counter() - t0:.0f} s")
    X = np.concatenate([x for x, _ in parts]) if parts else np.zeros((0, IN - 1), np.float32)
    t = np.concatenate([y for _, y in parts]) if parts else np.zeros(0, np.float32)
    return X, t


def train(net, X, t, epochs=20, batch=256, lr=2e-3, seed=0):
    rng = np.random.default_rng(seed)
    n = len(X)
    losses = []
    for _ in range(epochs):
        order = rng.permutation(n)
        tot, cnt = 0.0, 0
        for lo in range(0, n, batch):
            ix = order[lo:lo + batch]
            tot += net.train_step(X[ix], t[ix], lr) * len(ix)
            cnt += len(ix)
        losses.append(tot / max(cnt, 1))
    return losses


def descend_timed(searcher, seconds, kmin=2, kmax=6, proposer=None):
    """Notch descent under a wall-clock budget; returns [(t, feasible size)]
    (t in seconds from the start of the descent)."""
    s = searcher
    t0 = time.perf_counter()
    hist = [(0.0, len(s.code))]
    while time.perf_counter() - t0 < seconds and s.uncov == 0 and len(s.code) > 1:
        M = len(s.code) - 1
        s.ruin(1, "low")
        left = seconds - (time.perf_counter() - t0)
        s.lns(M, round_budget_s=max(left, 0.0), kmin=kmin, kmax=kmax,
              select="adaptive", proposer=proposer)
        if s.uncov == 0:
            hist.append((time.perf_counter() - t0, len(s.code)))
    return hist


def start_code(make_backend, seed):
    """The common start of both arms: lazy greedy from empty, then peel
    (Searcher seed ``seed``); returns the code."""
    b = make_backend()
    try:
        s = Searcher(b, seed=seed)
        s.load(np.zeros(0, dtype=np.int64))
        s.greedy_fill(10 ** 9)
        s.peel()
        return s.code.copy()
    finally:
        b.close()


def rank_auc(score, label):
    """Probability that a random positive scores above a random negative
    (ties count half), from ranks: O(n log n) for held-out sets of any size."""
    from scipy.stats import rankdata
    pos = label == 1
    npos, nneg = int(pos.sum()), int((~pos).sum())
    if not npos or not nneg:
        return None
    r = rankdata(score)
    return float((r[pos].sum() - npos * (npos + 1) / 2) / (npos * nneg))


def time_to_size(hist, size):
    """First time (s) a run's feasible size reached ``size``; None if never."""
    for t, m in hist:
        if m <= size:
            return t
    return None


def save_proposal(path, params, mu, sd, meta):
    """The trained proposal as a numpy archive: network parameters, the
    feature standardisation it was trained with, and a JSON note of how it
    was trained (records, epochs, held-out AUC)."""
    import json
    np.savez(path, params=np.asarray(params, np.float32),
             mu=np.asarray(mu, np.float32), sd=np.asarray(sd, np.float32),
             meta=np.array(json.dumps(meta)))


def load_proposal(path):
    """(params, mu, sd, meta) of a proposal written by :func:`save_proposal`
    (plain arrays only: nothing in the file is executed)."""
    import json
    with np.load(path, allow_pickle=False) as z:
        params = z["params"].astype(np.float32)
        if params.size != NPARAM:
            raise ValueError(f"{path}: {params.size} parameters, expected {NPARAM}")
        return params, z["mu"], z["sd"], json.loads(str(z["meta"]))


def learned_ab(make_backend, train_cells, seeds, seconds, train_rounds=2000,
               epochs=40, ncand=32, tau=0.25, log=None, records=None,
               model=None, save=None, prior=2.0, meta_extra=None, cell=None):
    """Time-to-size of the learned ruin proposal against plain adaptive LNS.

    1. Move logs: plain LNS (fixed operator mix) on every cell of
       ``train_cells`` (callables making backends), ``train_rounds`` rounds
       per cell, every round logged (region features, outcome).
    2. The network is trained on them on the GPU (own kernels).
    3. For every seed, both arms start from the same greedy + peel cover
       and run notch descent for ``seconds`` on ``make_backend``: plain =
       adaptive operator bandit; learned = the same bandit with the network's
       proposal as one more arm (the exact fields still judge every move).
    Per seed: F = the smallest size plain reached, T_plain = when plain first
    reached F, T_learned = when the learned arm first reached F; the ratio
    T_learned / T_plain (< 1: the learned arm got there sooner; None when it
    never did).

    ``model``: a trained proposal (params, mu, sd, meta) from
    :func:`load_proposal`; steps 1 and 2 are then skipped.  ``save``: write
    the proposal trained in step 2 there."""
    if model is not None:
        return _ab_runs(make_backend, seeds, seconds, model, ncand, tau, log,
                        prior=prior, cell=cell)
    t0 = time.perf_counter()
    if records is not None:
        X, t = records                       # a log collected beforehand
    else:
        Xs, ts = [], []
        for i, mk in enumerate(train_cells):
            X, t = collect_records(mk, seeds=[100 + i], rounds=train_rounds)
            Xs.append(X)
            ts.append(t)
        X = np.concatenate(Xs)
        t = np.concatenate(ts)
    collect_s = time.perf_counter() - t0
    # standardised features; target = the round improved (binary); a
    # quarter held out to measure how well the network ranks regions
    mu = X.mean(axis=0)
    sd = np.maximum(X.std(axis=0), 1e-6)
    Xn = ((X - mu) / sd).astype(np.float32)
    y = (t == 1.0).astype(np.float32)
    perm = np.random.default_rng(0).permutation(len(X))
    cut = int(0.75 * len(X))
    tr, te = perm[:cut], perm[cut:]
    net = Net(max_batch=max(1024, ncand), seed=7)
    try:
        t1 = time.perf_counter()
        losses = train(net, Xn[tr], y[tr], epochs=epochs, batch=256, lr=3e-3)
        train_s = time.perf_counter() - t1
        score = net.forward(Xn[te])
        auc = rank_auc(score, y[te])
        meta = {"records": int(len(X)), "improving": int(y.sum()),
                "heldout_auc": None if auc is None else round(auc, 3),
                "collect_s": round(collect_s, 2), "train_s": round(train_s, 2),
                "epochs": int(epochs),
                "train_loss": [round(float(losses[0]), 4), round(float(losses[-1]), 4)]}
        meta.update(meta_extra or {})
        params = net.get_params()
        if save:
            save_proposal(save, params, mu, sd, meta)
    finally:
        net.close()
    if not seeds:
        return meta
    return _ab_runs(make_backend, seeds, seconds, (params, mu, sd, meta), ncand,
                    tau, log, prior=prior, cell=cell)


def _arm(b, seed, code, seconds, net, model, ncand, tau, prior):
    """One arm of the A/B on backend ``b``: notch descent from ``code`` for
    ``seconds``; with ``net`` the proposal is one more bandit arm."""
    _, mu, sd, _ = model
    s = Searcher(b, seed=seed)
    s.load(code)
    prop = (LearnedProposer(net, RegionScorer(b), ncand=ncand, tau=tau,
                            norm=(mu, sd), prior=prior) if net is not None else None)
    hist = descend_timed(s, seconds, proposer=prop)
    return {"final_size": int(min(m for _, m in hist)),
            "rounds": int(s.stats["rounds"]),
            "hist": [(round(a, 3), int(m)) for a, m in hist]}


def _arm_job(job):
    """An A/B arm in its own process (its own engine cell and network)."""
    axes, R, seed, code, seconds, arm, model, ncand, tau, prior = job
    from .gpu import GpuBackend
    b = GpuBackend(axes, R)
    net = None
    try:
        if arm == "learned":
            net = Net(max_batch=max(1024, ncand), seed=7)
            net.set_params(model[0])
        return arm, _arm(b, seed, code, seconds, net, model, ncand, tau, prior)
    finally:
        if net is not None:
            net.close()
        b.close()


def _ab_runs(make_backend, seeds, seconds, model, ncand, tau, log, prior=2.0,
             cell=None):
    """Step 3 of :func:`learned_ab` for a trained proposal ``model``.  With
    ``cell`` ((axes, R) of the backends) the two arms of a seed run at the
    same time, each in a process of its own on the one GPU (same start
    code, same conditions); otherwise one after the other."""
    params, mu, sd, meta = model
    pool = net = None
    if cell is not None:
        import multiprocessing as mp
        pool = mp.get_context("spawn").Pool(2)
    else:
        net = Net(max_batch=max(1024, ncand), seed=7)
        net.set_params(params)
    try:
        runs = []
        for seed in seeds:
            code = start_code(make_backend, seed)
            res = {"seed": int(seed), "start_size": int(len(code))}
            if pool is not None:
                jobs = [(list(cell[0]), int(cell[1]), int(seed), code, seconds, arm,
                         model, ncand, tau, prior) for arm in ("plain", "learned")]
                for arm, r in pool.map(_arm_job, jobs, chunksize=1):
                    res[arm] = r
            else:
                for arm in ("plain", "learned"):
                    b = make_backend()
                    try:
                        res[arm] = _arm(b, seed, code, seconds,
                                        net if arm == "learned" else None, model,
                                        ncand, tau, prior)
                    finally:
                        b.close()
            F = res["plain"]["final_size"]
            tp = time_to_size(res["plain"]["hist"], F)
            tl = time_to_size(res["learned"]["hist"], F)
            res["target_size"] = F
            res["t_plain_s"] = tp
            res["t_learned_s"] = tl
            res["ratio"] = (None if tl is None or not tp else round(tl / tp, 3))
            runs.append(res)
            if log:
                log(f"seed {seed}: plain {F} at {tp}s, learned "
                    f"{res['learned']['final_size']} (reached {F} at {tl}s)")
        ratios = [r["ratio"] for r in runs]
        done = [x for x in ratios if x is not None]
        out = dict(meta)
        out.update({"ncand": ncand, "tau": tau, "prior": prior,
                    "arms": "concurrent, one process each" if pool is not None
                            else "one after the other"})
        out.update({"seconds_per_arm": seconds, "runs": runs,
                    "median_ratio": (float(np.median(done)) if len(done) == len(ratios)
                                     else None),
                    "reached": f"{len(done)}/{len(ratios)}"})
        return out
    finally:
        if net is not None:
            net.close()
        if pool is not None:
            pool.close()
            pool.join()

