"""Benchmark: field passes/s and achieved HBM GB/s on K_2(24,3).

Workload (the hot path at its published single-GPU scale): the binary
covering-code cell K_2(24,3), |X| = 2^24 words, R = 3 (4 int32 distance
layers), code = flatnonzero(default_rng(0).random(2^24) < 1e-4) (1,678
words, synthetic).  One step = one full exact gain-field pass
(dct_transform(1, 1): predicate cnt < mu seeded into layer 0, all 24 axis
passes, layer sum into the map).  Inputs are resident in HBM before the
timed region; each pass is one ABI call that synchronises on return, so
the time is the real per-call wall time through the C ABI.

Also reported: loss and recount passes, 10^5 single-word updates
(alternating +1 / -1, seed 1), 10^6 exact gain queries (seed 2, chunks of
65,536), the per-sweep schedule with its algorithmic bytes, the roofline of
the dominant sweep kernel (HIP start / stop events stamped by the kernel
launch on the engine's stream), and the CPU
baseline: the numpy axis-DP oracle on one core over a bounded sample.
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0        # MI355X HBM3E, published

#: the proposal network trained on 10^6 logged plain-LNS rounds (half
#: K_3(10,2), half K_4(8,2)) by the engine's own MLP kernels:
#: ``python bench.py --only learned-train`` writes it
PROPOSAL = os.path.join(ROOT, "coldcore_amd", "proposal_k3_10_2.npz")
#: the proposal for radius-4 binary cells (banded windows), trained on plain-
#: LNS logs of K_2(20,4): ``python bench.py --only learned-train-r4``
PROPOSAL_R4 = os.path.join(ROOT, "coldcore_amd", "proposal_k2_r4.npz")


def sweep_bytes(plan):
    """Algorithmic bytes per word of each sweep of one pass, from the
    engine's own schedule (dct_sweep_plan2): what the sweep must read (its
    seed -- the uint16 cnt predicate or the int32 code indicator -- or the
    previous sweep's compact layer words) plus what it must write (compact
    layer words, or the int32 map / uint16 cnt)."""
    return [sw["bytes_in"] + sw["bytes_out"] for sw in plan]


def kernel_name(plan, i, q):
    """The kernel a sweep launches (template instantiations are distinct
    kernels to the profiler)."""
    sw = plan[i]
    if sw["W"] == 32:
        return "k_sweep / k_sweep_r (unpacked layers)"
    last = i == len(plan) - 1
    name = {1: "k_psweep", 2: "k_pcol", 3: "k_pcol", 4: "k_pfirst",
            5: "k_pfirstp", 6: "k_pfirstw", 7: "k_pstrip"}[sw["kernel"]]
    return (f"{name}<Q={q},RR={sw['RR']},W={sw['W']},ND={sw['ND']},"
            f"{'sum' if last else 'compact'},{sw['bytes_in']}+{sw['bytes_out']}B>")


def describe(plan):
    return [f"axes {s['first_axis']}..{s['first_axis'] + s['axes'] - 1}"
            f" P={s['P']} TL={s['TL']} RR={s['RR']} W={s['W']} ND={s['ND']}"
            f" kernel={s['kernel']}"
            f" {s['bytes_in']}+{s['bytes_out']} B/word LDS={s['lds_bytes']}"
            for s in plan]


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor()


def run_cpu_baseline(code, passes):
    """numpy axis-DP oracle, one thread, a bounded sample of gain passes
    (the median of ``passes`` >= 3 after one warm-up)."""
    from oracle.dp import NumpyDP
    passes = max(3, passes)
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    c = NumpyDP([2] * 24, 3)
    c.recount(code)
    c.gain_map()                      # warm-up
    ts = []
    for _ in range(passes):
        t = time.perf_counter()
        c.gain_map()
        ts.append(time.perf_counter() - t)
    dt = float(np.median(ts))
    return dict(value=round(1.0 / dt, 4), unit="field passes/s", cores=1,
                kind="port",
                sample=f"median of {passes} gain passes of K_2(24,3) on the numpy "
                       f"axis-DP oracle after one warm-up, 1 thread "
                       f"({dt * 1e3:.0f} ms/pass, range "
                       f"{min(ts) * 1e3:.0f}-{max(ts) * 1e3:.0f}); host {cpu_model()}, "
                       f"os.cpu_count()={os.cpu_count()}")


#: the reference's published single-thread numpy gain-transform times
#: (reference docs/benchmarks.md:127-144): cell -> ms
CPU_CALIB_CELLS = [
    ("king 15x15 torus R=1", "torus_linf", [15, 15], 1, 0.04),
    ("lee 9x9x9 R=2", "lee", [9, 9, 9], 2, 0.12),
    ("K_2(12,2)", "hamming", [2] * 12, 2, 0.59),
    ("K_3(8,2)", "hamming", [3] * 8, 2, 0.48),
    ("K_2(16,3)", "hamming", [2] * 16, 3, 13.3),
    ("torus 64^3 R=2", "torus_linf", [64, 64, 64], 2, 1.6),
    ("K_5(8,3)", "hamming", [5] * 8, 3, 26.3),
    ("K_3(12,3)", "hamming", [3] * 12, 3, 73.1),
    ("K_7(8,3)", "hamming", [7] * 8, 3, 455.0),
]


def run_cpu_calibration(reps=3):
    """The numpy oracle's gain transform (one thread, median of ``reps``
    after a warm-up) on the cells the reference publishes single-thread
    numpy times for, beside those times: the CPU baseline's restatement is
    of the same speed class as the reference's own CPU backend."""
    from oracle.dp import NumpyDP
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    out = []
    for name, prob, axes, R, ref_ms in CPU_CALIB_CELLS:
        c = NumpyDP(axes, R, problem=prob)
        size = c.space.size
        rng = np.random.default_rng(0)
        c.recount(rng.choice(size, max(1, size // 100), replace=False).astype(np.int64))
        c.gain_map()
        ts = []
        for _ in range(reps):
            t = time.perf_counter()
            c.gain_map()
            ts.append(time.perf_counter() - t)
        ms = sorted(ts)[len(ts) // 2] * 1e3
        out.append({"cell": name, "words": int(size), "oracle_gain_ms": round(ms, 3),
                    "reference_cpu_gain_ms": ref_ms, "ratio": round(ms / ref_ms, 2)})
    return {"host": cpu_model(), "threads": 1, "cells": out}


def schedule_key(plan):
    return " | ".join(describe(plan))


TRAFFIC_PROFILE = os.path.join("profiles", "k_pcol_K2_24_3.md")


def measured_traffic(kernel):
    """HBM bytes per launch of a sweep kernel, taken from the copy-calibrated
    rocprofv3 FETCH_SIZE / WRITE_SIZE runs summarised in the profile file
    (not measured by this run: counters need their own profiler runs).
    Returns (bytes, source); (None, reason) when the profile does not cover
    the kernel this run found dominant, so a stale figure is never
    reported."""
    import re
    try:
        text = open(os.path.join(ROOT, TRAFFIC_PROFILE)).read()
    except OSError:
        return None, f"{TRAFFIC_PROFILE} missing"
    date = re.search(r"^measured: (.+)$", text, re.M)
    t = re.search(r"^traffic\[" + re.escape(kernel) + r"\]: (\d+)$", text, re.M)
    if not t:
        return None, f"{TRAFFIC_PROFILE} has no counter run of {kernel}"
    return int(t.group(1)), (f"{TRAFFIC_PROFILE}, kernel {kernel}, counters "
                             f"measured {date.group(1) if date else '(undated)'}")


def run_config1(GpuBackend, with_cpu):
    """K_3(8,2): seeded 40-word start, fields, lazy greedy, peel, 1,000 LNS
    rounds of notch descent, on the engine (and the numpy oracle)."""
    from coldcore_amd.search import config1_run
    start = np.random.default_rng(0).choice(3 ** 8, size=40, replace=False)
    out = {}
    g = GpuBackend([3] * 8, 2)
    try:
        t = time.perf_counter()
        r = config1_run(g, start, rounds=1000)
        out["gpu_wall_s"] = round(time.perf_counter() - t, 3)
    finally:
        g.close()
    out.update(greedy_size=r["greedy"], final_size=int(len(r["final"])),
               final_uncovered=r["uncov"], rounds=r["rounds"],
               best_feasible=r["sizes"][-1][1])
    if with_cpu:
        from oracle.dp import NumpyDP
        t = time.perf_counter()
        c = config1_run(NumpyDP([3] * 8, 2), start, rounds=1000)
        out["cpu_oracle_wall_s"] = round(time.perf_counter() - t, 3)
        out["trajectory_identical"] = bool(
            c["trace"] == r["trace"] and (c["final"] == r["final"]).all())
    return out


def batch_cells64():
    """The 64 uneven cells of the batched workload: Hamming K_3(n,R) for
    n = 8..12 and R = 1..3, Lee 9x9x9 R = 2, king tori k x k for
    k = 9..31 (R = 1), and 25 open-grid boxes with axes drawn from
    {2,3,4,6,8,9,12,16}, n <= 8, at most 2^20 words (seed 64, R = 1)."""
    cells = [("hamming", [3] * n, R) for n in range(8, 13) for R in (1, 2, 3)]
    cells.append(("lee", [9, 9, 9], 2))
    cells += [("torus_linf", [k, k], 1) for k in range(9, 32)]
    rng = np.random.default_rng(64)
    sizes = [2, 3, 4, 6, 8, 9, 12, 16]
    while len(cells) < 64:
        n = int(rng.integers(2, 9))
        axes = [int(x) for x in rng.choice(sizes, n)]
        while int(np.prod(axes)) > 1 << 20:
            axes.remove(max(axes))
        cells.append(("grid_linf", axes, 1))
    return cells


def run_batch64(GpuBackend, rounds=200, k=2, quick=False):
    """64 uneven cells in one batch: greedy sizes, seeded random starts at
    1.5 x greedy (seed = cell index), the batched transform (one persistent
    launch per sweep level) against the same cells' singleton transforms run
    one after the other, and LNS rounds/s with and without the captured HIP
    graph.  quick: no greedy (starts of |X|/64 random words), 20 rounds and
    no singleton comparison -- a short run for the profiler."""
    from coldcore_amd.batch import GpuBatch
    cells = batch_cells64()
    out = {"cells": len(cells),
           "words": int(sum(int(np.prod(a)) for _, a, _ in cells))}
    with GpuBatch(cells, cap=65535) as b:
        # the batch's own schedule autotuner (LDS budget per tile)
        out["autotune_budget"], out["autotune_us"] = b.autotune(reps=5)
        out.update({f"schedule_{k_}": v for k_, v in b.info().items()})
        empty = np.zeros(0, dtype=np.int64)
        for i in range(len(cells)):
            b.set_code(i, empty)
        b.recount()
        if quick:
            rounds = 20
            greedy = [max(2, int(np.prod(a)) // 96) for _, a, _ in cells]
        else:
            t = time.perf_counter()
            out["greedy_steps"] = b.greedy_cover(256)
            out["greedy_s"] = round(time.perf_counter() - t, 3)
            greedy = [b.stats(i)["M"] for i in range(len(cells))]
            out["greedy_words_total"] = int(sum(greedy))
        starts = []
        for i, (_, a, _) in enumerate(cells):
            size = int(np.prod(a))
            m = min(size, 65535, int(round(1.5 * greedy[i])))
            starts.append(np.random.default_rng(i).choice(size, m, replace=False))

        def reset():
            for i, c in enumerate(starts):
                b.set_code(i, c)
            b.recount()
            b.lns_init(0)

        reset()
        b.gain_maps()
        t = time.perf_counter()
        for _ in range(20):
            b.gain_maps()
        out["batched_gain_ms"] = round((time.perf_counter() - t) * 50, 4)
        # LNS: direct launches, then the captured graph from the same start
        reset()
        b.lns_run(1, k=k, graph=False)
        reset()
        ms = b.lns_run(rounds, k=k, graph=False)
        out["lns_rounds"] = rounds
        out["lns_k"] = k
        out["lns_direct_rounds_per_s"] = round(rounds / (ms * 1e-3), 1)
        direct = [b.stats(i) for i in range(len(cells))]
        reset()
        b.lns_run(1, k=k, graph=True)             # capture
        reset()
        ms = b.lns_run(rounds, k=k, graph=True)
        out["lns_graph_rounds_per_s"] = round(rounds / (ms * 1e-3), 1)
        graph = [b.stats(i) for i in range(len(cells))]
        out["graph_equals_direct"] = bool(direct == graph)
        out["cells_covered_at_end"] = int(sum(s["uncovered"] == 0 for s in graph))
        out["best_words_total_covered"] = int(sum(
            s["best"] for s in graph if s["best"] > 0))
    if quick:
        return out
    # the same cells one at a time through the singleton ABI (the cells
    # whose axes its v1 tables address: axes <= 16; the tori with k > 16
    # are left out, which only understates the batch's advantage)
    tot = 0.0
    nsingle = 0
    for i, (p, a, R) in enumerate(cells):
        if max(a) > 16:
            continue
        nsingle += 1
        g = GpuBackend(a, R, problem=p)
        try:
            g.recount(starts[i].astype(np.int64))
            g.gain_map()
            t = time.perf_counter()
            for _ in range(5):
                g.gain_map()
            tot += (time.perf_counter() - t) / 5
        finally:
            g.close()
    out["singleton_cells"] = nsingle
    out["singleton_gain_ms_sum"] = round(tot * 1e3, 4)
    out["batched_speedup"] = round(tot * 1e3 / out["batched_gain_ms"], 2)
    return out


def run_census(GpuBackend, reps=60):
    """The search loop's query pairs on K_2(24,3) (the headline code): a
    gain pass followed by map_max, and a recount followed by the deficit
    (uncovered), each timed through the ABI with the field scans and with
    the census folded into the transform (dct_set_fold); both answers are
    checked against a host scan of the fields."""
    size = 1 << 24
    code = np.flatnonzero(np.random.default_rng(0).random(size) < 1e-4).astype(np.int64)
    out = {"cell": "K_2(24,3)", "code_words": int(len(code)), "calls": reps}
    g = GpuBackend([2] * 24, 3)
    try:
        g.recount(code)
        g.gain_map()
        ref_max = max(0, int(g.map_all().max()))
        ref_unc = int((g.read_cnt() == 0).sum())
        for fold in (False, True):
            g.set_fold(fold)
            g.recount(code)
            assert g.uncovered() == ref_unc
            g.gain_map()
            assert g.map_max() == ref_max
            tg, tr = [], []
            for _ in range(reps):
                t = time.perf_counter()
                g.gain_map()
                g.map_max()
                tg.append(time.perf_counter() - t)
            for _ in range(reps // 2):
                t = time.perf_counter()
                g.recount(code)
                g.uncovered()
                tr.append(time.perf_counter() - t)
            key = "folded" if fold else "scans"
            out[key] = {"gain_plus_map_max_us": round(float(np.median(tg)) * 1e6, 1),
                        "recount_plus_uncovered_us": round(float(np.median(tr)) * 1e6, 1)}
        g.set_fold(False)
    finally:
        g.close()
    return out


def run_batch64_cells(top=16):
    """Where the 64-cell batched transform's time goes: the autotuned
    batched gain pass of each family's cells on their own, and of every cell
    alone (its critical path inside the full batch), slowest first (starts:
    |X| / 96 random words per cell, seed = cell index)."""
    from coldcore_amd.batch import GpuBatch
    cells = batch_cells64()

    def run(sub):
        with GpuBatch([c for _, c in sub], cap=65535) as b:
            for j, (i, (_, a, _)) in enumerate(sub):
                size = int(np.prod(a))
                b.set_code(j, np.random.default_rng(i).choice(
                    size, max(2, size // 96), replace=False))
            b.recount()
            b.autotune(reps=5)
            b.gain_maps()
            t = time.perf_counter()
            for _ in range(20):
                b.gain_maps()
            info = b.info()
            return round((time.perf_counter() - t) * 5e4, 1), info

    indexed = list(enumerate(cells))
    out = {"families": {}, "cells": []}
    for name, fams in (("all", None), ("hamming", ("hamming",)),
                       ("grid", ("grid_linf",)), ("torus+lee", ("torus_linf", "lee"))):
        sub = [x for x in indexed if fams is None or x[1][0] in fams]
        us, info = run(sub)
        out["families"][name] = {"cells": len(sub), "us": us, "levels": info["levels"],
                                 "threads": info["threads"],
                                 "lds_budget": info["lds_budget"]}
    rows = []
    for x in indexed:
        us, info = run([x])
        rows.append({"cell": x[0], "problem": x[1][0], "axes": x[1][1], "R": x[1][2],
                     "us": us, "levels": info["levels"], "threads": info["threads"],
                     "lds_budget": info["lds_budget"]})
    out["cells"] = sorted(rows, key=lambda r: -r["us"])[:top]
    return out


def run_config2_cell(GpuBackend, axes, R, steps=10):
    """The plain-transform recipe on one cell: code = flatnonzero(
    default_rng(0).random(|X|) < 1e-4); gain, loss and recount passes
    (median of ``steps`` after a warm-up that also tunes the schedule);
    10^5 single-word updates and 10^6 exact gain queries (run_walks)."""
    from coldcore_amd.patterns import ball_size
    size = int(np.prod(np.array(axes, dtype=np.int64)))
    code = np.flatnonzero(np.random.default_rng(0).random(size) < 1e-4).astype(np.int64)
    out = {"axes": list(axes), "R": R, "words": size, "code_words": int(len(code))}
    g = GpuBackend(axes, R)
    try:
        g.recount(code)
        for name, fn in (("gain", g.gain_map), ("loss", g.loss_map),
                         ("recount", lambda: g.recount(code))):
            fn()
            ts = []
            for _ in range(steps):
                t = time.perf_counter()
                fn()
                ts.append(time.perf_counter() - t)
            out[f"{name}_pass_ms"] = round(float(np.median(ts)) * 1e3, 4)
        out["gain_bytes_per_word"] = sum(sweep_bytes(g.sweep_plan2(0)))
        out["recount_bytes_per_word"] = sum(sweep_bytes(g.sweep_plan2(1)))
        ball = ball_size(axes[0], len(axes), R) if len(set(axes)) == 1 else None
        out.update(run_walks(g, code, size, ball=ball or 0))
    finally:
        g.close()
    return out


def run_config5(GpuBackend, streams=8, slice_rounds=1, max_restarts=4,
                ruin_frac=0.01, kmax=64, proposal=PROPOSAL_R4):
    """The largest single-GPU space: K_2(30,4), 2^30 words, the seeded
    40,000-word random start (default_rng(0).choice); its recount and the
    full gain / loss field passes, then the restart portfolio at M = 40,000
    (``streams`` time-shared restarts of plain LNS with a shared incumbent,
    ``slice_rounds`` rounds per slice, ``max_restarts`` restarts, a restart
    kicks ``ruin_frac`` M incumbents, LNS ruins 8..``kmax`` words; reference
    search.py:1536-1620), all on the HBM-resident cell."""
    from coldcore_amd.search import Searcher
    size = 1 << 30
    start = np.random.default_rng(0).choice(size, 40000, replace=False).astype(np.int64)
    out = {"words": size, "start_words": int(len(start)), "streams": streams,
           "slice_rounds": slice_rounds, "max_restarts": max_restarts,
           "ruin_frac": ruin_frac, "kmin": 8, "kmax": kmax}
    g = GpuBackend([2] * 30, 4)
    try:
        s = Searcher(g, seed=1)
        t = time.perf_counter()
        s.load(start)
        out["load_recount_ms"] = round((time.perf_counter() - t) * 1e3, 2)
        out["start_uncovered"] = int(s.uncov)
        for name, fn in (("gain", g.gain_map), ("loss", g.loss_map)):
            fn()
            t = time.perf_counter()
            fn()
            out[f"{name}_pass_ms"] = round((time.perf_counter() - t) * 1e3, 2)
        t = time.perf_counter()
        u = s.portfolio(len(start), streams=streams, slice_rounds=slice_rounds,
                        max_restarts=max_restarts, kmin=8, kmax=kmax,
                        ruin_frac=ruin_frac)
        out["portfolio_s"] = round(time.perf_counter() - t, 2)
        out["lns_rounds"] = int(s.stats["rounds"])
        out["restarts"] = int(s.stats["restarts"])
        out["incumbent_uncovered"] = int(u)
        out["incumbent_words"] = int(len(s.code))
        out["recount_check"] = bool(s.verify_full() == u)
        if proposal and os.path.exists(proposal):
            # the same portfolio from the same start with the learned
            # proposal as one more arm of each stream's adaptive operator
            # bandit (the exact fields judge every move)
            from coldcore_amd.learned import (LearnedProposer, Net, RegionScorer,
                                              load_proposal)
            params, mu, sd, meta = load_proposal(proposal)
            net = Net(max_batch=1024, seed=7)
            try:
                net.set_params(params)
                s = Searcher(g, seed=1)
                s.load(start)
                prop = LearnedProposer(net, RegionScorer(g), norm=(mu, sd))
                t = time.perf_counter()
                u = s.portfolio(len(start), streams=streams,
                                slice_rounds=slice_rounds,
                                max_restarts=max_restarts, kmin=8, kmax=kmax,
                                ruin_frac=ruin_frac, select="adaptive",
                                proposer=prop)
                out["learned"] = {
                    "proposal": os.path.relpath(proposal, ROOT),
                    "trained_on": meta.get("cells"), "records": meta.get("records"),
                    "heldout_auc": meta.get("heldout_auc"),
                    "portfolio_s": round(time.perf_counter() - t, 2),
                    "lns_rounds": int(s.stats["rounds"]),
                    "learned_rounds": int(s.stats.get("learned_rounds", 0)),
                    "restarts": int(s.stats["restarts"]),
                    "incumbent_uncovered": int(u),
                    "incumbent_words": int(len(s.code)),
                    "recount_check": bool(s.verify_full() == u)}
            finally:
                net.close()
    finally:
        g.close()
    return out


def run_walks(g, code, size, n_upd=100000, n_q=10 ** 6, ball=2325):
    """Incremental updates and exact queries on a recounted K_2(24,3) cell:
    n_upd single-word updates alternating +1 (a random non-code word, seed
    1) and -1 (a random current codeword), 20 x (+1, -1) batches of 1,024
    words, and n_q exact gain queries (seed 2) in chunks of 65,536."""
    out = {}
    rng = np.random.default_rng(1)
    live = list(code.tolist())
    liveset = set(live)
    adds = rng.integers(0, size, n_upd)
    picks = rng.random(n_upd)
    g.count_eq(0)
    t = time.perf_counter()
    for i in range(n_upd):
        if i % 2 == 0:
            w = int(adds[i])
            while w in liveset:
                w = (w + 1) % size
            g.ball_update(np.array([w], dtype=np.int64), +1)
            live.append(w)
            liveset.add(w)
        else:
            j = int(picks[i] * len(live))
            w = live[j]
            live[j] = live[-1]
            live.pop()
            liveset.discard(w)
            g.ball_update(np.array([w], dtype=np.int64), -1)
    g.count_eq(0)                          # drains the stream
    du = time.perf_counter() - t
    out["updates_per_s"] = round(n_upd / du, 1)
    out["update_us"] = round(du / n_upd * 1e6, 3)
    wb = rng.choice(size, 1024, replace=False).astype(np.int64)
    t = time.perf_counter()
    for _ in range(20):
        g.ball_update(wb, +1)
        g.ball_update(wb, -1)
    g.count_eq(0)
    out["update_words_per_s_batch1024"] = round(
        40 * 1024 / (time.perf_counter() - t), 1)
    q = np.random.default_rng(2).integers(0, size, n_q).astype(np.int64)
    g.ball_gather(q[:65536], 0)
    t = time.perf_counter()
    for lo in range(0, len(q), 65536):
        g.ball_gather(q[lo:lo + 65536], 0)
    dq = time.perf_counter() - t
    out["queries_per_s"] = round(n_q / dq, 1)
    out["query_cnt_reads_per_s"] = round(n_q * ball / dq, 1)
    return out


def run_near_cover(GpuBackend):
    """Set-side gathers and the swap-evaluation fast path against the walk
    (CC_SET_POLICY=0) on near-cover states: K_2(24,3) with 60,000 random
    code words (seed 3; a few thousand words stay uncovered), 65,536 gain
    queries (seed 4); K_5(8,3) with 1,000 random code words (seed 5), one
    polish move = swap_eval of 1 drop and 110 candidates (seed 6)."""
    out = {}
    size = 1 << 24
    code = np.random.default_rng(3).choice(size, 60000, replace=False)
    code = code.astype(np.int64)
    cands = np.random.default_rng(4).integers(0, size, 65536).astype(np.int64)
    res = {}
    try:
        for pol in ("1", "0"):
            os.environ["CC_SET_POLICY"] = pol
            g = GpuBackend([2] * 24, 3)
            try:
                g.recount(code)
                out["k2_24_3_uncovered"] = int(g.count_eq(0))
                t = time.perf_counter()
                r = g.ball_gather(cands, 0)            # lists Z on the set side
                first = (time.perf_counter() - t) * 1e3
                t = time.perf_counter()
                for _ in range(10):
                    r = g.ball_gather(cands, 0)
                res[pol] = (first, (time.perf_counter() - t) * 100, r)
            finally:
                g.close()
        out["gather65536_set_side_first_ms"] = round(res["1"][0], 3)
        out["gather65536_set_side_ms"] = round(res["1"][1], 3)
        out["gather65536_walk_ms"] = round(res["0"][1], 3)
        out["gather_identical"] = bool((res["1"][2] == res["0"][2]).all())
        s5 = 5 ** 8
        code5 = np.random.default_rng(5).choice(s5, 1000, replace=False)
        code5 = code5.astype(np.int64)
        rng = np.random.default_rng(6)
        drops = code5[:1]
        cand5 = [rng.integers(0, s5, 110).astype(np.int64)]
        for pol in ("1", "0"):
            os.environ["CC_SET_POLICY"] = pol
            g = GpuBackend([5] * 8, 3)
            try:
                g.recount(code5)
                ref = g.swap_eval(drops, cand5)
                t = time.perf_counter()
                for _ in range(50):
                    got = g.swap_eval(drops, cand5)
                us = (time.perf_counter() - t) * 1e6 / 50
                key = "fast" if pol == "1" else "fallback"
                out[f"swap_eval_k5_8_3_{key}_us"] = round(us, 1)
                res["swap" + pol] = (got[0].tolist(), [x.tolist() for x in got[1]])
                assert res["swap" + pol][0] == ref[0].tolist()
            finally:
                g.close()
        out["swap_identical"] = res["swap1"] == res["swap0"]
    finally:
        os.environ.pop("CC_SET_POLICY", None)
    return out




def run_learned_saved(GpuBackend, path=PROPOSAL, seconds=30.0,
                      seeds=(0, 1, 2, 3, 4), log=None):
    """Time-to-size A/B of the saved proposal against plain adaptive LNS on
    K_3(10,2): per seed both arms from the same greedy + peel cover, run at
    the same time in two processes on the one GPU for ``seconds`` each."""
    from coldcore_amd.learned import learned_ab, load_proposal
    model = load_proposal(path)
    out = learned_ab(lambda: GpuBackend([3] * 10, 2), [], list(seeds), seconds,
                     model=model, log=log, cell=([3] * 10, 2))
    out["proposal"] = os.path.relpath(path, ROOT)
    return _ab_summary(out)


def _ab_summary(out):
    for r in out["runs"]:                 # keep the JSON line short
        for arm in ("plain", "learned"):
            r[arm].pop("hist", None)
    out.pop("median_ratio")
    # median over the seeds, a seed whose learned arm never reached the
    # plain arm's final size counting as infinitely late
    ratios = sorted(np.inf if r["ratio"] is None else r["ratio"] for r in out["runs"])
    med = float(np.median(ratios))
    out["time_to_size_ratio"] = None if not np.isfinite(med) else round(med, 3)
    out["per_seed_ratio"] = [r["ratio"] for r in out["runs"]]
    out["learned_final_smaller_or_equal"] = int(sum(
        r["learned"]["final_size"] <= r["plain"]["final_size"] for r in out["runs"]))
    return out


def run_learned(GpuBackend, seconds=10.0, seeds=(0, 1, 2, 3, 4), rounds=12500,
                workers=8, epochs=8, per_cell=10, log=None, save=None):
    """Time-to-size of the learned ruin proposal against plain adaptive LNS
    on K_3(10,2): move logs of plain LNS on K_3(10,2) and K_4(8,2) (every
    round logged: region features from the exact layer fields, outcome),
    ``workers`` logging processes with their own cells, ``rounds`` rounds
    per process; the network trained on the GPU by its own kernels; then
    per seed both arms from the same greedy + peel cover for ``seconds`` of
    notch descent (the exact fields judge every move in both)."""
    from coldcore_amd.learned import collect_records_parallel, learned_ab
    t = time.perf_counter()
    per_cell = per_cell or max(1, workers // 2)
    X, y = collect_records_parallel([([3] * 10, 2), ([4] * 8, 2)], per_cell,
                                    rounds, workers=workers, log=log)
    collect_s = time.perf_counter() - t
    out = learned_ab(lambda: GpuBackend([3] * 10, 2), [], list(seeds), seconds,
                     epochs=epochs, records=(X, y), log=log, save=save)
    out["collect_s"] = round(collect_s, 2)
    out["logging_processes"] = workers
    out["log_jobs"] = 2 * per_cell
    out["rounds_per_job"] = rounds
    if not seeds:
        return out
    return _ab_summary(out)


def run_learned_train(cells, per_cell, rounds, save, workers=8, epochs=8,
                      log=None):
    """Move logs of plain LNS on ``cells`` ((axes, R) pairs, ``per_cell``
    logging jobs of ``rounds`` rounds each, ``workers`` processes), the
    network trained on them on the GPU, saved to ``save`` (no A/B)."""
    from coldcore_amd.learned import collect_records_parallel, learned_ab
    t = time.perf_counter()
    X, y = collect_records_parallel(cells, per_cell, rounds, workers=workers, log=log)
    collect_s = time.perf_counter() - t
    meta = learned_ab(None, [], [], 0.0, epochs=epochs, records=(X, y), save=save,
                      meta_extra={"cells": [[list(a), int(R)] for a, R in cells],
                                  "collect_s": round(collect_s, 2)})
    return meta


def copy_calibration(mib):
    """hipMemcpy device-to-device of a known size: its dispatch
    (__amd_rocclr_copyBuffer) reads and writes exactly mib MiB."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    n = mib << 20
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(n)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(n)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(n)) == 0
    assert hip.hipDeviceSynchronize() == 0
    # 3 untimed copies, then 20 between HIP events on the null stream (a
    # copy of a few tens of MiB takes about as long as a host round trip)
    e0, e1 = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipEventCreate(ctypes.byref(e0)) == 0
    assert hip.hipEventCreate(ctypes.byref(e1)) == 0
    for r in range(23):
        if r == 3:
            assert hip.hipEventRecord(e0, None) == 0
        assert hip.hipMemcpyAsync(b, a, ctypes.c_size_t(n), 3, None) == 0   # D2D
    assert hip.hipEventRecord(e1, None) == 0
    assert hip.hipEventSynchronize(e1) == 0
    ms = ctypes.c_float(0.0)
    assert hip.hipEventElapsedTime(ctypes.byref(ms), e0, e1) == 0
    hip.hipEventDestroy(e0)
    hip.hipEventDestroy(e1)
    hip.hipFree(a)
    hip.hipFree(b)
    us = ms.value * 1e3 / 20
    print(json.dumps({"calib_copy_bytes": n, "copies": 23, "timed_copies": 20,
                      "copy_us": round(us, 2),
                      "copy_GBps_read_plus_write": round(2 * n / (us * 1e-6) / 1e9, 1)}),
          flush=True)


#: further configured cells, timed as gain passes (name, axes, R, code recipe)
OTHER_CELLS = [
    ("K_5(10,3)", [5] * 10, 3, ("p", 1e-4)),
    ("mixed [4,5,6,7,8,4,5,6,7] R=3", [4, 5, 6, 7, 8, 4, 5, 6, 7], 3, ("p", 1e-4)),
    ("K_2(30,4)", [2] * 30, 4, ("choice", 40000)),
    # the reference's own published GPU gain-transform cells
    ("K_8(9,4)", [8] * 9, 4, ("p", 2e-6)),
    ("K_10(9,5)", [10] * 9, 5, ("p", 2e-6)),
    ("K_3(8,2)", [3] * 8, 2, ("choice", 40)),
    ("K_3(10,2)", [3] * 10, 2, ("choice", 400)),
    ("K_4(8,2)", [4] * 8, 2, ("choice", 400)),
    # L-inf cells with axes beyond the v1 nibble (long-axis tables)
    ("king torus 1024x1024 R=1", [1024, 1024], 1, ("p", 1e-3), "torus_linf"),
    ("grid 1000x1000 R=3", [1000, 1000], 3, ("p", 1e-3), "grid_linf"),
]


#: the reference's published gain-transform times on one H100-class GPU
#: (GH200), v0.3 pair schedule, and its bytes moved (reference
#: docs/performance.md:386-391, docs/benchmarks.md:103-104)
REFERENCE_GPU_MS = {"K_8(9,4)": (6.1, 18.5e9), "K_10(9,5)": (61.0, 154e9)}


def time_cells(GpuBackend, steps):
    out = []
    for name, axes, R, (kind, arg), *prob in OTHER_CELLS:
        problem = prob[0] if prob else "hamming"
        size = int(np.prod(np.array(axes, dtype=np.int64)))
        rng = np.random.default_rng(0)
        if kind == "p":
            code = np.flatnonzero(rng.random(size) < arg)
        else:
            code = rng.choice(size, arg, replace=False)
        g = GpuBackend(axes, R, problem=problem)
        try:
            g.recount(code.astype(np.int64))
            g.gain_map()                  # the first one tunes the schedule
            plan = g.sweep_plan2(0)
            ts = []
            for _ in range(steps):
                t = time.perf_counter()
                g.gain_map()
                ts.append(time.perf_counter() - t)
            ms = float(np.median(ts)) * 1e3
            bpw = sum(sweep_bytes(plan))
            gbs = bpw * size / (ms * 1e-3) / 1e9
            row = {"cell": name, "words": size, "code_words": int(len(code)),
                   "gain_pass_ms": round(ms, 4),
                   "passes_per_s": round(1e3 / ms, 1),
                   "bytes_per_word": bpw,
                   "GBps_algorithmic": round(gbs, 1),
                   "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 3),
                   "sweeps": len(plan), "schedule": describe(plan)}
            if name in REFERENCE_GPU_MS:
                ref_ms, ref_bytes = REFERENCE_GPU_MS[name]
                row.update(reference_gain_ms_gh200=ref_ms,
                           reference_bytes=ref_bytes,
                           speedup_vs_reference=round(ref_ms / ms, 2))
            out.append(row)
        finally:
            g.close()
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-passes", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--calib-copy", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB "
                         "(3 times): the reference kernel for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters")
    ap.add_argument("--only", choices=("near_cover", "batch64", "batch64-quick",
                                       "batch64-cells",
                                       "cells", "learned", "learned-1m", "walks",
                                       "cpu_calib", "config5", "census",
                                       "learned-train", "learned-train-r4"),
                    help="run one side section and print its JSON only")
    ap.add_argument("--out", help="learned-train: where to write the proposal "
                    "(default: the package's saved proposal)")
    ap.add_argument("--learned-seconds", type=float, default=30.0,
                    help="seconds per arm of the learned-proposal A/B")
    args = ap.parse_args()
    if args.only == "cpu_calib":
        print(json.dumps({"cpu_calib": run_cpu_calibration()}), flush=True)
        return
    if args.only:
        from coldcore_amd.gpu import GpuBackend
        if args.only == "near_cover":
            r = run_near_cover(GpuBackend)
        elif args.only == "cells":
            r = time_cells(GpuBackend, 10)
        elif args.only == "learned":
            # the saved proposal's A/B (seconds per arm: --learned-seconds)
            lg = lambda m: print(m, file=sys.stderr, flush=True)
            r = (run_learned_saved(GpuBackend, seconds=args.learned_seconds, log=lg)
                 if os.path.exists(PROPOSAL) else
                 run_learned(GpuBackend, seconds=args.learned_seconds, log=lg))
        elif args.only == "learned-1m":
            # config 4 at its configured scale: 10^6 logged rounds (80 jobs
            # of 12,500 rounds over 8 processes), 30 s per arm
            r = run_learned(GpuBackend, seconds=30.0, rounds=12500, epochs=8,
                            per_cell=40,
                            log=lambda m: print(m, file=sys.stderr, flush=True))
        elif args.only == "batch64-cells":
            r = run_batch64_cells()
        elif args.only == "learned-train-r4":
            # the radius-4 proposal: plain-LNS logs of K_2(20,4)
            r = run_learned_train([([2] * 20, 4)], per_cell=16, rounds=12500,
                                  save=args.out or PROPOSAL_R4,
                                  log=lambda m: print(m, file=sys.stderr, flush=True))
        elif args.only == "learned-train":
            # the saved proposal: 10^6 logged rounds, trained, no A/B
            r = run_learned(GpuBackend, seeds=(), rounds=12500, epochs=8,
                            per_cell=40, save=args.out or PROPOSAL,
                            log=lambda m: print(m, file=sys.stderr, flush=True))
        elif args.only == "config5":
            r = run_config5(GpuBackend)
        elif args.only == "census":
            r = run_census(GpuBackend)
        elif args.only == "walks":
            # a short run of the update / query kernels (for the profiler)
            size = 1 << 24
            code = np.flatnonzero(np.random.default_rng(0).random(size) < 1e-4)
            g = GpuBackend([2] * 24, 3)
            try:
                g.recount(code.astype(np.int64))
                r = run_walks(g, code.astype(np.int64), size, 20000, 200000)
            finally:
                g.close()
        else:
            r = run_batch64(GpuBackend, quick=args.only.endswith("quick"))
        print(json.dumps({args.only: r}), flush=True)
        return
    if args.calib_copy:
        copy_calibration(args.calib_copy)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: the hot path runs on a single "
              "MI355X (sharding across GPUs is described, not implemented); "
              "run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from coldcore_amd.gpu import GpuBackend
    n, R = 24, 3
    size = 1 << n
    code = np.flatnonzero(np.random.default_rng(0).random(size) < 1e-4)
    code = code.astype(np.int64)
    g = GpuBackend([2] * n, R)
    try:
        g.recount(code)
        for _ in range(max(1, args.warmup)):
            g.gain_map()                  # the first one tunes the schedule
        plan = g.sweep_plan2(0)           # the schedule the timed passes run
        rplan = g.sweep_plan2(1)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            g.gain_map()
        t1 = time.perf_counter()
        ms = (t1 - t0) * 1e3 / args.steps
        passes_s = 1e3 / ms

        # per-sweep event timing on the engine's stream (separate passes, so
        # the events do not perturb the timed region above)
        g.sweep_profile(True)
        per = []
        for _ in range(args.steps):
            g.gain_map()
            per.append(g.sweep_times())
        g.sweep_profile(False)
        per = np.array(per)                       # (steps, sweeps) ms
        sweep_ms = per.mean(axis=0)
        bpw = sweep_bytes(plan)
        pass_bytes = sum(bpw) * size
        kernel_ms = float(per.sum(axis=1).mean())
        # the dominant kernel: sweeps sharing one kernel instantiation are
        # one kernel to the profiler, so group them and take the group with
        # the largest time; per launch = its bytes and time over its launches
        groups = {}
        for i, sw in enumerate(plan):
            key = kernel_name(plan, i, 2)
            gb = groups.setdefault(key, [0.0, 0, 0])
            gb[0] += float(sweep_ms[i])
            gb[1] += bpw[i] * size
            gb[2] += 1
        dom_key = max(groups, key=lambda k: groups[k][0])
        dom_ms, dom_bytes, dom_n = groups[dom_key]
        dom_gbs = dom_bytes / (dom_ms * 1e-3) / 1e9
        traffic, traffic_src = measured_traffic(dom_key)

        extra = {}
        if not args.no_extra:
            # loss and recount passes
            for name, fn in (("loss", g.loss_map),
                             ("recount", lambda: g.recount(code))):
                fn()
                t = time.perf_counter()
                for _ in range(10):
                    fn()
                extra[f"{name}_pass_ms"] = round(
                    (time.perf_counter() - t) * 100, 4)
            extra["recount_bytes_per_word"] = sum(sweep_bytes(rplan))
            extra.update(run_walks(g, code, size))
    finally:
        g.close()

    rec = {
        "metric": "field passes/s + HBM GB/s on K_2(24,3)",
        "value": round(passes_s, 3),
        "unit": "field passes/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 5),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic: seeded random code, p=1e-4 (1,678 words)",
        "config": {
            "workload": "K_2(24,3) exact gain-field pass, 2^24 words, R=3",
            "axes": "[2]*24", "R": 3, "code_words": int(len(code)),
            "schedule": describe(plan),
        },
        "pass_bytes_algorithmic": int(pass_bytes),
        "pass_bytes_per_word": int(sum(bpw)),
        "pass_GBps_algorithmic": round(pass_bytes / (ms * 1e-3) / 1e9, 1),
        "pass_kernel_ms": round(kernel_ms, 5),
        "sweep_ms": [round(float(x), 5) for x in sweep_ms],
        "sweep_bytes_per_word": bpw,
        "roofline": {
            "bound": "hbm",
            "achieved": round(dom_gbs, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(dom_gbs / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "traffic_source": traffic_src,
            "kernel": f"{dom_key}: {dom_n} launches per pass, "
                      f"{dom_bytes / dom_n / 1e6:.1f} MB algorithmic per launch, "
                      f"{dom_ms / dom_n * 1e3:.1f} us average",
        },
    }
    rec.update(extra)
    t_sec = [time.perf_counter()]

    def section(name):
        # progress on stderr: one line per finished section
        now = time.perf_counter()
        print(f"bench section {name}: {now - t_sec[0]:.1f} s", file=sys.stderr,
              flush=True)
        t_sec[0] = now

    section("K_2(24,3) passes, walks")
    if not args.no_extra:
        rec["other_cells"] = time_cells(GpuBackend, 10)
        section("other cells")
        rec["config2_k5_10_3"] = run_config2_cell(GpuBackend, [5] * 10, 3)
        section("config 2 K_5(10,3)")
        rec["config5_k2_30_4"] = run_config5(GpuBackend)
        section("config 5 K_2(30,4)")
        rec["config1"] = run_config1(GpuBackend, not args.no_cpu)
        section("config 1")
        rec["near_cover"] = run_near_cover(GpuBackend)
        section("near cover")
        rec["census"] = run_census(GpuBackend)
        section("census")
        try:
            rec["batch64"] = run_batch64(GpuBackend)
        except RuntimeError as e:          # reported, never hidden
            rec["batch64"] = {"error": str(e)}
        section("batch64")
        # config 4: the learned proposal's time-to-size against plain LNS
        # (the saved network when there is one; else a shorter log, trained
        # here)
        lg = lambda m: print(m, file=sys.stderr, flush=True)
        rec["learned"] = (run_learned_saved(GpuBackend, seconds=args.learned_seconds,
                                            log=lg)
                          if os.path.exists(PROPOSAL) else
                          run_learned(GpuBackend, seconds=args.learned_seconds, log=lg))
        section("learned")
    if not args.no_cpu:
        rec["cpu_baseline"] = run_cpu_baseline(code, args.cpu_passes)
        rec["speedup_vs_cpu"] = round(passes_s / rec["cpu_baseline"]["value"], 1)
        section("cpu baseline")
        # the oracle against the reference's published single-thread numpy
        # times, on this (the GPU box's) host
        rec["cpu_calib"] = run_cpu_calibration()
        section("cpu calibration")
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
