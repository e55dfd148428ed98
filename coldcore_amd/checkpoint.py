"""Checkpoint / resume of one cell's search through its solution file.

Behaviour of the reference's job runner on the path that matters here
(reference src/coldcore/campaign.py:106-135): every feasible code the
search establishes is written at once, atomically, to the cell's solution
file (digit rows, ``coldcore_amd.io``); a rerun with ``resume`` starts from
that file instead of a fresh greedy build, so an interrupted run loses at
most the notch in progress and a resumed run never ends above the size it
started from.  ``seed_file`` starts from another file (a known code).

The search itself is the restated caller (``coldcore_amd.search``): exact
lazy greedy to a cover, peel, then notch descent (drop one incumbent, LNS
at M - 1 until feasible).  ``rounds`` bounds the LNS rounds (clock-free and
reproducible: the same seed and backend fields give the same files);
``budget_s`` bounds wall time instead.  The record mirrors the reference's
(M, uncov, solutions as [seconds, M] pairs, solution and log paths).
"""
import os
import time

import numpy as np

from .io import read_solution, write_solution_atomic
from .search import Searcher, descend_rounds


def run_cell(backend, solution_path, seed=1, rounds=1000, budget_s=None,
             resume=True, seed_file=None, log_path=None, kmin=8,
             select="fixed"):
    """Search one cell on ``backend`` (any object with the Backend
    protocol and a ``space`` with ``axes``), checkpointing every feasible
    code to ``solution_path``.  Returns the run record."""
    if rounds is None and budget_s is None:
        raise ValueError("run_cell needs a round bound or a time budget")
    t0 = time.time()
    rec = {"M": None, "uncov": None, "solutions": [], "resumed_from": None,
           "rounds": 0, "solution_file": os.fspath(solution_path),
           "log_file": log_path}
    log = open(log_path, "a", buffering=1) if log_path else None

    def say(msg):
        if log is not None:
            log.write(f"[{time.time() - t0:8.1f}s] {msg}\n")

    last = {"code": None}

    def on_solved(M, code):
        code = np.asarray(code, dtype=np.int64)
        last["code"] = code.copy()
        rec["M"] = int(M)
        rec["solutions"].append([round(time.time() - t0, 2), int(M)])
        write_solution_atomic(solution_path, backend.index_word(code))
        say(f"SOLVED M={M}")

    try:
        deadline = None if budget_s is None else t0 + float(budget_s)
        s = Searcher(backend, seed=seed, deadline=deadline, on_log=say,
                     on_solved=on_solved)
        start = seed_file or (solution_path if resume and
                              os.path.exists(solution_path) else None)
        if start is not None:
            digits = read_solution(start, axes=backend.space.axes)
            s.load(digits)
            rec["resumed_from"] = os.fspath(start)
            say(f"resumed from {start} at M={len(s.code)} (uncov={s.uncov})")
            if s.uncov != 0:
                before = s.stats["rounds"]
                s.lns(len(s.code), max_rounds=rounds, kmin=kmin, select=select)
                rec["rounds"] += s.stats["rounds"] - before
        else:
            s.load(np.zeros(0, dtype=np.int64))
            s.greedy_fill(10 ** 9)
        if s.uncov == 0:
            s.peel()
            on_solved(len(s.code), s.code)
            left = rounds - rec["rounds"] if rounds is not None else 10 ** 12
            used, _ = descend_rounds(s, max(0, left), kmin=kmin, select=select)
            rec["rounds"] += used
            if s.uncov != 0:          # the last notch failed: back to the file
                s.load(last["code"])
        else:
            say("no feasible cover established")
        rec["uncov"] = int(s.uncov)
        if rec["M"] is None:
            rec["M"] = len(s.code)
    finally:
        rec["secs"] = round(time.time() - t0, 2)
        say(f"done M={rec['M']} uncov={rec['uncov']} in {rec['secs']}s")
        if log is not None:
            log.close()
    return rec
