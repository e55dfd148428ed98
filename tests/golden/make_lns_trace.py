"""Writes tests/golden/lns_k3_8_2.json: the reference search core's own
trajectory on the config-1 cell, as a fixture for the restated search.

Run in the build container only (it imports the reference checkout, which
never travels anywhere): ``python tests/golden/make_lns_trace.py``.

Cell K_3(8,2) on the reference's numpy backend (reference
src/coldcore/dp.py:66), start = default_rng(0).choice(6561, 40,
replace=False), Searcher seed 1 (reference src/coldcore/search.py:173-208):
load, greedy_fill(10**9), peel(), then notch descent -- ruin(1, "low") and
lns(M - 1, fixed operator mix, kmin 8) until the pass is feasible, then the
next notch -- for 1,000 LNS rounds in total.  The reference's LNS is timed;
the round cap is its own time-out path (search.py:724-727): the searcher's
time_left() reports no time once the cap is reached, so the pass ends at the
top of the next round exactly as a deadline would end it.  Recorded per
round, from the pass's own per-round log line (search.py:774-777): the
target size M, the uncovered count after the accept / reject step and the
code size; at the end the sorted final code and its SHA-256.
"""
import hashlib
import json
import os
import sys

import numpy as np

REF = "/root/reference/src"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lns_k3_8_2.json")
ROUNDS = 1000


def main():
    sys.path.insert(0, REF)
    from coldcore.dp import NumpyDPBackend
    from coldcore.search import Searcher

    b = NumpyDPBackend([3] * 8, 2)
    start = np.random.default_rng(0).choice(3 ** 8, size=40, replace=False)
    trace = []
    holder = {}

    def on_log(msg):
        if msg.startswith("lns M="):
            s = holder["s"]
            M = int(msg.split()[1].split("=")[1])
            trace.append([M, int(s.uncov), int(len(s.code))])

    s = Searcher(b, seed=1, on_log=on_log)
    holder["s"] = s
    s.time_left = lambda: (-np.inf if s.stats["rounds"] >= ROUNDS else np.inf)
    s.load(start)
    s.greedy_fill(10 ** 9)
    greedy = int(len(s.code))
    s.peel()
    peeled = int(len(s.code))
    sizes = [[0, peeled]]
    while s.stats["rounds"] < ROUNDS and s.uncov == 0 and len(s.code) > 1:
        M = len(s.code) - 1
        s.ruin(1, "low")
        s.lns(M, np.inf, kmin=8, log_every=1, select="fixed")
        if s.uncov == 0:
            sizes.append([int(s.stats["rounds"]), int(len(s.code))])
    final = np.sort(s.code.astype(np.int64))
    rec = {"cell": "K_3(8,2)", "start_seed": 0, "searcher_seed": 1,
           "rounds": int(s.stats["rounds"]), "greedy": greedy, "peeled": peeled,
           "sizes": sizes, "final_uncovered": int(s.uncov),
           "final_code": final.tolist(),
           "final_sha256": hashlib.sha256(final.tobytes()).hexdigest(),
           "trace_M_uncov_size": trace}
    with open(OUT, "w") as f:
        json.dump(rec, f, separators=(",", ":"))
    print(f"wrote {OUT}: {len(trace)} rounds, greedy {greedy}, peeled {peeled}, "
          f"final size {len(final)} (uncovered {s.uncov})")


if __name__ == "__main__":
    main()
