"""Batched multi-cell engine: ctypes binding of the ``ccb_*`` entry points.

Many uneven cells (Hamming, Lee, king torus, open grid) live on the GPU at
once; one persistent launch per sweep level transforms all of them (the
l-th sweep of every cell together), and a batched LNS round
(ruin k random words per cell, greedy recreate by exact gain, keep or
revert) replays as one captured HIP graph.  There is no counterpart in the
reference, whose plugin holds one cell per process (reference
docs/PLUGIN.md:73-76); the per-cell fields are the ones the reference's
``dct_transform`` defines (reference src/include/coldcore/plugin.h:90-104).
The library is the same in-tree ``libcoldcore_covering.so`` as the
singleton backend; there is no CPU fallback.
"""
import ctypes

import numpy as np

from .gpu import load_plugin, locality_order, _check
from .patterns import PROBLEM_IDS, support_patterns, support_patterns_wide
from .space import ProductSpace

_LL = ctypes.c_longlong
_LLP = ctypes.POINTER(ctypes.c_longlong)
_IP = ctypes.POINTER(ctypes.c_int)
_U64P = ctypes.POINTER(ctypes.c_uint64)

#: every batch symbol the library exports
BATCH_SYMBOLS = ("ccb_init", "ccb_free", "ccb_set_code", "ccb_transform",
                 "ccb_read_cnt", "ccb_read_map", "ccb_greedy", "ccb_lns_init",
                 "ccb_lns_run", "ccb_get_code", "ccb_stats", "ccb_info",
                 "ccb_autotune")

KMAX = 8
MAX_R = 3


def _bind_batch(lib):
    if getattr(lib, "_ccb_bound", False):
        return lib
    for f in BATCH_SYMBOLS:
        getattr(lib, f).restype = ctypes.c_int
    lib.ccb_init.argtypes = [ctypes.c_int, _IP, _IP, _IP, _IP, _U64P, _LLP, _IP,
                             ctypes.c_int]
    lib.ccb_set_code.argtypes = [ctypes.c_int, _LLP, ctypes.c_int]
    lib.ccb_transform.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.ccb_read_cnt.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint16)]
    lib.ccb_read_map.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
    lib.ccb_greedy.argtypes = [ctypes.c_int]
    lib.ccb_lns_init.argtypes = [ctypes.c_ulonglong]
    lib.ccb_lns_run.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                ctypes.POINTER(ctypes.c_float)]
    lib.ccb_get_code.argtypes = [ctypes.c_int, ctypes.c_int, _LLP, ctypes.c_int]
    lib.ccb_stats.argtypes = [ctypes.c_int, _LLP]
    lib.ccb_info.argtypes = [_LLP]
    lib.ccb_autotune.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_float)]
    lib._ccb_bound = True
    return lib


def _ints(v):
    return (ctypes.c_int * len(v))(*[int(x) for x in v])


class GpuBatch:
    """A batch of cells resident in HBM.

    cells: sequence of (problem, axes, R) with problem in "hamming" | "lee"
    | "torus_linf" | "grid_linf" and R in 1..3; cap: code words per cell.
    """

    def __init__(self, cells, cap=4096, plugin_path=None):
        self._lib = lib = _bind_batch(load_plugin(plugin_path))
        self.cells = [(p, list(map(int, a)), int(r)) for p, a, r in cells]
        self.spaces = [ProductSpace(a) for _, a, _ in self.cells]
        self.cap = int(cap)
        probs, ns, axes, Rs, fmts, npats, tables = [], [], [], [], [], [], []
        for (p, a, r), sp in zip(self.cells, self.spaces):
            if p not in PROBLEM_IDS:
                raise ValueError(f"unknown problem {p}")
            if not 1 <= r <= MAX_R:
                raise ValueError("batch cells take R in 1..3")
            probs.append(PROBLEM_IDS[p])
            ns.append(len(a))
            axes.extend(a)
            Rs.append(r)
            if p in ("hamming", "lee"):
                wide = len(a) > 16
                t = (locality_order(support_patterns_wide(a, r, p), a, bits=10)
                     if wide else locality_order(support_patterns(a, r, p), a))
                fmts.append(1 if wide else 0)
                npats.append(len(t))
                tables.append(np.asarray(t, dtype=np.uint64))
            else:
                fmts.append(0)
                npats.append(0)
        pats = (np.concatenate(tables) if tables
                else np.zeros(1, dtype=np.uint64))
        pats = np.ascontiguousarray(pats, dtype=np.uint64)
        npl = (_LL * len(npats))(*npats)
        rc = lib.ccb_init(len(self.cells), _ints(probs), _ints(ns), _ints(axes),
                          _ints(Rs), pats.ctypes.data_as(_U64P), npl,
                          _ints(fmts), self.cap)
        _check("ccb_init", rc)
        self._closed = False

    def close(self):
        if not self._closed:
            self._lib.ccb_free()
            self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __len__(self):
        return len(self.cells)

    # ---------------- code and fields ----------------
    def set_code(self, cell, idx):
        idx = np.ascontiguousarray(idx, dtype=np.int64)
        _check("ccb_set_code", self._lib.ccb_set_code(
            int(cell), idx.ctypes.data_as(_LLP), len(idx)))

    def recount(self):
        _check("ccb_transform(code->cnt)", self._lib.ccb_transform(0, 2))

    def gain_maps(self):
        _check("ccb_transform(uncov->gain)", self._lib.ccb_transform(1, 1))

    def loss_maps(self):
        _check("ccb_transform(cnt1->loss)", self._lib.ccb_transform(2, 1))

    def read_cnt(self, cell):
        out = np.empty(self.spaces[cell].size, dtype=np.uint16)
        _check("ccb_read_cnt", self._lib.ccb_read_cnt(
            int(cell), out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16))))
        return out

    def read_map(self, cell):
        out = np.empty(self.spaces[cell].size, dtype=np.int32)
        _check("ccb_read_map", self._lib.ccb_read_map(
            int(cell), out.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))))
        return out

    # ---------------- search ----------------
    def greedy(self, steps):
        """steps greedy steps on every cell; returns the number of cells
        still uncovered."""
        rc = self._lib.ccb_greedy(int(steps))
        if rc < 0:
            raise RuntimeError(f"ccb_greedy rc={rc}")
        return rc

    def greedy_cover(self, chunk=256, limit=1 << 20):
        """greedy steps until every cell is covered; returns the steps run"""
        done = 0
        while done < limit:
            open_cells = self.greedy(chunk)
            done += chunk
            if open_cells == 0:
                return done
        raise RuntimeError("greedy did not cover every cell")

    def lns_init(self, seed=0):
        _check("ccb_lns_init", self._lib.ccb_lns_init(int(seed)))

    def lns_run(self, rounds, k=2, graph=True):
        ms = ctypes.c_float(0.0)
        _check("ccb_lns_run", self._lib.ccb_lns_run(
            int(rounds), int(k), 1 if graph else 0, ctypes.byref(ms)))
        return ms.value

    def code(self, cell, best=False):
        out = np.empty(self.cap, dtype=np.int64)
        m = self._lib.ccb_get_code(int(cell), 1 if best else 0,
                                   out.ctypes.data_as(_LLP), self.cap)
        if m < 0:
            raise RuntimeError(f"ccb_get_code rc={m}")
        return out[:m].copy()

    def stats(self, cell):
        out = (_LL * 6)()
        _check("ccb_stats", self._lib.ccb_stats(int(cell), out))
        keys = ("M", "uncovered", "best", "rounds", "accepted", "reverted")
        return dict(zip(keys, (int(v) for v in out)))

    def info(self):
        out = (_LL * 7)()
        _check("ccb_info", self._lib.ccb_info(out))
        return dict(zip(("cells", "sweeps", "items", "levels", "workgroups",
                         "lds_budget", "threads"), (int(v) for v in out)))

    def autotune(self, reps=5):
        """Time the batched gain transform under each LDS budget and level
        workgroup size and keep the fastest plan; returns (budget bytes, us
        per transform)."""
        us = ctypes.c_float(0.0)
        r = self._lib.ccb_autotune(int(reps), ctypes.byref(us))
        if r < 0:
            raise RuntimeError(f"ccb_autotune rc={r}")
        return int(r), float(us.value)
