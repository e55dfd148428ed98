"""GPU backend: ctypes binding of the MI355X engine (plugin ABI v1).

Mirrors reference src/coldcore/gpu.py (``_bind`` 52-103, ``locality_order``
106-129, ``GpuBackend`` 207-537): same class, method names, argument meaning
and error behaviour (a nonzero rc raises ``RuntimeError``; ``-100-e`` is
device error e).  The library is the in-tree ``libcoldcore_covering.so``
built from ``coldcore_amd/csrc`` (or ``COLDCORE_PLUGIN_PATH``).  There is no
CPU fallback: a missing or unloadable library raises.
"""
import ctypes
import os

import numpy as np

from .patterns import (PROBLEM_IDS, support_patterns, support_patterns_long,
                       support_patterns_wide)
from .space import ProductSpace

_LL = ctypes.c_longlong
_LLP = ctypes.POINTER(ctypes.c_longlong)
_I32P = ctypes.POINTER(ctypes.c_int32)
_U64P = ctypes.POINTER(ctypes.c_uint64)
_libs = {}

#: every symbol the library must export (the v1 surface plus the optional
#: and additive ones this engine implements)
REQUIRED = ("dct_init3", "dct_free", "dct_set_code", "dct_transform",
            "dct_ball_update", "dct_ball_gather", "dct_count_eq",
            "dct_map_max", "dct_map_extract", "dct_map_extract64",
            "dct_map_read_at", "dct_read_cnt", "dct_init", "dct_init2")
OPTIONAL = ("dct_map_hist", "dct_set_mu", "dct_get_mu", "dct_ball_gather_mu",
            "dct_count_deficit", "dct_host_bytes", "dct_loss_owner",
            "dct_set_tbl2", "dct_use_fmt", "dct_ball_update_delta",
            "dct_swap_eval", "dct_stats")
ADDITIVE = ("dct_init3w", "dct_init3x", "dct_sweep_plan", "dct_sweep_plan2",
            "dct_plan_probe", "dct_autotune", "dct_sweep_profile",
            "dct_sweep_times", "dct_sweep_force", "dct_map_read",
            "dct_transform_layers", "dct_layers_read_at", "dct_region_features",
            "dct_region_propose", "dct_set_fold")


def _rc_msg(rc):
    if rc <= -100:
        return f"rc={rc} (HIP error {-(rc + 100)})"
    return f"rc={rc}"


def _check(name, rc):
    if rc != 0:
        raise RuntimeError(f"{name} {_rc_msg(rc)}")


def default_plugin_path():
    env = os.environ.get("COLDCORE_PLUGIN_PATH")
    if env:
        return env
    return os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "libcoldcore_covering.so")


def _bind(path):
    if not os.path.exists(path):
        raise RuntimeError(
            f"engine library {path} is missing: run python -m coldcore_amd.build")
    lib = ctypes.CDLL(path)
    for f in REQUIRED + OPTIONAL + ADDITIVE:
        getattr(lib, f)                      # every symbol must resolve
    for f in ("dct_free", "dct_set_code", "dct_transform", "dct_ball_update",
              "dct_ball_gather", "dct_map_max", "dct_map_read_at",
              "dct_read_cnt", "dct_map_hist", "dct_set_mu", "dct_get_mu",
              "dct_ball_gather_mu", "dct_loss_owner", "dct_set_tbl2",
              "dct_use_fmt", "dct_stats", "dct_sweep_plan",
              "dct_sweep_plan2", "dct_plan_probe", "dct_autotune",
              "dct_sweep_profile", "dct_sweep_force", "dct_map_read",
              "dct_sweep_times", "dct_transform_layers",
              "dct_layers_read_at", "dct_region_features", "dct_region_propose",
              "dct_set_fold"):
        getattr(lib, f).restype = ctypes.c_int
    for f in ("dct_count_eq", "dct_count_deficit", "dct_map_extract",
              "dct_map_extract64", "dct_host_bytes"):
        getattr(lib, f).restype = _LL
    lib.dct_read_cnt.argtypes = [_LL, _LL, ctypes.POINTER(ctypes.c_uint16)]
    lib.dct_map_read.argtypes = [_LL, _LL, _I32P]
    lib.dct_region_propose.argtypes = [_LLP, ctypes.c_int, _I32P, ctypes.c_int,
                                       ctypes.c_int, ctypes.c_uint64, _I32P,
                                       ctypes.POINTER(ctypes.c_float)]
    lib.dct_init.restype = ctypes.c_int
    lib.dct_init.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _U64P,
                             _LL, ctypes.c_int, _LLP]
    lib.dct_init2.restype = ctypes.c_int
    lib.dct_init2.argtypes = lib.dct_init.argtypes + [ctypes.c_int] * 3
    for f in ("dct_init3", "dct_init3w", "dct_init3x"):
        getattr(lib, f).restype = ctypes.c_int
        getattr(lib, f).argtypes = [ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                    _U64P, _LL, ctypes.c_int, _LLP,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.dct_ball_update_delta.restype = ctypes.c_int
    lib.dct_ball_update_delta.argtypes = [_LLP, ctypes.c_int, ctypes.c_int,
                                          _LLP]
    lib.dct_swap_eval.restype = ctypes.c_int
    lib.dct_swap_eval.argtypes = [_LLP, ctypes.c_int, _LLP, _LLP, _I32P,
                                  _I32P]
    return lib


def load_plugin(path=None):
    path = os.path.abspath(path or default_plugin_path())
    if path not in _libs:
        _libs[path] = _bind(path)
    return _libs[path]


def locality_order(pats, axes, bits=8):
    """Sort a v1 table by unwrapped offset so neighbouring table entries
    touch neighbouring words (reference gpu.py:106-129).  Walks are
    order-free, so results are unchanged; entry 0 stays the centre.
    ``bits`` is the field width: 8 (v1), 10 (wide), 16 (long-axis, whose
    fields are pos << 10 | delta).
    ``COLDCORE_BALL_ORDER=native`` keeps the builder's order."""
    if os.environ.get("COLDCORE_BALL_ORDER", "locality") == "native":
        return pats
    pats = np.ascontiguousarray(pats, dtype=np.uint64)
    stride = np.cumprod([1] + list(axes[:-1])).astype(np.int64)
    key = np.zeros(len(pats), dtype=np.int64)
    ps = 10 if bits == 16 else 4
    for k in range(64 // bits):
        byte = ((pats >> np.uint64(bits * k))
                & np.uint64((1 << bits) - 1)).astype(np.int64)
        pos, delta = byte >> ps, byte & ((1 << ps) - 1)
        key += np.where(byte != 0,
                        delta * stride[np.minimum(pos, len(axes) - 1)], 0)
    return np.ascontiguousarray(pats[np.argsort(key, kind="stable")])


def _ptr(a, t):
    return a.ctypes.data_as(t)


#: candidates and drops one dct_swap_eval call takes (uint16 staging)
_SWAP_LIMIT = (1 << 16) - 1

#: dct_stats slots, in slot order; the names are the reference binding's
#: (interface: callers read them)
STAT_KEYS = ("gather_set_side", "gather_walk", "set_scans", "set_overflows",
             "swap_eval_fast", "swap_eval_fallback", "count_eq0_listed",
             "gather_set_tiled")


def _swap_batches(sizes, limit):
    """[lo, hi) runs of consecutive drops whose candidate total and drop
    count stay within ``limit`` (each drop's own count is within it)."""
    k = len(sizes)
    lo = 0
    while lo < k:
        run = np.cumsum(sizes[lo:lo + limit])
        hi = lo + max(1, int(np.searchsorted(run, limit, side="right")))
        yield lo, hi
        lo = hi


class GpuBackend:
    """One problem cell resident in MI355X HBM (reference gpu.py:207-537).

    problem: "hamming" | "torus_linf" | "lee" | "grid_linf"; mu: covering
    multiplicity.  layers_mode / cnt_mode are accepted for compatibility;
    every plane lives in HBM.
    """

    def __init__(self, axes, R, problem="hamming", plugin_path=None,
                 extract_cap=1 << 22, layers_mode=None, cnt_mode=None,
                 use_owner=False, mu=1):
        self.problem = problem
        self.R = R
        self.mu = mu = int(mu)
        if not 1 <= mu <= 65535:
            raise ValueError("mu must be in 1..65535")
        self.space = ProductSpace(axes)
        self.extract_cap = extract_cap
        self.indep_dist = 2 * R if problem in ("hamming", "lee") else None
        self.supports_hist = True
        self._lib = lib = load_plugin(plugin_path)
        self._ext_bufs = None
        # v1 tables address 16 axes; larger Hamming / Lee spaces use the
        # additive wide format (10-bit fields) through dct_init3w, L-inf
        # cells with axes longer than 16 the long-axis format (16-bit
        # fields) through dct_init3x
        self.wide = self.space.n > 16
        self.longax = max(self.space.axes) > 16
        if self.longax:
            pats = locality_order(
                support_patterns_long(self.space.axes, R, problem),
                self.space.axes, bits=16)
            init = lib.dct_init3x
        elif self.wide:
            pats = locality_order(
                support_patterns_wide(self.space.axes, R, problem),
                self.space.axes, bits=10)
            init = lib.dct_init3w
        else:
            pats = locality_order(
                support_patterns(self.space.axes, R, problem), self.space.axes)
            init = lib.dct_init3
        self._pats = pats
        self.layers_mode = layers_mode or 0
        self.cnt_mode = cnt_mode or 0
        by = _LL(0)
        ax = (ctypes.c_int * self.space.n)(*self.space.axes)
        rc = init(PROBLEM_IDS[problem], self.space.n, ax, R,
                           _ptr(pats, _U64P), len(pats), extract_cap,
                           ctypes.byref(by), self.layers_mode, self.cnt_mode,
                           1 if use_owner else 0)
        _check("dct_init3", rc)
        _check("dct_set_mu", lib.dct_set_mu(mu))
        self.gpu_bytes = by.value
        self.host_bytes = int(lib.dct_host_bytes())
        self._closed = False

    def close(self):
        if not self._closed:
            self._lib.dct_free()
            self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def word_index(self, words):
        return self.space.word_index(words)

    def index_word(self, idx):
        return self.space.index_word(idx)

    def _set_code(self, idx):
        idx = np.ascontiguousarray(np.atleast_1d(idx), dtype=np.int64)
        _check("dct_set_code", self._lib.dct_set_code(_ptr(idx, _LLP), len(idx)))

    def recount(self, idx):
        self._set_code(idx)
        _check("dct_transform(0, 2)", self._lib.dct_transform(0, 2))

    def gain_map(self):
        _check("dct_transform(1, 1)", self._lib.dct_transform(1, 1))

    def loss_map(self):
        _check("dct_transform(2, 1)", self._lib.dct_transform(2, 1))

    def ball_update(self, idx, delta):
        idx = np.ascontiguousarray(np.atleast_1d(idx), dtype=np.int64)
        _check("dct_ball_update", self._lib.dct_ball_update(
            _ptr(idx, _LLP), len(idx), int(delta)))

    def ball_gather(self, idx, target):
        """Exact per-candidate support count; target is a coverage ROLE
        (0 deficient = placement gain, 1 critical = removal loss), else a
        literal cnt value (reference gpu.py:323-348)."""
        idx = np.ascontiguousarray(np.atleast_1d(idx), dtype=np.int64)
        out = np.empty(len(idx), dtype=np.int32)
        if self.mu == 1 or target not in (0, 1):
            _check("dct_ball_gather", self._lib.dct_ball_gather(
                _ptr(idx, _LLP), len(idx), int(target), _ptr(out, _I32P)))
        else:
            _check("dct_ball_gather_mu", self._lib.dct_ball_gather_mu(
                _ptr(idx, _LLP), len(idx), int(target), _ptr(out, _I32P)))
        return out

    def ball_update_delta(self, idx, delta):
        idx = np.ascontiguousarray(np.atleast_1d(idx), dtype=np.int64)
        dd = _LL(0)
        _check("dct_ball_update_delta", self._lib.dct_ball_update_delta(
            _ptr(idx, _LLP), len(idx), int(delta), ctypes.byref(dd)))
        return int(dd.value)

    def swap_eval(self, drops, cands):
        """Exact costs of many swaps without changing state (reference
        gpu.py:368-440): loss[i] of removing drops[i] and the gain of every
        candidate once drops[i] is gone.  ``cands`` is one array shared by
        every drop (gains come back as a (k, m) array) or one array per drop
        (a list of arrays).  The ABI call takes at most 65,535 drops and
        65,535 candidates, so consecutive drops are packed into calls up to
        that limit."""
        drops = np.ascontiguousarray(np.atleast_1d(drops), dtype=np.int64)
        k = len(drops)
        shared = not isinstance(cands, (list, tuple))
        if shared:
            per = [np.ascontiguousarray(np.atleast_1d(cands), dtype=np.int64)] * k
        elif len(cands) != k:
            raise ValueError(f"{len(cands)} candidate arrays for {k} drops")
        else:
            per = [np.ascontiguousarray(np.atleast_1d(c), dtype=np.int64)
                   for c in cands]
        limit = _SWAP_LIMIT
        sizes = np.fromiter((len(c) for c in per), dtype=np.int64, count=k)
        if k and int(sizes.max()) > limit:
            raise ValueError(f"a drop has {int(sizes.max())} candidates; "
                             f"one call takes at most {limit}")
        loss = np.empty(k, dtype=np.int32)
        gains = []
        for lo, hi in _swap_batches(sizes, limit):
            bounds = np.concatenate(([0], np.cumsum(sizes[lo:hi]))).astype(np.int64)
            flat = (np.concatenate(per[lo:hi]) if bounds[-1]
                    else np.zeros(1, dtype=np.int64))
            got = np.empty(max(int(bounds[-1]), 1), dtype=np.int32)
            part = np.ascontiguousarray(drops[lo:hi])
            _check("dct_swap_eval", self._lib.dct_swap_eval(
                _ptr(part, _LLP), hi - lo, _ptr(flat, _LLP), _ptr(bounds, _LLP),
                _ptr(got, _I32P), _ptr(loss[lo:hi], _I32P)))
            gains.extend(np.split(got[:bounds[-1]], bounds[1:-1]))
        if shared:
            m = len(per[0]) if k else len(np.atleast_1d(cands))
            return loss, (np.vstack(gains) if k else np.zeros((0, m), dtype=np.int32))
        return loss, gains

    def gpu_stats(self):
        """The engine's event counters (dct_stats slots 0..7) under the
        reference binding's key names (reference gpu.py:451-456)."""
        slots = (_LL * len(STAT_KEYS))()
        self._lib.dct_stats(slots, len(STAT_KEYS))
        return dict(zip(STAT_KEYS, (int(v) for v in slots)))

    def count_eq(self, target):
        r = int(self._lib.dct_count_eq(int(target)))
        if r < 0:
            _check("dct_count_eq", r)
        return r

    def uncovered(self):
        """Total deficit sum_x max(0, mu - cnt(x)); 0 iff feasible."""
        r = int(self._lib.dct_count_eq(0) if self.mu == 1
                else self._lib.dct_count_deficit())
        if r < 0:
            _check("dct_count_eq/deficit", r)
        return r

    def map_max(self):
        v = ctypes.c_int32(0)
        _check("dct_map_max", self._lib.dct_map_max(ctypes.byref(v)))
        return int(v.value)

    def map_hist(self, nbins, vmax):
        out = np.zeros(nbins, dtype=np.int32)
        _check("dct_map_hist", self._lib.dct_map_hist(
            int(nbins), int(vmax), _ptr(out, _I32P)))
        return out

    def map_extract(self, thr, cap=None):
        """(idx, val, found) of map >= thr, ascending index order; found may
        exceed cap (only cap entries are returned)."""
        cap = min(cap or self.extract_cap, self.extract_cap)
        wide = self.space.size > 2 ** 32 - 2
        oi = np.empty(max(cap, 1), dtype=np.int64 if wide else np.uint32)
        ov = np.empty(max(cap, 1), dtype=np.int32)
        if wide:
            found = self._lib.dct_map_extract64(
                int(thr), _ptr(oi, ctypes.POINTER(ctypes.c_int64)),
                _ptr(ov, _I32P), cap)
        else:
            found = self._lib.dct_map_extract(
                int(thr), _ptr(oi, ctypes.POINTER(ctypes.c_uint32)),
                _ptr(ov, _I32P), cap)
        if found < 0:
            _check("dct_map_extract", int(found))
        m = min(int(found), cap)
        return oi[:m].astype(np.int64), ov[:m], int(found)

    def map_read_at(self, idx):
        idx = np.ascontiguousarray(np.atleast_1d(idx), dtype=np.int64)
        out = np.empty(len(idx), dtype=np.int32)
        _check("dct_map_read_at", self._lib.dct_map_read_at(
            _ptr(idx, _LLP), len(idx), _ptr(out, _I32P)))
        return out

    def map_all(self, off=0, length=None):
        """Map-plane entries [off, off+length) (default: the whole plane)."""
        length = self.space.size - off if length is None else length
        out = np.empty(length, dtype=np.int32)
        _check("dct_map_read", self._lib.dct_map_read(int(off), int(length),
                                                  _ptr(out, _I32P)))
        return out

    def read_cnt(self, off=0, length=None):
        length = self.space.size - off if length is None else length
        out = np.empty(length, dtype=np.uint16)
        _check("dct_read_cnt", self._lib.dct_read_cnt(
            int(off), int(length), _ptr(out, ctypes.POINTER(ctypes.c_uint16))))
        return out

    def layers_at(self, init_mode, idx, code=None):
        """(k, R+1) exact distance layers at idx: entry [i, d] is the number
        of seed words at distance exactly d from idx[i]; init_mode 0 seeds
        with ``code`` (default: nothing), 1 the deficient words, 2 the
        critical words.  cnt is unchanged."""
        idx = np.ascontiguousarray(np.atleast_1d(idx), dtype=np.int64)
        if init_mode == 0:
            self._set_code(np.zeros(0, dtype=np.int64) if code is None else code)
        _check("dct_transform_layers", self._lib.dct_transform_layers(int(init_mode)))
        L = self.R + 1
        out = np.empty((len(idx), L), dtype=np.int32)
        _check("dct_layers_read_at", self._lib.dct_layers_read_at(
            _ptr(idx, _LLP), len(idx), _ptr(out, _I32P)))
        return out

    def region_features(self, code, regions):
        """(regions, 5 (R+1)) float32 features of ruin regions (rows of word
        indices), computed on the device from the three exact layer fields
        of ``code`` (dct_region_features)."""
        code = np.ascontiguousarray(np.atleast_1d(code), dtype=np.int64)
        W = np.ascontiguousarray(regions, dtype=np.int64)
        nreg, k = W.shape
        out = np.empty((nreg, 5 * (self.R + 1)), dtype=np.float32)
        _check("dct_region_features", self._lib.dct_region_features(
            _ptr(code, _LLP), len(code), _ptr(W, _LLP), nreg, k,
            out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return out

    def region_propose(self, code, anchors, k, seed):
        """((na, k) region positions into ``code``, (na, 5 (R+1)) features):
        the k code words nearest each anchor position and their feature
        windows, picked and computed on the device (dct_region_propose)."""
        code = np.ascontiguousarray(np.atleast_1d(code), dtype=np.int64)
        anc = np.ascontiguousarray(np.atleast_1d(anchors), dtype=np.int32)
        regions = np.empty((len(anc), k), dtype=np.int32)
        feats = np.empty((len(anc), 5 * (self.R + 1)), dtype=np.float32)
        _check("dct_region_propose", self._lib.dct_region_propose(
            _ptr(code, _LLP), len(code), _ptr(anc, _I32P), len(anc), int(k),
            int(seed) & (2 ** 64 - 1), _ptr(regions, _I32P),
            feats.ctypes.data_as(ctypes.POINTER(ctypes.c_float))))
        return regions, feats

    def sweep_plan(self):
        buf = (ctypes.c_int * (8 * 17))()
        ns = self._lib.dct_sweep_plan(buf, 16)
        if ns < 0:
            _check("dct_sweep_plan", ns)
        keys = ("first_axis", "axes", "P", "TL", "layers_in", "layers_out",
                "round_q", "lds_bytes")
        out = [dict(zip(keys, buf[8 * i:8 * i + 8])) for i in range(ns)]
        for sw in out:
            sw["layer_bytes_field"] = buf[8 * ns]
            sw["layer_bytes_recount"] = buf[8 * ns + 1]
        return out

    def sweep_plan2(self, kind=0):
        """Schedule a transform runs now: kind 0 gain / loss, 1 recount of
        the pending code.  Per sweep: axes, tile, register round and field
        format (W = 32: unpacked planes), algorithmic bytes per word."""
        buf = (ctypes.c_int * (13 * 41))()
        ns = self._lib.dct_sweep_plan2(kind, buf, 40)
        if ns < 0:
            _check("dct_sweep_plan2", ns)
        keys = ("first_axis", "axes", "P", "TL", "RR", "W", "ND", "bytes_in",
                "bytes_out", "lds_bytes", "threads", "rounds", "kernel")
        return [dict(zip(keys, buf[13 * i:13 * i + 13])) for i in range(ns)]

    def autotune(self, reps=3):
        """Time the candidate transform schedules on this cell and keep the
        fastest (overwrites the map plane).  Returns (schedules timed,
        microseconds per transform of the winner)."""
        us = ctypes.c_float(0.0)
        n = self._lib.dct_autotune(int(reps), ctypes.byref(us))
        if n < 0:
            _check("dct_autotune", n)
        return n, float(us.value)

    def sweep_force(self, cuts):
        """Pin the packed schedule's sweep boundaries (ascending axis
        counts; [] clears) and return the gain / loss sweep count; raises
        when no packed schedule has these boundaries."""
        c = np.ascontiguousarray(list(cuts), dtype=np.int32)
        ns = self._lib.dct_sweep_force(
            c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(c))
        if ns < 0:
            raise RuntimeError(f"dct_sweep_force{tuple(c)} rc={ns}")
        return ns

    def sweep_profile(self, on=True):
        _check("dct_sweep_profile", self._lib.dct_sweep_profile(1 if on else 0))

    def set_fold(self, on=True):
        """Fold the map maximum / cnt deficit into later transforms (read
        by map_max / count_eq(0) / uncovered without a field scan); returns
        the previous setting."""
        r = int(self._lib.dct_set_fold(1 if on else 0))
        if r < 0:
            _check("dct_set_fold", r)
        return bool(r)

    def sweep_times(self):
        buf = (ctypes.c_float * 16)()
        ns = self._lib.dct_sweep_times(buf, 16)
        if ns < 0:
            _check("dct_sweep_times", ns)
        return [float(buf[i]) for i in range(ns)]


PLAN_KEYS = ("first_axis", "axes", "P", "TL", "RR", "W", "ND", "bytes_in",
             "bytes_out", "lds_bytes", "threads", "rounds", "kernel")


def plan_probe(axes, R, kind=0, s0=1, M=1, path=None, cuts=None):
    """The packed schedule the engine would plan for a homogeneous Hamming
    cell, computed on the host (no device, no live cell).  [] when the cell
    runs the unpacked kernels.  ``cuts`` pins the sweep boundaries (as
    GpuBackend.sweep_force)."""
    lib = load_plugin(path)
    ax = np.ascontiguousarray(axes, dtype=np.int32)
    buf = (ctypes.c_int * (13 * 41))()
    c = np.ascontiguousarray(list(cuts or ()), dtype=np.int32)
    if lib.dct_sweep_force(c.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                           len(c)) < 0:
        raise RuntimeError(f"dct_sweep_force{tuple(c)}: bad cuts")
    try:
        ns = lib.dct_plan_probe(len(ax), ax.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                                int(R), int(kind), int(s0), int(M), buf, 40)
    finally:
        lib.dct_sweep_force(None, 0)
    if ns < 0:
        raise RuntimeError(f"plan_probe rc={ns}")
    return [dict(zip(PLAN_KEYS, buf[13 * i:13 * i + 13])) for i in range(ns)]


class GpuCoveringBackend(GpuBackend):
    """Homogeneous covering-code cell K_q(n,R) (reference gpu.py:540-545)."""

    def __init__(self, q, n, R, **kw):
        super().__init__([q] * n, R, problem="hamming", **kw)
        self.q, self.n = q, n
