"""Optional and legacy ABI symbols on the device: the ox table format
(dct_set_tbl2 / dct_use_fmt) and the homogeneous legacy inits (dct_init,
dct_init2), called through ctypes as the reference's own tests call them.

* ox (reference src/plugins/covering/dct_hc.cu:430-450): after
  dct_set_tbl2 + dct_use_fmt(1) every walk (update, gather of every role)
  equals the v1 walk and the oracle; the grid refuses use_fmt(1) with the
  table itself accepted (reference tests/test_gpu_parity_grid.py:129-146).
* dct_init / dct_init2 (reference dct_hc.cu:376-391) build the same cell as
  dct_init3: recount, gain and loss equal the oracle.
"""
import ctypes

import numpy as np
import pytest

from oracle.dp import NumpyDP
from coldcore_amd.patterns import support_patterns, support_patterns_ox

pytestmark = pytest.mark.gpu

OX_CELLS = [("hamming", [4] * 5, 2), ("hamming", [2, 3, 4, 5, 3], 2),
            ("lee", [5, 6, 7], 2), ("torus_linf", [9, 9], 1),
            ("torus_linf", [5, 6, 7], 1)]


def _install_ox(g, axes, R, problem):
    mask, offs = support_patterns_ox(axes, R, problem)
    return g._lib.dct_set_tbl2(
        mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
        offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
        ctypes.c_longlong(len(mask)))


@pytest.mark.parametrize("problem,axes,R", OX_CELLS,
                         ids=[f"{p}-{a}-R{r}" for p, a, r in OX_CELLS])
def test_ox_walks_equal_v1_and_oracle(gpu_cell, problem, axes, R):
    g = gpu_cell(axes, R, problem=problem)
    c = NumpyDP(axes, R, problem=problem)
    size = c.space.size
    rng = np.random.default_rng(3)
    code = rng.integers(0, size, max(2, size // 25)).astype(np.int64)
    g.recount(code)
    c.recount(code)
    assert _install_ox(g, axes, R, problem) == 0
    probe = rng.integers(0, size, 300).astype(np.int64)
    moves = rng.integers(0, size, 40).astype(np.int64)
    for fmt in (1, 0, 1):
        assert g._lib.dct_use_fmt(ctypes.c_int(fmt)) == 0
        for role in (0, 1, 2):
            assert (g.ball_gather(probe, role) == c.ball_gather(probe, role)).all(), \
                (fmt, role)
    # updates walked with the ox table, then undone with the v1 table
    assert g._lib.dct_use_fmt(ctypes.c_int(1)) == 0
    g.ball_update(moves, +1)
    c.ball_update(moves, +1)
    assert (g.read_cnt() == c.read_cnt()).all()
    for role in (0, 1):
        assert (g.ball_gather(probe, role) == c.ball_gather(probe, role)).all()
    assert g._lib.dct_use_fmt(ctypes.c_int(0)) == 0
    g.ball_update(moves, -1)
    c.ball_update(moves, -1)
    assert (g.read_cnt() == c.read_cnt()).all()
    # the fields do not depend on the table format
    g.gain_map()
    c.gain_map()
    assert (g.map_all() == c.map).all()


def test_ox_rejects_a_table_of_the_wrong_length(gpu_cell):
    g = gpu_cell([4] * 5, 2)
    mask, offs = support_patterns_ox([4] * 5, 2)
    rc = g._lib.dct_set_tbl2(
        mask.ctypes.data_as(ctypes.POINTER(ctypes.c_uint16)),
        offs.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64)),
        ctypes.c_longlong(len(mask) - 1))
    assert rc == -1
    assert g._lib.dct_use_fmt(ctypes.c_int(1)) == -1   # no ox table installed


def test_grid_rejects_the_circulant_ox_format(gpu_cell):
    g = gpu_cell([8, 8], 1, problem="grid_linf")
    assert _install_ox(g, [8, 8], 1, "grid_linf") == 0   # the table is fine
    assert g._lib.dct_use_fmt(ctypes.c_int(1)) != 0      # its walk is not
    assert g._lib.dct_use_fmt(ctypes.c_int(0)) == 0


@pytest.mark.parametrize("legacy", ["dct_init", "dct_init2"])
@pytest.mark.parametrize("q,n,R", [(3, 8, 2), (2, 12, 3), (5, 6, 3)])
def test_legacy_inits_build_the_dct_init3_cell(legacy, q, n, R):
    from coldcore_amd.gpu import load_plugin, locality_order
    lib = load_plugin()
    axes = [q] * n
    pats = locality_order(support_patterns(axes, R), axes)
    by = ctypes.c_longlong(0)
    pp = pats.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))
    if legacy == "dct_init":
        rc = lib.dct_init(q, n, R, pp, len(pats), 1 << 20, ctypes.byref(by))
    else:
        rc = lib.dct_init2(q, n, R, pp, len(pats), 1 << 20, ctypes.byref(by),
                           0, 0, 0)
    assert rc == 0 and by.value > 0
    try:
        c = NumpyDP(axes, R)
        size = c.space.size
        code = np.random.default_rng(q * 100 + n).integers(
            0, size, max(2, size // 40)).astype(np.int64)
        c.recount(code)
        assert lib.dct_set_code(code.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                len(code)) == 0
        assert lib.dct_transform(0, 2) == 0
        cnt = np.empty(size, dtype=np.uint16)
        assert lib.dct_read_cnt(0, size, cnt.ctypes.data_as(
            ctypes.POINTER(ctypes.c_uint16))) == 0
        assert (cnt == c.read_cnt()).all()
        out = np.empty(size, dtype=np.int32)
        for init, which in ((1, "gain_map"), (2, "loss_map")):
            assert lib.dct_transform(init, 1) == 0
            assert lib.dct_map_read(0, size, out.ctypes.data_as(
                ctypes.POINTER(ctypes.c_int32))) == 0
            getattr(c, which)()
            assert (out == c.map).all(), which
        probe = np.arange(0, size, 7, dtype=np.int64)
        g0 = np.empty(len(probe), dtype=np.int32)
        assert lib.dct_ball_gather(probe.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong)),
                                   len(probe), 0,
                                   g0.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))) == 0
        assert (g0 == c.ball_gather(probe, 0)).all()
    finally:
        lib.dct_free()
