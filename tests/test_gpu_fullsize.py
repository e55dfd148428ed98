"""Full-size parity on the benchmark cells, by size-independent checks.

* K_2(24,3) (2^24 words): cnt and the gain field against the integer
  Walsh-Hadamard convolution (an oracle that shares nothing with the axis
  DP), plus the translation-invariance sums sum cnt = M |B| and
  sum gain = |B| #{cnt < mu}.
* K_5(10,3) and the mixed-radix cell [4,5,6,7,8,4,5,6,7] R=3 against the
  numpy axis-DP oracle (a few seconds each on one core).
* walks at full size: gain_map == ball_gather(., 0) on sampled words and
  update-then-undo is the identity.
"""
import numpy as np
import pytest

from oracle.dp import NumpyDP
from oracle.walsh import binary_field
from coldcore_amd.patterns import ball_size

pytestmark = pytest.mark.gpu


def _code(size, p, seed=0):
    return np.flatnonzero(np.random.default_rng(seed).random(size) < p).astype(np.int64)


def test_k2_24_3_against_walsh_hadamard(gpu_cell):
    n, R = 24, 3
    size = 1 << n
    code = _code(size, 1e-4)
    g = gpu_cell([2] * n, R)
    g.recount(code)
    cnt = g.read_cnt()
    ind = np.zeros(size, dtype=np.int64)
    np.add.at(ind, code, 1)
    want = binary_field(ind, n, R)
    assert (cnt.astype(np.int64) == want).all()
    B = ball_size(2, n, R)
    assert int(cnt.sum(dtype=np.int64)) == len(code) * B
    g.gain_map()
    gain = g.map_all()
    want_gain = binary_field((want < 1).astype(np.int64), n, R)
    assert (gain.astype(np.int64) == want_gain).all()
    assert int(gain.sum(dtype=np.int64)) == B * int((want < 1).sum())
    g.loss_map()
    loss = g.map_all()
    want_loss = binary_field((want == 1).astype(np.int64), n, R)
    assert (loss.astype(np.int64) == want_loss).all()
    # walks against the field at sampled words
    rng = np.random.default_rng(2)
    probe = rng.integers(0, size, 4096).astype(np.int64)
    g.gain_map()
    assert (g.ball_gather(probe, 0) == g.map_read_at(probe)).all()
    w = probe[:64]
    g.ball_update(w, +1)
    g.ball_update(w, -1)
    assert (g.read_cnt() == cnt).all()


@pytest.mark.parametrize("axes,R", [([5] * 10, 3),
                                    ([4, 5, 6, 7, 8, 4, 5, 6, 7], 3)],
                         ids=["K5(10,3)", "mixed9-R3"])
def test_large_cells_against_numpy_oracle(gpu_cell, axes, R):
    c = NumpyDP(axes, R)
    size = c.space.size
    code = _code(size, 1e-4)
    g = gpu_cell(axes, R)
    g.recount(code)
    c.recount(code)
    assert (g.read_cnt() == c.read_cnt()).all()
    g.gain_map()
    c.gain_map()
    assert (g.map_all() == c.map).all()
    g.loss_map()
    c.loss_map()
    assert (g.map_all() == c.map).all()
    mx = c.map_max()
    assert g.map_max() == mx
    gi, gv, gf = g.map_extract(max(1, mx // 2), cap=1 << 18)
    ci, cv, cf = c.map_extract(max(1, mx // 2), cap=1 << 18)
    assert gf == cf and (gi == ci).all() and (gv == cv).all()
    assert (g.map_hist(1024, mx) == c.map_hist(1024, mx)).all()
