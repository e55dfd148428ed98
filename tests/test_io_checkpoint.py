"""Solution files and checkpoint / resume (reference src/coldcore/io.py:1-34,
campaign.py:106-135; its tests/test_campaign.py:51-64 and 95-103 check the
same behaviour: the artefact is a real cover of size M, and a rerun resumes
from it and never ends larger).

The file format has no fixture in the reference, so its bytes are pinned
by hand-written examples; a cover is checked by the oracle's recount.
"""
import os

import numpy as np
import pytest

from coldcore_amd.checkpoint import run_cell
from coldcore_amd.io import read_solution, write_solution_atomic
from oracle.dp import NumpyDP


# ------------------------------------------------------------------ format
def test_rows_are_space_separated_digits(tmp_path):
    p = tmp_path / "sol.txt"
    write_solution_atomic(str(p), np.array([[0, 1, 2], [2, 2, 0], [10, 0, 1]]))
    assert p.read_bytes() == b"0 1 2\n2 2 0\n10 0 1\n"
    assert not os.path.exists(str(p) + ".tmp")
    got = read_solution(str(p), n=3)
    assert got.dtype == np.uint8
    assert got.tolist() == [[0, 1, 2], [2, 2, 0], [10, 0, 1]]


def test_comments_and_blank_lines_are_skipped(tmp_path):
    p = tmp_path / "sol.txt"
    p.write_text("# K_3(3,1) cover\n\n0 0 0\n  1 1 1  \n# end\n2 2 2\n")
    assert read_solution(str(p), q=3, n=3).tolist() == [[0, 0, 0], [1, 1, 1], [2, 2, 2]]


def test_validation_errors_name_the_file(tmp_path):
    p = tmp_path / "sol.txt"
    p.write_text("0 1 2\n2 2 3\n")
    with pytest.raises(ValueError, match="want n=4"):
        read_solution(str(p), n=4)
    with pytest.raises(ValueError, match="out of range for q=3"):
        read_solution(str(p), q=3)
    assert read_solution(str(p), q=4).shape == (2, 3)
    # mixed radix: axis 2 has size 3, so the digit 3 is out of range
    with pytest.raises(ValueError, match="axis 2 of size 3"):
        read_solution(str(p), axes=[3, 3, 3])
    assert read_solution(str(p), axes=[3, 3, 4]).shape == (2, 3)
    p.write_text("0 1\n2 2 3\n")
    with pytest.raises(ValueError, match="different widths"):
        read_solution(str(p))


def test_empty_file_has_zero_rows(tmp_path):
    p = tmp_path / "sol.txt"
    write_solution_atomic(str(p), np.zeros((0, 5), dtype=np.uint8))
    assert p.read_bytes() == b""
    assert read_solution(str(p), n=5).shape == (0, 5)


def test_a_failed_write_leaves_the_old_file(tmp_path, monkeypatch):
    p = tmp_path / "sol.txt"
    write_solution_atomic(str(p), np.array([[1, 2]]))
    with pytest.raises(ValueError):            # rejected before any file op
        write_solution_atomic(str(p), np.array([[1, -2]]))
    with pytest.raises(ValueError):
        write_solution_atomic(str(p), np.arange(3))   # not an (M, n) array

    def boom(src, dst):
        raise OSError("disk gone")
    monkeypatch.setattr(os, "replace", boom)
    with pytest.raises(OSError):
        write_solution_atomic(str(p), np.array([[0, 0], [1, 1]]))
    monkeypatch.undo()
    assert p.read_bytes() == b"1 2\n"
    assert not os.path.exists(str(p) + ".tmp")


# ------------------------------------------------------------------ resume
def _covers(axes, R, digits):
    o = NumpyDP(axes, R)
    o.recount(o.word_index(np.asarray(digits, dtype=np.int64)))
    return int((o.cnt == 0).sum()) == 0


def test_run_cell_checkpoints_a_real_cover(tmp_path):
    axes, R = [3] * 4, 1
    sol, logp = str(tmp_path / "k3_4_1.txt"), str(tmp_path / "k3_4_1.log")
    rec = run_cell(NumpyDP(axes, R), sol, seed=1, rounds=200, log_path=logp)
    assert rec["uncov"] == 0 and rec["resumed_from"] is None
    code = read_solution(sol, n=4)
    assert len(code) == rec["M"] and _covers(axes, R, code)
    sizes = [m for _, m in rec["solutions"]]
    assert sizes == sorted(sizes, reverse=True) and sizes[-1] == rec["M"]
    assert rec["M"] == 9                       # K_3(4,1) = 9 (a perfect code)
    assert "SOLVED M=9" in open(logp).read()


def test_rerun_resumes_and_never_ends_larger(tmp_path):
    axes, R = [2] * 6, 1
    sol, logp = str(tmp_path / "k2_6_1.txt"), str(tmp_path / "k2_6_1.log")
    first = run_cell(NumpyDP(axes, R), sol, seed=3, rounds=20, log_path=logp)
    again = run_cell(NumpyDP(axes, R), sol, seed=4, rounds=60, log_path=logp)
    assert again["resumed_from"] == sol and "resumed from" in open(logp).read()
    assert again["uncov"] == 0 and again["M"] <= first["M"]
    assert _covers(axes, R, read_solution(sol, n=6))
    fresh_log = str(tmp_path / "fresh.log")
    fresh = run_cell(NumpyDP(axes, R), sol, seed=3, rounds=20, resume=False,
                     log_path=fresh_log)
    assert fresh["resumed_from"] is None
    assert "resumed from" not in open(fresh_log).read()
    assert fresh["M"] == first["M"]            # clock-free: the same run again


def test_unbounded_run_is_refused(tmp_path):
    with pytest.raises(ValueError):
        run_cell(NumpyDP([3] * 3, 1), str(tmp_path / "x.txt"), rounds=None)


def test_seed_file_start_and_mixed_radix(tmp_path):
    axes, R = [2, 3, 4], 1
    o = NumpyDP(axes, R)
    every = o.index_word(np.arange(o.space.size, dtype=np.int64))
    seed = str(tmp_path / "all.txt")
    write_solution_atomic(seed, every)         # the whole space: a cover
    sol = str(tmp_path / "mixed.txt")
    rec = run_cell(NumpyDP(axes, R), sol, seed=2, rounds=100, seed_file=seed)
    assert rec["resumed_from"] == seed and rec["uncov"] == 0
    code = read_solution(sol, axes=axes)
    assert len(code) == rec["M"] < o.space.size and _covers(axes, R, code)


@pytest.mark.gpu
def test_gpu_run_writes_the_oracle_files(gpu_cell, tmp_path):
    """The same seeded, clock-free run through the engine's C ABI writes
    the same checkpoints, byte for byte, as on the numpy oracle."""
    for i, (axes, R) in enumerate((([2] * 6, 1), ([3] * 4, 1), ([3] * 5, 1))):
        a, b = str(tmp_path / f"gpu{i}.txt"), str(tmp_path / f"cpu{i}.txt")
        g = gpu_cell(axes, R)
        rg = run_cell(g, a, seed=5, rounds=120)
        g.close()
        rc = run_cell(NumpyDP(axes, R), b, seed=5, rounds=120)
        assert (rg["M"], rg["uncov"], rg["rounds"]) == (rc["M"], rc["uncov"], rc["rounds"])
        assert [m for _, m in rg["solutions"]] == [m for _, m in rc["solutions"]]
        assert open(a, "rb").read() == open(b, "rb").read()
