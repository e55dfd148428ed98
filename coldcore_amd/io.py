"""Solution files: one code word per line as its digits, space-separated.

The on-disk format of a covering code (reference src/coldcore/io.py:1-34):
row i holds the n digits of code word i (axis 0 first, the order
``ProductSpace.index_word`` returns), separated by single spaces; blank
lines and lines starting with ``#`` are skipped on reading.  Files written
by the reference read here and the other way round.

Writes are crash-safe: the rows go to ``path + ".tmp"``, which is flushed
and fsync'ed before ``os.replace`` moves it over ``path`` in one step, so
a reader (or a resumed run) sees either the previous complete file or the
new complete file, never a torn one.
"""
import os

import numpy as np


def _rows_text(words):
    arr = np.asarray(words)
    if arr.ndim == 1 and arr.size == 0:
        arr = arr.reshape(0, 0)
    if arr.ndim != 2:
        raise ValueError(f"solution words must be an (M, n) digit array, got {arr.shape}")
    if arr.size and (arr.min() < 0):
        raise ValueError("solution digits must be non-negative")
    return "".join(" ".join(str(int(x)) for x in row) + "\n" for row in arr)


def write_solution_atomic(path, words):
    """Write the (M, n) digit array ``words`` to ``path`` through a
    temporary file and an atomic rename (reference io.py:11-17).  The text
    is built before the temporary file is opened, so a bad array leaves
    both ``path`` and the directory untouched."""
    text = _rows_text(words)
    tmp = os.fspath(path) + ".tmp"
    try:
        with open(tmp, "w") as f:
            f.write(text)
            f.flush()
            os.fsync(f.fileno())
        os.replace(tmp, path)
    except BaseException:
        if os.path.exists(tmp):
            os.remove(tmp)
        raise


def read_solution(path, q=None, n=None, axes=None):
    """Read a digit-row file into an (M, n) uint8 array (reference
    io.py:20-34).  ``n`` checks the row width, ``q`` the digit bound of a
    homogeneous space, ``axes`` the per-axis bounds of a mixed-radix one;
    each mismatch raises ValueError naming the file.  Rows of unequal
    width raise ValueError too.  An empty file gives shape (0, n)."""
    rows = []
    with open(path) as f:
        for line in f:
            s = line.strip()
            if not s or s.startswith("#"):
                continue
            rows.append([int(x) for x in s.split()])
    if not rows:
        width = n if n is not None else (len(axes) if axes is not None else 0)
        return np.zeros((0, width), dtype=np.uint8)
    widths = {len(r) for r in rows}
    if len(widths) != 1:
        raise ValueError(f"{path}: rows of different widths {sorted(widths)}")
    arr = np.array(rows, dtype=np.int64)
    if n is not None and arr.shape[1] != n:
        raise ValueError(f"{path}: rows have {arr.shape[1]} digits, want n={n}")
    if arr.min() < 0:
        raise ValueError(f"{path}: negative digit {arr.min()}")
    if q is not None and arr.max() >= q:
        raise ValueError(f"{path}: digit {arr.max()} out of range for q={q}")
    if axes is not None:
        ax = np.asarray(axes, dtype=np.int64)
        if ax.size != arr.shape[1]:
            raise ValueError(f"{path}: rows have {arr.shape[1]} digits, want n={ax.size}")
        bad = np.nonzero((arr >= ax[None, :]).any(axis=0))[0]
        if bad.size:
            j = int(bad[0])
            raise ValueError(f"{path}: digit {arr[:, j].max()} out of range for "
                             f"axis {j} of size {ax[j]}")
    if arr.max() > 255:
        raise ValueError(f"{path}: digit {arr.max()} does not fit uint8")
    return arr.astype(np.uint8)
