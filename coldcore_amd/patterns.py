"""Host-side support-table ("ball") builders: the v1 and ox wire formats
and the additive wide (n > 16) and long-axis (ax > 16) formats.

The packed formats are part of the plugin ABI v1 (reference
docs/PLUGIN.md:142-152, src/include/coldcore/plugin.h:56-61), so they are
kept byte-identical to the reference's; the enumeration order is the
reference's too (weight ascending, position subsets in lexicographic order,
per-position deltas with the last position varying fastest; reference
src/coldcore/patterns.py:71-121), which makes tables interchangeable
between the two engines.

v1 : one uint64 per pattern; byte k = (pos << 4 | delta), delta a circular
     shift in 1..ax_pos-1, zero byte terminates, <= 8 bytes; entry 0 is the
     empty pattern (the centre).
ox : (mask uint16, nibble offsets uint64); nibble p of offs = delta at p.

The builder is vectorised per weight: all position subsets of one weight
share one delta grid on homogeneous axes, so K_2(30,4)'s 31,931-entry table
is built without a Python loop over subsets.
"""
from itertools import combinations
from math import comb

import numpy as np

PROBLEM_IDS = {"hamming": 0, "torus_linf": 1, "lee": 2, "grid_linf": 3}


def _lee_wt(delta, a):
    return np.minimum(delta, a - delta)


def delta_sets(axes, R, problem):
    """Admissible circular shifts per position (reference patterns.py:34-50)."""
    out = []
    for a in axes:
        if problem == "hamming":
            out.append(list(range(1, a)))
        elif problem == "torus_linf":
            out.append(list(range(1, R + 1)) + list(range(a - R, a)))
        elif problem == "grid_linf":
            out.append(sorted({d for d in range(1, a) if d <= R or a - d <= R}))
        elif problem == "lee":
            out.append([d for d in range(1, a) if min(d, a - d) <= R])
        else:
            raise ValueError(problem)
    return out


def max_pattern_weight(n, R, problem):
    """Max changed positions per pattern (reference patterns.py:53-59)."""
    if problem == "hamming":
        return R
    if problem == "lee":
        return min(n, R)
    return n


def _delta_grid(sets):
    grids = np.meshgrid(*[np.asarray(s, dtype=np.int64) for s in sets],
                        indexing="ij")
    return np.stack([g.ravel() for g in grids], axis=1)


def _enumerate(axes, R, problem):
    """Yield (subs (S, w) int64, deltas (S, K, w) or (K, w) int64) blocks."""
    n = len(axes)
    ds = delta_sets(axes, R, problem)
    maxw = min(max_pattern_weight(n, R, problem), n)
    if maxw > 8:
        raise ValueError("support patterns cap at 8 changed positions")
    homog = len(set(axes)) == 1
    for w in range(1, maxw + 1):
        subs = np.array(list(combinations(range(n), w)), dtype=np.int64)
        if homog:
            d = _delta_grid([ds[0]] * w)
            if problem == "lee":
                d = d[_lee_wt(d, axes[0]).sum(axis=1) <= R]
            if len(d):
                yield subs, d
            continue
        for row in subs:
            d = _delta_grid([ds[p] for p in row])
            if problem == "lee":
                wt = sum(_lee_wt(d[:, k], axes[p]) for k, p in enumerate(row))
                d = d[wt <= R]
            if len(d):
                yield row[None, :], d


def _pack(axes, R, problem, bits, nfields, maxpos, pshift=4, maxax=16):
    axes = [int(a) for a in axes]
    if problem not in PROBLEM_IDS:
        raise ValueError(problem)
    if max(axes) > maxax:
        raise ValueError(f"this table format needs ax_j <= {maxax}")
    if len(axes) > maxpos:
        raise ValueError(f"this table format addresses at most {maxpos} axes")
    parts = [np.zeros(1, dtype=np.uint64)]
    for subs, d in _enumerate(axes, R, problem):
        w = subs.shape[1]
        if w > nfields:
            raise ValueError(f"this table format holds at most {nfields} "
                             "changed positions per pattern")
        out = np.zeros((len(subs), len(d)), dtype=np.uint64)
        for k in range(w):
            fld = ((subs[:, k][:, None] << pshift) | d[None, :, k]).astype(np.uint64)
            out |= fld << np.uint64(bits * k)
        parts.append(out.ravel())
    return np.concatenate(parts)


def support_patterns(axes, R, problem="hamming"):
    """Packed v1 support table (any problem, mixed radix, n <= 16)."""
    return _pack(axes, R, problem, 8, 8, 16)


def support_patterns_wide(axes, R, problem="hamming"):
    """Wide support table: the same enumeration, each pattern a uint64 of up
    to six 10-bit fields (pos << 4 | delta) at bits [10k, 10k+10), zero
    field terminates.  Positions up to 63, so Hamming and Lee cells with
    more than 16 axes (K_2(24,3), K_2(30,4)) have a table; the engine takes
    it through the additive ``dct_init3w``."""
    return _pack(axes, R, problem, 10, 6, 64)


def support_patterns_long(axes, R, problem="torus_linf"):
    """Long-axis support table: the same enumeration, each pattern a uint64
    of up to four 16-bit fields (pos << 10 | delta) at bits [16k, 16k+16),
    zero field terminates.  Deltas up to 1023, so king tori and grid boxes
    with axes longer than 16 (the v1 / ox nibble limit) have a table; the
    engine takes it through the additive ``dct_init3x``."""
    return _pack(axes, R, problem, 16, 4, 64, pshift=10, maxax=1024)


def support_patterns_ox(axes, R, problem="hamming"):
    """ox-format table: the same enumeration as :func:`support_patterns`."""
    axes = [int(a) for a in axes]
    masks = [np.zeros(1, dtype=np.uint16)]
    offs = [np.zeros(1, dtype=np.uint64)]
    for subs, d in _enumerate(axes, R, problem):
        w = subs.shape[1]
        m = np.zeros((len(subs), len(d)), dtype=np.uint64)
        o = np.zeros((len(subs), len(d)), dtype=np.uint64)
        for k in range(w):
            m |= (np.uint64(1) << subs[:, k].astype(np.uint64))[:, None]
            o |= (d[None, :, k].astype(np.uint64)
                  << (np.uint64(4) * subs[:, k].astype(np.uint64))[:, None])
        masks.append(m.ravel().astype(np.uint16))
        offs.append(o.ravel())
    return np.concatenate(masks), np.concatenate(offs)


def ball_patterns(q, n, R):
    """Homogeneous Hamming ball, v1 format (reference patterns.py:71-82)."""
    return support_patterns([q] * n, min(R, n), "hamming")


def ball_size(q, n, R):
    """|B_R| = sum_{r<=R} C(n,r) (q-1)^r (homogeneous Hamming)."""
    return sum(comb(n, r) * (q - 1) ** r for r in range(min(R, n) + 1))


def lee_ball_size(axes, R):
    """Lee ball cardinality: grades 0..R of prod_j (per-axis enumerator)."""
    coeff = [1] + [0] * R
    for a in axes:
        ka = [0] * (R + 1)
        for w in range(min(a // 2, R) + 1):
            ka[w] = 1 if (w == 0 or 2 * w == a) else 2
        coeff = [sum(coeff[i] * ka[d - i] for i in range(d + 1))
                 for d in range(R + 1)]
    return sum(coeff)


def decode_v1(pat):
    """One packed v1 pattern -> {pos: delta}."""
    return decode_fields(pat, 8)


def decode_fields(pat, bits):
    """One packed pattern of ``bits``-wide fields -> {pos: delta} (16-bit
    fields hold pos << 10 | delta, narrower ones pos << 4 | delta)."""
    out = {}
    p = int(pat)
    m = (1 << bits) - 1
    ps = 10 if bits == 16 else 4
    while p:
        b = p & m
        out[b >> ps] = b & ((1 << ps) - 1)
        p >>= bits
    return out
