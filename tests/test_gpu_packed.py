"""Packed-layer sweeps: every field bit-exact against the CPU oracle on the
cell shapes that select each register format, compact inter-sweep format
and tile shape.

The packed path keeps all R+1 distance layers of a word in bit fields of
one to four dwords (8-bit fields when the closed-form layer bounds allow,
16-bit otherwise), so a wrong bound, a field that overflows into its
neighbour or a wrong shift shows up as a wrong count somewhere in the
field.  Shapes: binary cells up to R = 6 (one and two dword formats, 8- and
16-bit fields), q = 3, 4, 5 (fiber sums T in the fields), multiset codes
(seed values above 1), mu = 2 predicates, and small LDS budgets that force
short first sweeps and multi-round strided tiles.
"""
import numpy as np
import pytest

from oracle.dp import NumpyDP

pytestmark = pytest.mark.gpu

CELLS = [([2] * 16, 3), ([2] * 14, 4), ([2] * 12, 6), ([2] * 13, 1),
         ([2] * 11, 2), ([2] * 10, 3), ([2] * 18, 3),
         ([3] * 9, 3), ([3] * 8, 1), ([4] * 7, 3), ([5] * 6, 3), ([5] * 6, 2)]
IDS = [f"{a[0]}^{len(a)}-R{r}" for a, r in CELLS]
# schedule variants: LDS budgets that force short first sweeps and
# multi-round strided tiles; the column kernel off, or whole-cube only; the
# first-sweep kernels; narrow (16 / 24-bit) inter-sweep words
VARIANTS = {"default": {}, "lds16k": {"CC_SWEEP_LDS": "16384"},
            "lds4k": {"CC_SWEEP_LDS": "4096"}, "nocol": {"CC_PK_COL": "0"},
            "col1": {"CC_PK_COL": "1"}, "nofirst": {"CC_PK_FIRST": "0"},
            "xlanefirst": {"CC_PK_FIRST": "2"}, "narrow": {"CC_PK_NARROW": "1"}}


def _check_cell(g, c, rng, mults):
    size = c.space.size
    every = np.arange(size, dtype=np.int64)
    for m in (max(2, size // 400), max(3, size // 60)):
        base = rng.integers(0, size, size=m).astype(np.int64)
        code = np.repeat(base, rng.integers(1, mults + 1, size=m))
        g.recount(code)
        c.recount(code)
        assert (g.read_cnt() == c.read_cnt()).all(), ("cnt", m)
        for which in ("gain_map", "loss_map"):
            getattr(g, which)()
            getattr(c, which)()
            assert (g.map_read_at(every) == c.map_read_at(every)).all(), (which, m)


@pytest.mark.parametrize("variant", list(VARIANTS))
@pytest.mark.parametrize("axes,R", CELLS, ids=IDS)
def test_packed_fields_match_oracle(gpu_cell, monkeypatch, axes, R, variant):
    for k, v in VARIANTS[variant].items():
        monkeypatch.setenv(k, v)
    rng = np.random.default_rng(len(axes) * 10 + R)
    g = gpu_cell(axes, R)
    plan = g.sweep_plan2(0)
    assert all(sw["W"] in (8, 16) for sw in plan), plan   # packed path taken
    assert sum(sw["axes"] for sw in plan) == len(axes)
    _check_cell(g, NumpyDP(axes, R), rng, mults=3)


@pytest.mark.parametrize("axes,R", [([2] * 12, 3), ([3] * 7, 2)],
                         ids=["2^12-R3", "3^7-R2"])
def test_packed_mu2_predicates(gpu_cell, axes, R):
    rng = np.random.default_rng(7)
    g = gpu_cell(axes, R, mu=2)
    _check_cell(g, NumpyDP(axes, R, mu=2), rng, mults=2)


@pytest.mark.parametrize("lcmp", ["1", "0"])
@pytest.mark.parametrize("axes,R", [([5] * 8, 3), ([3] * 12, 3)],
                         ids=["5^8-R3", "3^12-R3"])
def test_compact_lds_two_round_sweep(gpu_cell, monkeypatch, axes, R, lcmp):
    """Two-round LDS-tiled sweeps with two-dword register words: the tile
    crosses LDS as one compact dword per word (CC_PK_LCMP=1, half the LDS)
    or as both register dwords; fields exact either way."""
    monkeypatch.setenv("CC_PK_LCMP", lcmp)
    monkeypatch.setenv("CC_AUTOTUNE", "0")
    print(f"compact LDS {axes[0]}^{len(axes)} R{R} lcmp={lcmp}", flush=True)
    g = gpu_cell(axes, R)
    last = g.sweep_plan2(0)[-1]
    assert (last["kernel"], last["ND"], last["rounds"], last["TL"]) == (1, 2, 2, 64)
    per_word = 4 if lcmp == "1" else 8
    words = last["P"] * last["TL"]
    assert per_word * words <= last["lds_bytes"] <= per_word * (words + words // 64 + 1)
    rng = np.random.default_rng(len(axes) + R)
    _check_cell(g, NumpyDP(axes, R), rng, mults=3)


@pytest.mark.parametrize("fold", ["api", "env", "off"])
@pytest.mark.parametrize("axes,R,mu", [([2] * 16, 3, 1), ([2] * 10, 3, 1),
                                       ([5] * 8, 3, 1), ([3] * 9, 3, 2),
                                       ([2] * 14, 4, 2)],
                         ids=["2^16-R3", "2^10-R3-one-sweep", "5^8-R3",
                              "3^9-R3-mu2", "2^14-R4-mu2"])
def test_folded_census_equals_field_scans(gpu_cell, monkeypatch, axes, R, mu, fold):
    """The last packed sweep folds the map maximum (gain / loss) and the cnt
    deficit (recount) into the transform; map_max, count_eq(0) and the
    deficit then equal what a scan of the fields gives, also after updates
    change cnt (the folded deficit no longer applies) and across gain /
    loss / recount in any order.  Off (the default) scans every time."""
    if fold == "env":
        monkeypatch.setenv("CC_FOLD", "1")
    monkeypatch.setenv("CC_AUTOTUNE", "0")
    print(f"folded census {axes[0]}^{len(axes)} R{R} mu{mu} fold={fold}", flush=True)
    rng = np.random.default_rng(len(axes) * 7 + R)
    g = gpu_cell(axes, R, mu=mu)
    if fold == "api":
        assert g.set_fold(True) is False          # off at init
        assert g.set_fold(True) is True
    size = int(np.prod(axes))

    def census():
        cnt = g.read_cnt().astype(np.int64)
        assert g.count_eq(0) == int((cnt == 0).sum())
        assert g.uncovered() == int(np.maximum(0, mu - cnt).sum())

    for m in (3, max(4, size // 200), max(5, size // 20)):
        code = rng.integers(0, size, size=m).astype(np.int64)
        g.recount(code)
        census()
        for which in ("gain_map", "loss_map", "gain_map"):
            getattr(g, which)()
            assert g.map_max() == max(0, int(g.map_all().max())), which
        census()
        g.ball_update(rng.integers(0, size, size=2).astype(np.int64), +1)
        census()
        g.loss_map()
        g.ball_update(code[:1], -1)
        assert g.map_max() == max(0, int(g.map_all().max()))
        census()


def test_packed_equals_unpacked_engine_k2_20_3(monkeypatch):
    """The two GPU paths agree on a 2^20-word cell (gain, loss, recount)."""
    from coldcore_amd.gpu import GpuBackend
    axes, R = [2] * 20, 3
    code = np.flatnonzero(np.random.default_rng(3).random(1 << 20) < 3e-4)
    out = {}
    for mode in ("1", "0"):
        monkeypatch.setenv("CC_PACKED", mode)
        with GpuBackend(axes, R) as g:
            packed = any(sw["W"] != 32 for sw in g.sweep_plan2(0))
            assert packed == (mode == "1")
            g.recount(code)
            cnt = g.read_cnt().copy()
            g.gain_map()
            gain = g.map_all().copy()
            g.loss_map()
            out[mode] = (cnt, gain, g.map_all().copy())
    for a, b in zip(out["1"], out["0"]):
        assert (a == b).all()


def test_k2_24_3_schedule_is_three_packed_sweeps(gpu_cell):
    """K_2(24,3): axes 0-11 with 8-bit fields in one dword, then two
    six-axis strided sweeps with 16-bit fields; compact inter-sweep words of
    4 bytes, so a gain pass moves 2+4 + 4+4 + 4+4 = 22 bytes per word."""
    g = gpu_cell([2] * 24, 3)
    plan = g.sweep_plan2(0)
    assert [(s["first_axis"], s["axes"]) for s in plan] == [(0, 12), (12, 6), (18, 6)]
    assert [s["W"] for s in plan] == [8, 16, 16]
    assert sum(s["bytes_in"] + s["bytes_out"] for s in plan) == 22


def test_autotune_keeps_fields_exact(monkeypatch):
    """The autotuner only changes which schedule runs: after it, gain, loss
    and recount equal the planner's default schedule bit for bit, on a
    2^22-word cell (where the engine also tunes by itself)."""
    from coldcore_amd.gpu import GpuBackend
    monkeypatch.setenv("CC_AUTOTUNE", "0")
    axes, R = [2] * 22, 3
    code = np.flatnonzero(np.random.default_rng(4).random(1 << 22) < 2e-4)
    with GpuBackend(axes, R) as g:
        def fields():
            g.recount(code)
            cnt = g.read_cnt().copy()
            g.gain_map()
            gain = g.map_all().copy()
            g.loss_map()
            return cnt, gain, g.map_all().copy()
        before = fields()
        n, us = g.autotune(2)
        assert n >= 2 and us > 0
        after = fields()
    for a, b in zip(before, after):
        assert (a == b).all()


def test_autotune_unpacked_budgets_keep_fields_exact(monkeypatch):
    """A mixed-radix cell has no packed schedule: the autotuner times the
    unpacked plan rebuilt under each LDS budget, and the winner's fields
    equal the default plan's bit for bit."""
    from coldcore_amd.gpu import GpuBackend
    monkeypatch.setenv("CC_AUTOTUNE", "0")
    axes, R = [4, 5, 6, 7, 8, 4, 5, 6], 3
    size = int(np.prod(axes))
    code = np.flatnonzero(np.random.default_rng(6).random(size) < 3e-4)
    with GpuBackend(axes, R) as g:
        def fields():
            g.recount(code)
            cnt = g.read_cnt().copy()
            g.gain_map()
            gain = g.map_all().copy()
            g.loss_map()
            return cnt, gain, g.map_all().copy()
        before = fields()
        n, us = g.autotune(2)
        assert n >= 3 and us > 0          # several unpacked plans timed
        after = fields()
    for a, b in zip(before, after):
        assert (a == b).all()


# ---- the kernel instantiations of the 2^30-word R = 4 schedules ----------
# K_2(30,4) runs, depending on the autotuner's pick, a 10- or 12-axis first
# sweep (k_pfirstp with 8-bit fields in 2 dwords, or 16-bit fields in 3)
# and five- or six-axis column sweeps (one lane or a lane pair per cube)
# whose compact inter-sweep words are one dword plane up to 15 axes and two
# planes from 16 axes on (field d takes bitlen(C(c,d)) bits).  The cuts
# below reproduce every one of those kernels, with one- and two-plane words
# in and out, on 2^22 / 2^23-word cells, checked against the Walsh-Hadamard
# convolution.  Rows per sweep: (first axis, axes, kernel, RR, W, ND, bytes
# in, bytes out); kernel 5 = persistent first sweep, 2 / 3 = column kernel
# with 1 / 2 lanes per cube, 1 = LDS-tiled.
FORCED_R4 = {
    "first12-col2": (22, [12, 17], "2", [(0, 12, 5, 5, 16, 3, 2, 4),
                                         (12, 5, 3, 5, 16, 3, 4, 8),
                                         (17, 5, 3, 5, 16, 3, 8, 4)]),
    "first12-col1": (22, [12, 17], "1", [(0, 12, 5, 5, 16, 3, 2, 4),
                                         (12, 5, 2, 5, 16, 3, 4, 8),
                                         (17, 5, 2, 5, 16, 3, 8, 4)]),
    "first10-col1": (22, [10, 15, 20], "1", [(0, 10, 5, 4, 8, 2, 2, 4),
                                             (10, 5, 2, 5, 16, 3, 4, 4),
                                             (15, 5, 2, 5, 16, 3, 4, 8),
                                             (20, 2, 2, 2, 16, 3, 8, 4)]),
    "first10-col2": (22, [10, 15, 20], "2", [(0, 10, 5, 4, 8, 2, 2, 4),
                                             (10, 5, 3, 5, 16, 3, 4, 4),
                                             (15, 5, 3, 5, 16, 3, 4, 8),
                                             (20, 2, 2, 2, 16, 3, 8, 4)]),
    "mid5-col1": (22, [10, 16, 21], "1", [(0, 10, 5, 4, 8, 2, 2, 4),
                                          (10, 6, 1, 5, 16, 3, 4, 8),
                                          (16, 5, 2, 5, 16, 3, 8, 8),
                                          (21, 1, 1, 1, 16, 3, 8, 4)]),
    "mid5-col2": (22, [10, 16, 21], "2", [(0, 10, 5, 4, 8, 2, 2, 4),
                                          (10, 6, 3, 6, 16, 3, 4, 8),
                                          (16, 5, 3, 5, 16, 3, 8, 8),
                                          (21, 1, 1, 1, 16, 3, 8, 4)]),
    "pair6-last": (22, [10, 16], "2", [(0, 10, 5, 4, 8, 2, 2, 4),
                                       (10, 6, 3, 6, 16, 3, 4, 8),
                                       (16, 6, 3, 6, 16, 3, 8, 4)]),
    # the LDS-free first sweep (k_pfirst: 4 axes in-thread, 6 across the
    # lanes), which the autotuner may pick for K_2(30,4)'s first 10 axes
    "xlane10-col1": (22, [10, 15, 20], "1", [(0, 10, 4, 4, 8, 2, 2, 4),
                                             (10, 5, 2, 5, 16, 3, 4, 4),
                                             (15, 5, 2, 5, 16, 3, 4, 8),
                                             (20, 2, 2, 2, 16, 3, 8, 4)], "2"),
    "pair6-mid": (23, [10, 16, 22], "2", [(0, 10, 5, 4, 8, 2, 2, 4),
                                          (10, 6, 3, 6, 16, 3, 4, 8),
                                          (16, 6, 3, 6, 16, 3, 8, 8),
                                          (22, 1, 1, 1, 16, 3, 8, 4)]),
}
_WALSH = {}


def _walsh_fields(n, R):
    """(code, cnt, gain, loss) of a seeded multiset code on K_2(n,R) by the
    Walsh-Hadamard convolution; computed once per n."""
    from oracle.walsh import binary_field
    if (n, R) not in _WALSH:
        rng = np.random.default_rng(100 + n)
        base = np.flatnonzero(rng.random(1 << n) < 4e-4).astype(np.int64)
        code = np.repeat(base, rng.integers(1, 4, size=len(base)))
        ind = np.bincount(code, minlength=1 << n).astype(np.int64)
        cnt = binary_field(ind, n, R)
        gain = binary_field((cnt == 0).astype(np.int64), n, R)
        loss = binary_field((cnt == 1).astype(np.int64), n, R)
        _WALSH[(n, R)] = (code, cnt, gain, loss)
    return _WALSH[(n, R)]


@pytest.mark.parametrize("case", list(FORCED_R4))
def test_r4_schedule_kernels_match_walsh(gpu_cell, monkeypatch, case):
    n, cuts, col, rows, *first = FORCED_R4[case]
    monkeypatch.setenv("CC_PK_COL", col)
    if first:
        monkeypatch.setenv("CC_PK_FIRST", first[0])
    print(f"forced R=4 schedule {case}: n={n} cuts={cuts} col={col}", flush=True)
    code, cnt, gain, loss = _walsh_fields(n, 4)
    g = gpu_cell([2] * n, 4)
    assert g.sweep_force(cuts) == len(rows)
    got = [(s["first_axis"], s["axes"], s["kernel"], s["RR"], s["W"], s["ND"],
            s["bytes_in"], s["bytes_out"]) for s in g.sweep_plan2(0)]
    assert got == rows
    g.recount(code)
    assert (g.read_cnt().astype(np.int64) == cnt).all()
    g.gain_map()
    assert (g.map_all().astype(np.int64) == gain).all()
    g.loss_map()
    assert (g.map_all().astype(np.int64) == loss).all()


# ---- the two-sweep shapes (CC_PK_WIDE): k_pfirstw and k_pstrip ----------
# A 13- or 14-axis first sweep whose byte tile widens to 16-bit fields in
# its last round (kernel 6) and a strided sweep of 2 RA axes in two
# register rounds through a compact LDS tile (kernel 7), against the
# Walsh-Hadamard convolution; a set code (its recount takes the widening
# first sweep too, seeded from the sorted word list) and mu = 2 predicates.
WIDE = {"k2_22_3": (22, 3, 1, None, [(0, 12, 5), (12, 10, 7)]),
        "k2_22_3_cut14": (22, 3, 1, [14], [(0, 14, 6), (14, 8, 7)]),
        "k2_23_3": (23, 3, 1, None, [(0, 13, 6), (13, 10, 7)]),
        "k2_22_3_cut14_mu2": (22, 3, 2, [14], [(0, 14, 6), (14, 8, 7)]),
        "k2_24_3": (24, 3, 1, None, [(0, 14, 6), (14, 10, 7)])}


@pytest.mark.parametrize("case", list(WIDE))
def test_wide_schedule_kernels_match_walsh(gpu_cell, monkeypatch, case):
    from oracle.walsh import binary_field
    n, R, mu, cuts, rows = WIDE[case]
    monkeypatch.setenv("CC_PK_WIDE", "1")
    monkeypatch.setenv("CC_AUTOTUNE", "0")
    print(f"two-sweep schedule {case}", flush=True)
    rng = np.random.default_rng(7 + n)
    code = np.flatnonzero(rng.random(1 << n) < 2e-4).astype(np.int64)
    ind = np.zeros(1 << n, dtype=np.int64)
    ind[code] = 1
    cnt = binary_field(ind, n, R)
    g = gpu_cell([2] * n, R, mu=mu)
    if cuts:
        g.sweep_force(cuts)
    got = [(s["first_axis"], s["axes"], s["kernel"]) for s in g.sweep_plan2(0)]
    assert got == rows
    g.recount(code)
    assert [s["kernel"] for s in g.sweep_plan2(1)] == [r[2] for r in rows]
    assert (g.read_cnt().astype(np.int64) == cnt).all()
    g.gain_map()
    gain = binary_field((cnt < mu).astype(np.int64), n, R)
    assert (g.map_all().astype(np.int64) == gain).all()
    g.loss_map()
    loss = binary_field(((cnt >= 1) & (cnt <= mu)).astype(np.int64), n, R)
    assert (g.map_all().astype(np.int64) == loss).all()
