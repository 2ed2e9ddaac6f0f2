"""The oracle against the reference's own outputs (tests/golden/, written by
tests/golden/make_golden.py from the reference compiled from its sources).

Same compiler flags and the same libm: the restatement is expected to be bit-identical
to the reference -- IGRF-14 epoch setup, IGRF synthesis, T96, TS05, the generic SH
synthesis (CHAOS form) with seeded tables, and the complete
text of cutoff runs (every allowed-trajectory line and the Rl/Ru/Rc summary).
"""
from pathlib import Path

import numpy as np
import pytest

import oracle_lib as L

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def vec():
    return np.load(GOLD / "field_vectors.npz")


def test_recalc_bitwise(oracle, vec):
    for k, e in enumerate(vec["recalc_epochs"]):
        rc, p = oracle.recalc(*[int(v) for v in e])
        assert rc == 0
        for f in ("G", "H", "REC", "A1", "A2", "A3"):
            assert np.array_equal(np.array(getattr(p, f)), vec[f"recalc{k}_{f}"]), (tuple(e), f)


def test_recalc_rejects_out_of_range_years(oracle):
    assert oracle.recalc(1899, 1, 1, 1, 0, 0, 0)[0] == -1
    assert oracle.recalc(2031, 1, 1, 1, 0, 0, 0)[0] == -1


def test_igrf_bitwise(oracle, vec):
    _, p = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    got = oracle.igrf_geo(p, vec["igrf_pts"])
    assert np.array_equal(got, vec["igrf_b"])


@pytest.mark.parametrize("k", range(4))
def test_ts05_bitwise(oracle, k):
    """lib/ts05/*.c at four driver sets (quiet to storm, W1..W6 on and off), bit for bit."""
    tv = np.load(GOLD / "ts05_vectors.npz")
    got = oracle.ts05(tv[f"ts05_params{k}"], tv["ts05_pts"])
    want = tv[f"ts05_b{k}"]
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    assert same.all(), np.argwhere(~same)[:5]


@pytest.mark.parametrize("k", range(2))
def test_chaos_bitwise(oracle, k):
    """chaos_geo (lib/chaos/chaos.c:120-227) with seeded degree-20 and degree-23 tables,
    bit for bit, including the NaN the reference returns exactly on the polar axis."""
    cv = np.load(GOLD / "chaos_vectors.npz")
    got = oracle.chaos(cv[f"chaos_g{k}"], cv[f"chaos_h{k}"], cv["chaos_pts"])
    want = cv[f"chaos_b{k}"]
    assert np.isnan(want).any()
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    assert same.all(), np.argwhere(~same)[:5]


@pytest.mark.parametrize("k", range(4))
def test_t96_bitwise(oracle, vec, k):
    got = oracle.t96(vec[f"t96_params{k}"], vec["t96_pts"])
    want = vec[f"t96_b{k}"]
    same = (got == want) | (np.isnan(got) & np.isnan(want))
    assert same.all(), np.argwhere(~same)[:5]


def _cutoff_run(oracle, name, models, parmod, limit):
    txt = (GOLD / f"{name}.out").read_text()
    ref_lines = [ln for ln in txt.splitlines() if len(ln.split()) == 9 and _numeric(ln)]
    _, ref_cut = L.parse_cutoff_output(txt)
    inp = (GOLD / f"{name}.in").read_text().split()
    rig, zn, rk, r0, the0, fi0, the1, fi1 = map(float, inp[:8])
    nk1, dl = int(inp[15]), float(inp[18])
    _, p = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    for k, v in enumerate(parmod):
        p.PARMOD[k] = v
    sc = L.OScan(r0, the0, fi0, the1, fi1, rig, dl, rk, zn, nk1, 0)
    tr, cut = oracle.trace_scan(sc, models, p, limit)
    lines = []
    for i, t in enumerate(tr):
        if t["status"] == 0:
            r = rig + i * dl
            lines.append("%10.6f%15.10f%12.6f%10.3f%10.3f%10.3f%10.3f%12.6f%16.2f" % (
                r, t["vc"], t["r"], t["td"], t["fei"], t["ast"], t["asf"], t["time"], t["length_m"] / 1000.))
    summary = "%12.5f%12.5f%12.5f" % tuple(cut)
    return lines, ref_lines, summary, txt


def _numeric(ln):
    try:
        [float(v) for v in ln.split()]
        return True
    except ValueError:
        return False


@pytest.mark.parametrize("name,models,parmod,limit", [
    ("config1_igrf_nk100", dict(internal=1, external=0), [], 1000000),
    ("t96_equator_nk400", dict(internal=1, external=1), [2, -20, 0, -2], 1000000),
    ("t96_mid_nk400", dict(internal=1, external=1), [2, -20, 0, -2], 1000000),
    ("t96_storm_limit_nk100", dict(internal=1, external=1), [6, -150, 4, -8], 20000),
    ("ts05_storm_nk100", dict(internal=1, external=2), [3, -100, 2, -5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5], 20000),
])
def test_cutoff_output_identical(oracle, name, models, parmod, limit):
    lines, ref_lines, summary, txt = _cutoff_run(oracle, name, L.models(**models), parmod, limit)
    assert lines == ref_lines
    assert summary in txt
