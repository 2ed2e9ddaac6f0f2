"""Epoch series on the GPU: cor_trace_plan_set_epoch and cor_gpu_trace_series.

The reference traces one epoch per run (src/main.c:38-86: igrf14 recalc for the
header's date, the model's driver line, then trajectory_simulation_cutoff); a storm
sequence is one run per epoch.  Here the same scans are traced at a list of epochs,
several in flight at once.  Bar: every epoch of a series is bit-identical to a single
trace of that epoch (FP64 and per-lane mixed are deterministic), a plan re-pointed
with set_epoch is bit-identical to a fresh plan, and the TS05 storm epochs agree with
the oracle at the FP64 trace bar.
"""
import numpy as np
import pytest

import oracle_lib as L
from test_gpu_trace import compare, scan

pytestmark = pytest.mark.gpu


def storm_epochs(cor, minutes, seed=4):
    """Epochs 2020-01-01 00:00 + the given minutes, with seeded TS05 drivers: Pdyn
    U(0.5, 10) nPa, Dst a random walk clamped to [-300, 0] nT, By, Bz ~ N(0, 5) nT,
    W1..W6 U(0, 2).  Returns the library's and the oracle's parameter blocks."""
    rng = np.random.default_rng(seed)
    dst = -20.0
    out = []
    for m in minutes:
        dst = float(np.clip(dst + rng.normal(0.0, 25.0), -300.0, 0.0))
        parmod = [rng.uniform(0.5, 10.0), dst, rng.normal(0, 5), rng.normal(0, 5)] + list(rng.uniform(0, 2, 6))
        t = (2020, 1, 1, 1, int(m) // 60, int(m) % 60, 0)
        out.append((t, parmod))
    return out


def both_params(cor, oracle, t, parmod):
    p = cor.recalc(*t, parmod=parmod)
    _, po = oracle.recalc(*t)
    for k, v in enumerate(parmod):
        po.PARMOD[k] = v
    return p, po


def site_rows(cor, sites, rig=1.0, dl=0.5, rk=20.0, nk1=100):
    return cor.scans_array([(r0, la, lo, la, lo, rig, dl, rk, -1.0, nk1) for (r0, la, lo) in sites])


SITES = [(1.00310, 52.0, 30.0), (1.00300, -20.0, 150.0), (1.00290, 65.0, 250.0)]


@pytest.mark.parametrize("in_flight", [1, 3])
def test_ts05_series_matches_single_epochs(cor, in_flight):
    """Four storm epochs (00:00, 03:05, 12:30, 23:55) x three sites, TS05 + IGRF, FP64:
    the series equals one cor_gpu_trace per epoch, bit for bit."""
    ep = storm_epochs(cor, [0, 185, 750, 1435])
    params = [cor.recalc(*t, parmod=pm) for t, pm in ep]
    rows = site_rows(cor, SITES)
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_TS05)
    tr, cut = cor.trace_series(ms, params, rows, in_flight=in_flight, step_limit=8000)
    assert tr.shape == (4, 3 * 39) and cut.shape == (4, 3)
    for e, p in enumerate(params):
        t1, c1 = cor.trace(ms, p, rows, step_limit=8000)
        assert np.array_equal(tr[e], t1), e
        assert np.array_equal(cut[e], c1), e
    print("series cutoffs (Rc):", cut["rc"].tolist())
    # the epochs differ: the drivers move the cutoffs
    assert len({tuple(np.round(c["rc"], 6)) for c in cut}) > 1


@pytest.mark.parametrize("precision", [1, 0], ids=["mixed-cooperative", "fp64"])
def test_t96_series_graph_replays_match_eager_traces(cor, precision):
    """IGRF + T96 at five epochs with two in flight: each in-flight slot captures its
    epoch once as a graph (epoch upload, lane tables, reset, trace, penumbra, result
    copies) and replays it for the later epochs, with the host collecting results and
    rebuilding the epoch block between replays.  Every replay equals an eager trace of
    its epoch bit for bit -- in the mixed path with the cooperative tail on (every
    trajectory handed over after 20 steps), whose park records carry a run id that the
    device advances at every replay."""
    drivers = [[2, -20, 0, -2], [8, -250, 4, -12], [4, -80, -3, 5], [2, -20, 0, -2], [6, -150, 1, -8]]
    params = [cor.recalc(2020, 1, 1, 1, 3 * k, 10 * k, 0, parmod=pm) for k, pm in enumerate(drivers)]
    rows = site_rows(cor, SITES, rig=0.5, dl=0.25, nk1=100)
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    kw = dict(step_limit=100000 if precision else 20000, precision=precision, handover_steps=20 if precision else 0)
    tr, cut = cor.trace_series(ms, params, rows, in_flight=2, **kw)
    for e, p in enumerate(params):
        t1, c1 = cor.trace(ms, p, rows, **kw)
        assert np.array_equal(tr[e], t1), e
        assert np.array_equal(cut[e], c1), e
    if precision:
        assert (tr["handover"] > 0).any()
    assert len({tuple(np.round(c["rc"], 6)) for c in cut}) > 1


@pytest.mark.parametrize("precision,ext", [(0, 2), (1, 2), (1, 0)], ids=["fp64-ts05", "mixed-ts05", "mixed-igrf"])
def test_multi_epoch_launch_equals_single_epochs(cor, precision, ext):
    """Six storm epochs x three sites in ONE launch (cor_gpu_trace_epochs: a single queue
    of (epoch, scan, rigidity) trajectories, each lane reading its epoch's block): every
    epoch equals a single-epoch trace of it bit for bit, in FP64 and in mixed precision
    (mixed with both queue orders); the plan form gives the same."""
    ep = storm_epochs(cor, [0, 95, 185, 750, 1200, 1435])
    params = [cor.recalc(*t, parmod=pm) for t, pm in ep]
    rows = site_rows(cor, SITES)
    ms = cor.model_set(cor.INTERNAL_IGRF, ext)
    for qo in ((1,) if precision == 0 else (0, 1)):
        tr, cut = cor.trace_epochs(ms, params, rows, step_limit=8000, precision=precision, queue_order=qo)
        assert tr.shape == (6, 3 * 39) and cut.shape == (6, 3)
        for e, p in enumerate(params):
            t1, c1 = cor.trace(ms, p, rows, step_limit=8000, precision=precision, coop=-1)
            assert np.array_equal(tr[e], t1), (qo, e)
            assert np.array_equal(cut[e], c1), (qo, e)
    plan = cor.Plan(ms, params, rows, step_limit=8000, precision=precision, queue_order=1)
    plan.run()
    tp, cp = plan.results()
    st = plan.stats()
    plan.close()
    assert np.array_equal(tp, tr) and np.array_equal(cp, cut)
    assert st.n_traj == 6 * 3 * 39 and st.steps == int(tr["steps"].sum())
    if ext == 2:
        assert len({tuple(np.round(c["rc"], 6)) for c in cut}) > 1


def test_multi_epoch_launch_rejects_t96(cor):
    """The multi-epoch tracer has no T96 model set (its per-lane evaluator keeps the
    epoch's T96 constants in registers): the call reports it instead of tracing."""
    params = [cor.recalc(2020, 1, 1, 1, h, 0, 0, parmod=[2, -20, 0, -2]) for h in (0, 1)]
    with pytest.raises(cor.CorError, match="cor_gpu_trace_epochs"):
        cor.trace_epochs(cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96), params, site_rows(cor, SITES[:1]))


def test_ts05_series_against_oracle(cor, oracle):
    """Two storm epochs at one site against the oracle, FP64 bar."""
    ep = storm_epochs(cor, [0, 750])
    rows = site_rows(cor, SITES[:1])
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_TS05)
    params = [cor.recalc(*t, parmod=pm) for t, pm in ep]
    tr, cut = cor.trace_series(ms, params, rows, step_limit=20000)
    r0, la, lo = SITES[0]
    for e, (t, pm) in enumerate(ep):
        _, po = both_params(cor, oracle, t, pm)
        osc = scan(la, lo, r0=r0, rig=1.0, dl=0.5, rk=20.0, nk1=100)
        tc, cc = oracle.trace_scan(osc, L.models(internal=1, external=2), po, 20000)
        print("epoch", t, "gpu", cut[e][0], "cpu", cc)
        compare(rows[0], tr[e], cut[e][0], tc, cc, rc_tol=0.5)


@pytest.mark.parametrize("precision,coop,hs", [(0, 0, 0), (1, -1, 0), (1, 0, 0), (1, 0, 20)],
                         ids=["fp64", "mixed-per-lane", "mixed-default", "mixed-handover-20"])
def test_set_epoch_equals_fresh_plan(cor, precision, coop, hs):
    """IGRF + T96: one plan re-pointed quiet -> storm -> quiet gives the fresh plans'
    results bit for bit -- also in the mixed path's cooperative mode, whose lane tables
    are rebuilt from each epoch and whose sums have one order for every group width."""
    rows = site_rows(cor, SITES, rig=0.5, dl=0.25, nk1=100)
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    quiet = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
    storm = cor.recalc(2020, 1, 1, 1, 13, 20, 0, parmod=[8, -250, 4, -12])
    kw = dict(step_limit=100000 if precision else 30000, precision=precision, coop=coop, handover_steps=hs)
    plan = cor.Plan(ms, quiet, rows, **kw)
    got = []
    for p in (quiet, storm, quiet):
        if got:
            plan.set_epoch(p)
        plan.run()
        got.append(plan.results())
    plan.close()
    for (tg, cg), p in zip(got, (quiet, storm, quiet)):
        tf, cf = cor.trace(ms, p, rows, **kw)
        assert np.array_equal(tg, tf) and np.array_equal(cg, cf)
    assert not np.array_equal(got[0][1]["rc"], got[1][1]["rc"])


def test_set_epoch_cooperative_tables_follow_the_epoch(cor):
    """Mixed precision with every trajectory in the cooperative mode: its lane tables are
    rebuilt from the new epoch (T96 drivers changed), so a re-pointed plan gives the fresh
    plan's results bit for bit, and both agree with FP64 at the mixed trace bar (at least
    0.25 GV: five rigidity steps of this scan, or half the penumbra)."""
    rows = site_rows(cor, SITES, rig=0.5, dl=0.05, nk1=100)
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    quiet = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
    storm = cor.recalc(2020, 1, 1, 1, 13, 20, 0, parmod=[8, -250, 4, -12])
    kw = dict(step_limit=60000, precision=cor.PREC_MIXED, handover_steps=1)
    plan = cor.Plan(ms, quiet, rows, **kw)
    plan.run()
    plan.results()
    plan.set_epoch(storm)
    plan.run()
    tg, cg = plan.results()
    plan.close()
    tf, cf = cor.trace(ms, storm, rows, **kw)
    assert np.array_equal(tg, tf) and np.array_equal(cg, cf)
    _, c64 = cor.trace(ms, storm, rows, step_limit=60000)
    for a, d in zip(cg, c64):
        tol64 = max(0.25, 0.5 * (d["ru"] - d["rl"]))
        print("cooperative", a["rc"], "fp64", d["rc"], "tol", tol64)
        assert abs(a["rc"] - d["rc"]) <= tol64 + 1e-9
