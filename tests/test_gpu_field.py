"""GPU field evaluators against the oracle (calculate_field, src/models.c:194-207).

FP64 path: the same expression order as the reference, compiled without FMA
contraction; the remaining difference is OCML vs glibc transcendentals (a few ULP),
so the bar is 1e-12 relative to |B|.

Mixed paths (one point per lane, and one point per wavefront as the tracer's
cooperative mode evaluates it): FP32 field.  T96's shielding fits cancel terms 1e4-1e5 times larger than
their sum, so its FP32 error is measured against the field the particle sees: |dB|
divided by |B| + B_dip(r) (B_dip = 30574 nT / r^3, the dipole scale, which is what
IGRF adds on top of the external field).  Bar: max 1e-3, 99th percentile 1e-4.
"""
import numpy as np
import pytest

import oracle_lib as L

pytestmark = pytest.mark.gpu

RE = 6.3712e6


def points(seed, n, rmin=1.0, rmax=25.0):
    rng = np.random.default_rng(seed)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    r = rng.uniform(rmin, rmax, size=(n, 1))
    return d * r * RE


def box_points(seed, n):
    """GSM points spread over the T96 domain, tail included (x down to -40 Re)."""
    rng = np.random.default_rng(seed)
    p = np.column_stack([rng.uniform(-40, 12, n), rng.uniform(-18, 18, n), rng.uniform(-18, 18, n)])
    p = p[np.linalg.norm(p, axis=1) > 1.0]
    return p * RE


CASES = [
    ("igrf", dict(internal=1, external=0), None, points(1, 4000)),
    ("t96", dict(internal=0, external=1), [2, -20, 0, -2], box_points(2, 6000)),
    ("igrf+t96", dict(internal=1, external=1), [2, -20, 0, -2], box_points(3, 6000)),
    ("igrf+t96 storm", dict(internal=1, external=1), [6, -150, 4, -8], box_points(4, 4000)),
    ("igrf+t96 quiet", dict(internal=1, external=1), [0.8, 5, -3, 2], box_points(5, 4000)),
]


@pytest.mark.parametrize("name,kinds,parmod,pts", CASES, ids=[c[0] for c in CASES])
def test_field_fp64_matches_oracle(cor, oracle, name, kinds, parmod, pts):
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[parmod[0], parmod[1], parmod[2], parmod[3]] if parmod else None)
    _, po = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    if parmod:
        for k, v in enumerate(parmod):
            po.PARMOD[k] = v
    got = cor.field(cor.model_set(**kinds), p, pts, cor.PREC_FP64)
    want = oracle.field(L.models(**kinds), po, pts)
    scale = np.linalg.norm(want, axis=1)
    err = np.linalg.norm(got - want, axis=1) / scale
    print(name, "max rel err", err.max())
    assert np.all(np.isfinite(got))
    assert err.max() < 1e-12


@pytest.mark.parametrize("name,kinds,parmod,pts", CASES, ids=[c[0] for c in CASES])
def test_field_mixed_close_to_oracle(cor, oracle, name, kinds, parmod, pts):
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=parmod)
    _, po = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    if parmod:
        for k, v in enumerate(parmod):
            po.PARMOD[k] = v
    got = cor.field(cor.model_set(**kinds), p, pts, cor.PREC_MIXED)
    want = oracle.field(L.models(**kinds), po, pts)
    r = np.linalg.norm(pts, axis=1) / RE
    scale = np.linalg.norm(want, axis=1) + 30574.0 / r ** 3
    err = np.linalg.norm(got - want, axis=1) / scale
    print(name, "mixed max rel err", err.max(), "p99", np.quantile(err, 0.99),
          "max abs nT", np.linalg.norm(got - want, axis=1).max())
    assert np.all(np.isfinite(got))
    assert err.max() < 1e-3
    assert np.quantile(err, 0.99) < 1e-4


@pytest.mark.parametrize("lanes", [64, 32, 16, 8, 4])
@pytest.mark.parametrize("name,kinds,parmod,pts", CASES, ids=[c[0] for c in CASES])
def test_field_wave_close_to_oracle(cor, oracle, name, kinds, parmod, pts, lanes):
    """The tracer's cooperative evaluator (one wavefront, or one group of 32, 16, 8 or 4
    lanes, per point; term families in SIMD passes; one cross-lane sum): same FP32 bar
    as the per-lane mixed path.  With smaller groups, the points of a wavefront are
    different points, so the groups follow different model branches."""
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=parmod)
    _, po = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    if parmod:
        for k, v in enumerate(parmod):
            po.PARMOD[k] = v
    prec = {64: cor.PREC_MIXED_WAVE, 32: cor.PREC_MIXED_GROUP32, 16: cor.PREC_MIXED_GROUP16,
            8: cor.PREC_MIXED_GROUP, 4: cor.PREC_MIXED_GROUP4}[lanes]
    got = cor.field(cor.model_set(**kinds), p, pts, prec)
    want = oracle.field(L.models(**kinds), po, pts)
    r = np.linalg.norm(pts, axis=1) / RE
    scale = np.linalg.norm(want, axis=1) + 30574.0 / r ** 3
    err = np.linalg.norm(got - want, axis=1) / scale
    print(name, lanes, "wave max rel err", err.max(), "p99", np.quantile(err, 0.99))
    assert np.all(np.isfinite(got))
    assert err.max() < 1e-3
    assert np.quantile(err, 0.99) < 1e-4


@pytest.mark.parametrize("name,kinds,parmod,pts", CASES, ids=[c[0] for c in CASES])
def test_cooperative_field_same_bits_in_every_group_width(cor, name, kinds, parmod, pts):
    """The cooperative evaluator sums its terms in one canonical order (term t in slot
    brev6(t) of 64, slots summed pairwise in slot order) whatever the number of lanes
    per point, so 64, 32, 16, 8 and 4 lanes give bit-identical fields -- the property that
    makes a trajectory's results independent of how many trajectories shared its
    wavefront.  The points include the magnetopause blend, the Region-1 band edges and
    the Region-2 xi bands (the box spans the T96 domain)."""
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=parmod)
    ms = cor.model_set(**kinds)
    outs = {lanes: cor.field(ms, p, pts, prec) for lanes, prec in
            ((64, cor.PREC_MIXED_WAVE), (32, cor.PREC_MIXED_GROUP32), (16, cor.PREC_MIXED_GROUP16),
             (8, cor.PREC_MIXED_GROUP), (4, cor.PREC_MIXED_GROUP4))}
    for lanes in (32, 16, 8, 4):
        same = outs[lanes] == outs[64]
        print(name, lanes, "lanes: identical", same.all(axis=1).mean())
        assert same.all(), np.argwhere(~same.all(axis=1))[:5].ravel()


@pytest.mark.parametrize("name,kinds,parmod,pts", CASES, ids=[c[0] for c in CASES])
def test_cooperative_field_fp64(cor, oracle, name, kinds, parmod, pts):
    """The cooperative evaluator in FP64 (the FP64 tracer's cooperative tail): the same
    terms and canonical order as the mixed one, so 64 and 8 lanes per point give the
    same bits; against the oracle the error is FP64 rounding in another summation order
    (T96's shielding fits cancel by up to 1e5, hence the bar relative to |B| + the
    dipole scale: 1e-10)."""
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=parmod)
    _, po = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    if parmod:
        for k, v in enumerate(parmod):
            po.PARMOD[k] = v
    ms = cor.model_set(**kinds)
    w = cor.field(ms, p, pts, cor.PREC_FP64_WAVE)
    g = cor.field(ms, p, pts, cor.PREC_FP64_GROUP)
    assert np.array_equal(w, g)
    want = oracle.field(L.models(**kinds), po, pts)
    r = np.linalg.norm(pts, axis=1) / RE
    err = np.linalg.norm(w - want, axis=1) / (np.linalg.norm(want, axis=1) + 30574.0 / r ** 3)
    print(name, "fp64 cooperative max rel err", err.max())
    assert err.max() < 1e-10


def cone_axis_points():
    """Nightside points within 0.002 Re of the GSM y = 0 plane where TS05's Region-1/2
    cone coordinates (lib/ts05/birk_tot.c:252-377) put the point within ~3e-4 rad of the
    cone's axis: three points where a trajectory of the 1-degree map's workload met it,
    and for each a line of 401 points across y = 0."""
    base = [(-7.454434, -15.121027), (-4.367026, -9.75418), (-3.81447, -8.748972)]
    ys = np.linspace(-0.002, 0.002, 401)
    return np.array([(x, y, z) for (x, z) in base for y in ys]) * RE


TS05_CASES = [
    ("ts05", dict(internal=0, external=2), [2, -20, 0, -2, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5], box_points(6, 4000)),
    ("ts05 cone axis", dict(internal=0, external=2), [3, -100, 2, -5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5],
     cone_axis_points()),
    ("igrf+ts05 storm", dict(internal=1, external=2), [3, -100, 2, -5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5],
     box_points(7, 4000)),
    ("igrf+ts05 drivers", dict(internal=1, external=2), [8, -250, -4, 6, 1.5, 0.2, 1.9, 0.7, 1.1, 0.3],
     box_points(8, 3000)),
]


@pytest.mark.parametrize("name,kinds,parmod,pts", TS05_CASES, ids=[c[0] for c in TS05_CASES])
def test_field_ts05(cor, oracle, name, kinds, parmod, pts):
    """TS05 (lib/ts05/*.c).  FP64: the reference's central differences (steps 1e-6 and
    1e-4) divide the few-ulp OCML/glibc differences of the differenced potentials by
    the step, ~1e-11 of |B| typically (p99 bar 1e-9); where TS05 alone is weak against
    the potentials it differences the relative figure grows, so the worst point's bar
    is 1e-7 of |B|.  Mixed: the T96 bar; its derivatives are exact (dual numbers) and
    its cancelling harmonic sums FP64."""
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=parmod)
    _, po = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    for k, v in enumerate(parmod):
        po.PARMOD[k] = v
    want = oracle.field(L.models(**kinds), po, pts)
    got = cor.field(cor.model_set(**kinds), p, pts, cor.PREC_FP64)
    err = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
    print(name, "fp64 max rel err", err.max(), "p99", np.quantile(err, 0.99), "at |B|",
          np.linalg.norm(want[np.argmax(err)]))
    assert np.all(np.isfinite(got)) and err.max() < 1e-7 and np.quantile(err, 0.99) < 1e-9
    got = cor.field(cor.model_set(**kinds), p, pts, cor.PREC_MIXED)
    r = np.linalg.norm(pts, axis=1) / RE
    scale = np.linalg.norm(want, axis=1) + 30574.0 / r ** 3
    err = np.linalg.norm(got - want, axis=1) / scale
    print(name, "mixed max rel err", err.max(), "p99", np.quantile(err, 0.99))
    assert np.all(np.isfinite(got)) and err.max() < 1e-3 and np.quantile(err, 0.99) < 1e-4


def seeded_sh(seed, nm=21):
    """A degree-20 table (nm = 21 rows) with an IGRF-like spectrum, row n, column m, nT."""
    rng = np.random.default_rng(seed)
    g = np.zeros((nm, nm))
    h = np.zeros((nm, nm))
    for n in range(1, nm):
        sc = 30000.0 * 0.3 ** n
        for m in range(n + 1):
            g[n, m] = sc * rng.normal()
            h[n, m] = sc * rng.normal() if m else 0.0
    return g, h


@pytest.mark.parametrize("name,ext,parmod", [
    ("sh", 0, None),
    ("sh+ts05", 2, [3, -100, 2, -5, 0.5, 0.5, 0.5, 0.5, 0.5, 0.5]),
])
def test_field_sh(cor, oracle, name, ext, parmod):
    """Generic SH synthesis (lib/chaos/chaos.c:118-227) to degree 20, alone and with
    TS05 as in the 1-degree map configuration: FP64 within 1e-12 of |B| (1e-8 with
    TS05's differences), mixed at the FP32 bar."""
    g, h = seeded_sh(11)
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=parmod)
    _, po = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    for k, v in enumerate(parmod or []):
        po.PARMOD[k] = v
    pts = points(9, 3000, rmin=1.0, rmax=12.0)
    ms = cor.model_set(cor.INTERNAL_NONE, ext, sh=(g, h))
    want = oracle.field(L.models(internal=0, external=ext, crustal=1, chaos=(g, h)), po, pts)
    got = cor.field(ms, p, pts, cor.PREC_FP64)
    err = np.linalg.norm(got - want, axis=1) / np.linalg.norm(want, axis=1)
    print(name, "fp64 max rel err", err.max())
    assert np.all(np.isfinite(got)) and err.max() < (1e-12 if ext == 0 else 1e-8)
    got = cor.field(ms, p, pts, cor.PREC_MIXED)
    r = np.linalg.norm(pts, axis=1) / RE
    scale = np.linalg.norm(want, axis=1) + 30574.0 / r ** 3
    err = np.linalg.norm(got - want, axis=1) / scale
    print(name, "mixed max rel err", err.max(), "p99", np.quantile(err, 0.99))
    assert np.all(np.isfinite(got)) and err.max() < 1e-3 and np.quantile(err, 0.99) < 1e-4


def test_field_chaos716(cor, oracle):
    """The registry's CHAOS-7.16 crustal model (degree 185, lib/chaos/chaos.c:118-227
    with lib/chaos/coefs/chaos_7_16.h) near the surface: FP64 within 1e-10 of the
    crustal field alone (with a 1e-6 nT floor); IGRF + CHAOS-7.16 mixed at the FP32 field
    bar."""
    from test_host_lib import chaos716_tables
    _, g, h = chaos716_tables()
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0)
    _, po = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    pts = points(21, 300, rmin=1.0, rmax=1.5)
    want = oracle.field(L.models(internal=0, crustal=1, chaos=(g, h)), po, pts)
    got = cor.field(cor.model_set(cor.INTERNAL_NONE, sh=(g, h)), p, pts, cor.PREC_FP64)
    # relative to |B| with a 1e-6 nT floor (the degree-21+ field falls as r^-23 and is
    # about 1e-3 nT at 1.5 Re)
    err = np.linalg.norm(got - want, axis=1) / (np.linalg.norm(want, axis=1) + 1e-6)
    print("chaos7.16 fp64 max rel err", err.max(), "|B| median", np.median(np.linalg.norm(want, axis=1)))
    assert np.all(np.isfinite(got)) and err.max() < 1e-10
    want = oracle.field(L.models(internal=1, crustal=1, chaos=(g, h)), po, pts)
    got = cor.field(cor.model_set(cor.INTERNAL_IGRF, sh=(g, h)), p, pts, cor.PREC_MIXED)
    r = np.linalg.norm(pts, axis=1) / RE
    scale = np.linalg.norm(want, axis=1) + 30574.0 / r ** 3
    err = np.linalg.norm(got - want, axis=1) / scale
    print("igrf+chaos7.16 mixed max rel err", err.max(), "p99", np.quantile(err, 0.99))
    assert np.all(np.isfinite(got)) and err.max() < 1e-3 and np.quantile(err, 0.99) < 1e-4
