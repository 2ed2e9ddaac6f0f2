"""Oblique arrival directions, asymptotic directions and penumbra refinement on the GPU
against the oracle.

The reference takes the arrival direction as a GEO latitude/longitude (the1, fi1,
src/io.c:7-11, src/trajectory.c:39-43) and reports the asymptotic direction of every
allowed rigidity (src/trajectory.c:91-112); a finer scan over the penumbra is another
run with a smaller del.  Directions given as local zenith/azimuth are converted with
cor_direction_to_geo (FP64, host); refined scans come from cor_refine_scan.  Bars as
in test_gpu_trace: FP64 flags equal outside the union penumbra band, |dRc| within one
coarse step, asymptotic directions within 0.05 degree; mixed per-lane: flags equal
outside the union band, |dRc| within half that band (at least one coarse step),
directions within 0.5 degree.
"""
import numpy as np
import pytest

import oracle_lib as L
from test_gpu_trace import compare

pytestmark = pytest.mark.gpu

T96 = [2.0, -20.0, 0.0, -2.0]
# (geocentric radius [Re], geocentric lat, lon) of two seeded stations, and
# (zenith, azimuth) of three directions: vertical, 40 deg from the east, 70 deg from
# the south-west
STATIONS = [(1.00300, -23.4, 133.9), (1.00296, 38.6, 291.7)]
DIRS = [(0.0, 0.0), (40.0, 90.0), (70.0, 225.0)]


def direction_rows(cor, rig=0.5, dl=0.25, rk=20.0, nk1=100):
    rows = []
    for (r0, la, lo) in STATIONS:
        for (ze, az) in DIRS:
            t1, f1 = cor.direction_to_geo(la, lo, ze, az)
            rows.append((r0, la, lo, t1, f1, rig, dl, rk, -1.0, nk1))
    return cor.scans_array(rows)


def oscan(row):
    return L.OScan(*[float(row[k]) for k in ("r0", "the0", "fi0", "the1", "fi1", "rig", "del_", "rk", "zn")],
                   int(row["nk1"]), 0)


def params(cor, oracle):
    p = cor.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=T96)
    _, po = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    for k, v in enumerate(T96):
        po.PARMOD[k] = v
    return p, po


@pytest.mark.parametrize("precision", [0, 1], ids=["fp64", "mixed-per-lane"])
def test_oblique_directions_against_oracle(cor, oracle, precision):
    """Two stations x three directions, IGRF + T96, 0.5..20 GV in 0.25."""
    rows = direction_rows(cor)
    p, po = params(cor, oracle)
    ms = cor.model_set(cor.INTERNAL_IGRF, cor.EXTERNAL_T96)
    tg, cg = cor.trace(ms, p, rows, step_limit=200000, precision=precision, coop=-1)
    off = 0
    for k, row in enumerate(rows):
        n = cor.scan_count(row)
        tc, cc = oracle.trace_scan(oscan(row), L.models(internal=1, external=1), po, 200000)
        print("scan", k, "gpu", cg[k], "cpu", cc)
        if precision == 0:
            compare(row, tg[off:off + n], cg[k], tc, cc, rc_tol=0.25)
        else:
            # half the union penumbra band of the two runs (at least one coarse step):
            # inside the band single trajectories are chaotic, and a last-bit change of
            # the FP32 field moves the lowest allowed rigidity by whole steps
            band = max(cg[k]["ru"], cc[1]) - min(cg[k]["rl"], cc[0])
            compare(row, tg[off:off + n], cg[k], tc, cc, rc_tol=max(0.25, 0.5 * band), dir_tol=0.5,
                    dir_margin=1.0, exit_tol=5.0)
        off += n
    # oblique directions see other cutoffs than the vertical one at the same station
    assert len({round(float(c), 6) for c in cg["rc"][:3]}) > 1


def test_direction_conversion_closed_forms(cor):
    """Zenith 0 is the station's own direction; zenith 90 towards azimuth 0 from the
    equator points at the north pole; azimuth 90 on the equator points east."""
    assert np.allclose(cor.direction_to_geo(37.0, 200.0, 0.0, 123.0), (37.0, 200.0))
    t, f = cor.direction_to_geo(0.0, 10.0, 90.0, 0.0)
    assert np.isclose(t, 90.0)
    t, f = cor.direction_to_geo(0.0, 10.0, 90.0, 90.0)
    assert np.isclose(t, 0.0, atol=1e-12) and np.isclose(f, 100.0)


def test_refined_penumbra_against_oracle(cor, oracle):
    """The IGRF-only site at 47N 8E: coarse scan 0.01..20 in 0.01, then [Rl - 0.01, Ru]
    in 0.001 on both sides (the same refined scan given to the oracle)."""
    r0, lat = cor.geodetic_to_geocentric(47.0, 20.0)
    coarse = cor.scans_array([(r0, lat, 8.0, lat, 8.0, 0.01, 0.01, 20.0, -1.0, 100)])
    p, po = params(cor, oracle)
    ms = cor.model_set(cor.INTERNAL_IGRF)
    _, cg = cor.trace(ms, p, coarse, step_limit=1000000)
    fine = cor.refine_scans(coarse, cg, 10)
    n = cor.scan_count(fine[0])
    assert n == int(round((fine[0]["rk"] - fine[0]["rig"]) / 0.001)) + 1
    tg, cf = cor.trace(ms, p, fine, step_limit=1000000)
    tc, cc = oracle.trace_scan(oscan(fine[0]), L.models(internal=1), po, 1000000)
    print("coarse", cg[0], "refined gpu", cf[0], "cpu", cc)
    compare(fine[0], tg, cf[0], tc, cc, rc_tol=0.01)
    # the refined band lies within the refined range (Ru is the last forbidden
    # rigidity plus one fine step)
    assert cf[0]["rl"] >= fine[0]["rig"] - 1e-9 and cf[0]["ru"] <= fine[0]["rk"] + 0.001 + 1e-9
