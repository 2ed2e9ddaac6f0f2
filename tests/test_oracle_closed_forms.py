"""Pins of the oracle that do not depend on the reference: closed forms, invariants and
independent algorithms.  Each check is chosen so that a plausible slip in the
restatement (a dropped term, a sign, a transposed index, a wrong normalisation)
fails it.
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy.special import factorial, lpmv

import oracle_lib as L

RE = 6.3712e6


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def schmidt_potential(n, m, g, h, x, y, z):
    """a (a/r)^(n+1) P_n^m(cos t) (g cos m f + h sin m f), Schmidt semi-normalised,
    no Condon-Shortley phase; a = 1 (positions in Earth radii)."""
    r = math.sqrt(x * x + y * y + z * z)
    t = math.acos(z / r)
    f = math.atan2(y, x)
    p = lpmv(m, n, math.cos(t)) * (-1) ** m
    if m > 0:
        p *= math.sqrt(2.0 * factorial(n - m) / factorial(n + m))
    return (1.0 / r) ** (n + 1) * p * (g * math.cos(m * f) + h * math.sin(m * f))


def fd_field(pot, x, y, z, eps=1e-5):
    """B = -grad V by central differences."""
    gx = (pot(x + eps, y, z) - pot(x - eps, y, z)) / (2 * eps)
    gy = (pot(x, y + eps, z) - pot(x, y - eps, z)) / (2 * eps)
    gz = (pot(x, y, z + eps) - pot(x, y, z - eps)) / (2 * eps)
    return -np.array([gx, gy, gz])


@pytest.mark.parametrize("n,m", [(1, 0), (1, 1), (2, 1), (3, 3), (5, 2), (8, 7), (13, 0)])
def test_sh_single_harmonic_matches_potential(oracle, n, m):
    """chaos_geo (lib/chaos/chaos.c:118-227) restated: one Schmidt harmonic against the
    finite-difference gradient of its potential computed with scipy's Legendre."""
    nm = 14
    g = np.zeros((nm, nm))
    h = np.zeros((nm, nm))
    g[n, m], h[n, m] = 1000.0, (700.0 if m else 0.0)
    rng = np.random.default_rng(n * 31 + m)
    for _ in range(6):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        p = d * rng.uniform(1.1, 4.0)
        got = oracle.chaos(g, h, p[None, :])[0]
        want = fd_field(lambda x, y, z: schmidt_potential(n, m, g[n, m], h[n, m], x, y, z), *p)
        assert np.linalg.norm(got - want) <= 1e-6 * np.linalg.norm(want) + 1e-9


def _schmidt_tables(oracle):
    """IGRF 2020.0 rows as [n][m] Schmidt tables (G2020_14 row of the chain)."""
    rows = np.ctypeslib.as_array(C.cast(oracle.lib.o_igrf14_rows, C.POINTER(C.c_double)),
                                 shape=(L.NCOEF * 54,)).reshape(54, L.NCOEF)
    return rows


def test_igrf_recursion_equals_independent_synthesis(oracle):
    """internal_geo's recursion form with internal_recalc's scaling against the
    standard three-term Schmidt synthesis on the unscaled 2020.0 coefficients."""
    rows = np.ctypeslib.as_array((C.c_double * (54 * L.NCOEF)).in_dll(oracle.lib, "o_igrf14_rows")).reshape(54, L.NCOEF)
    g20, h20 = rows[48], rows[49]          # G2020_14, H2020_14 (chain positions 48, 49)
    G = np.zeros((14, 14))
    H = np.zeros((14, 14))
    for n in range(1, 14):
        for m in range(n + 1):
            G[n, m] = g20[n * (n + 1) // 2 + m]
            H[n, m] = h20[n * (n + 1) // 2 + m]
    _, p = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    rng = np.random.default_rng(11)
    d = rng.normal(size=(300, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pts = d * rng.uniform(1.0, 1.99, size=(300, 1))   # below R = 2: full degree 13
    a = oracle.igrf_geo(p, pts)
    b = oracle.chaos(G, H, pts)
    err = np.linalg.norm(a - b, axis=1) / np.linalg.norm(b, axis=1)
    assert err.max() < 1e-11


def test_centred_dipole(oracle):
    """g10 alone: B_r = 2 g10 cos t / r^3, B_t = g10 sin t / r^3."""
    _, p = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    g10 = -29400.0
    for k in range(L.NCOEF):
        p.G[k] = 0.0
        p.H[k] = 0.0
    p.G[1] = g10
    rng = np.random.default_rng(3)
    for _ in range(50):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        x = d * rng.uniform(1.0, 20.0)
        b = oracle.igrf_geo(p, x[None, :])[0]
        r = np.linalg.norm(x)
        ct, st = x[2] / r, math.hypot(x[0], x[1]) / r
        br, bt = 2 * g10 * ct / r ** 3, g10 * st / r ** 3
        rho = math.hypot(x[0], x[1])
        cf, sf = x[0] / rho, x[1] / rho
        want = np.array([(br * st + bt * ct) * cf, (br * st + bt * ct) * sf, br * ct - bt * st])
        assert np.allclose(b, want, rtol=1e-12, atol=1e-9)


def test_igrf_is_solenoidal_and_irrotational(oracle):
    _, p = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    rng = np.random.default_rng(5)
    e = 1e-5
    for _ in range(20):
        d = rng.normal(size=3)
        d /= np.linalg.norm(d)
        x = d * rng.uniform(1.1, 1.9)
        J = np.zeros((3, 3))
        for j in range(3):
            dx = np.zeros(3)
            dx[j] = e
            J[:, j] = (oracle.igrf_geo(p, (x + dx)[None])[0] - oracle.igrf_geo(p, (x - dx)[None])[0]) / (2 * e)
        scale = np.linalg.norm(oracle.igrf_geo(p, x[None])[0]) / np.linalg.norm(x)
        assert abs(np.trace(J)) < 1e-6 * scale
        curl = np.array([J[2, 1] - J[1, 2], J[0, 2] - J[2, 0], J[1, 0] - J[0, 1]])
        assert np.linalg.norm(curl) < 1e-6 * scale


def test_gsm_frame(oracle):
    """A1..A3 orthonormal; the dipole axis lies in the GSM x-z plane (A2 . m = 0);
    A1 is the sun direction: declination ~0 at the 2020 March equinox (20 Mar 03:50 UT)
    and ~+23.44 deg at the June solstice (20 Jun 21:44 UT)."""
    for date in [(2020, 3, 20, 80, 3, 50, 0), (2020, 6, 20, 172, 21, 44, 0), (1975, 11, 2, 306, 7, 0, 0)]:
        _, p = oracle.recalc(*date)
        A = np.array([p.A1[:], p.A2[:], p.A3[:]])
        assert np.allclose(A @ A.T, np.eye(3), atol=1e-14)
        m = -np.array([p.G[2], p.H[2], p.G[1]])   # -(g11, h11, g10): dipole axis in GEO
        assert abs(A[1] @ m) < 1e-12 * np.linalg.norm(m)
    out = np.zeros(4)
    oracle.lib.o_sun(2020, 80, 3, 50, 0, _dp(out))
    assert abs(math.degrees(out[3])) < 0.05
    oracle.lib.o_sun(2020, 172, 21, 44, 0, _dp(out))
    assert abs(math.degrees(out[3]) - 23.44) < 0.05


def test_rk_uniform_field_is_fifth_order(oracle):
    """src/trajectory.c:275-356 restated.  In a uniform field the velocity rotates about
    B at rate d|B|: the step's angle error falls as h^6 (order 5), |v| is kept to the
    same order, and the position increment is v (e^h - 1) -- the reference's stages
    integrate f(x) = v + x -- to order h^6."""
    B = np.array([0.3, -0.5, 2.0])
    d = 1.7
    v0 = np.array([1.0, 0.4, -0.2])
    errs = []
    for h in (0.2, 0.1, 0.05):
        vel = v0.copy()
        pos = np.zeros(3)
        oracle.lib.o_rk_step(h, d, _dp(B.copy()), _dp(vel), _dp(pos))
        w = d * np.linalg.norm(B)
        u = -B / np.linalg.norm(B)
        th = w * h   # dv/dt = d v x B: rotation by +w h about -B
        exact = v0 * math.cos(th) + np.cross(u, v0) * math.sin(th) + u * (u @ v0) * (1 - math.cos(th))
        errs.append(np.linalg.norm(vel - exact))
        assert abs(np.linalg.norm(vel) - np.linalg.norm(v0)) < 10 * h ** 6
        assert np.allclose(pos, v0 * math.expm1(h), atol=5 * h ** 6)
    assert errs[0] / errs[1] > 40 and errs[1] / errs[2] > 40


def test_penumbra_algebra(oracle):
    """src/trajectory.c:156-160: Rl = lowest allowed, Ru = highest forbidden + del,
    Rc = Ru - del * (allowed count in [Rl, Ru))."""
    sc = L.OScan(1.0, 0, 0, 0, 0, 1.0, 0.5, 10.0, -1, 100, 0)
    n = 19
    status = np.zeros(n, dtype=np.int32)
    status[:6] = 1            # 1.0 .. 3.5 forbidden (re-entry)
    status[8] = 2             # 5.0 forbidden (step limit)
    status[10] = 1            # 6.0 forbidden
    tr = np.zeros(n, dtype=L.OTRAJ_DTYPE)
    tr["status"] = status
    cut = np.zeros(3)
    oracle.lib.o_penumbra(C.byref(sc), tr.ctypes.data, n, _dp(cut))
    rl, ru = 4.0, 6.0 + 0.5
    allowed_in_band = np.sum(status[6:11] == 0)   # 4.0, 4.5, 5.5 allowed
    assert cut[0] == rl and cut[1] == ru
    assert abs(cut[2] - (ru - 0.5 * allowed_in_band)) < 1e-12
    # nothing allowed: Rl = rk; nothing forbidden: Ru = del
    tr["status"] = 1
    oracle.lib.o_penumbra(C.byref(sc), tr.ctypes.data, n, _dp(cut))
    assert cut[0] == 10.0 and cut[1] == 10.0 + 0.5
    tr["status"] = 0
    oracle.lib.o_penumbra(C.byref(sc), tr.ctypes.data, n, _dp(cut))
    assert cut[0] == 1.0 and cut[1] == 0.5 and abs(cut[2] - 1.0) < 1e-12


def test_stormer_vertical_cutoff_at_dipole_equator(oracle):
    """g10-only field, vertical incidence on the dipole equator at r0 = 1.003 Re: the
    Stormer cutoff is |g10| a c / (4 r0^2) (14.04 GV / r0^2 for |g10| = 29403 nT).  The
    traced effective cutoff must sit within 0.1 GV of it."""
    _, p = oracle.recalc(2020, 1, 1, 1, 0, 0, 0)
    g10 = p.G[1]
    for k in range(L.NCOEF):
        p.G[k] = 0.0
        p.H[k] = 0.0
    p.G[1] = g10
    r0 = 1.003
    stormer = abs(g10) * 1e-9 * 6.3712e6 * 2.99725e8 / 4 / r0 ** 2 / 1e9
    sc = L.OScan(r0, 0.0, 30.0, 0.0, 30.0, stormer - 1.0, 0.02, stormer + 1.0, -1.0, 400, 0)
    tr, cut = oracle.trace_scan(sc, L.models(1, 0), p, 200000)
    assert abs(cut[2] - stormer) < 0.1, (cut, stormer)
    # East-west effect: protons (back-traced as zn = -1) arriving from the west at
    # zenith 30 deg see a lower cutoff than from the east (Stormer: 0.81 and 1.37
    # times the vertical one).  A sign slip in the charge or the Lorentz force swaps them.
    lon = math.radians(30.0)
    up = np.array([math.cos(lon), math.sin(lon), 0.0])
    west = np.array([math.sin(lon), -math.cos(lon), 0.0])
    cuts = {}
    for name, side in (("west", west), ("east", -west)):
        d = math.cos(math.radians(30)) * up + math.sin(math.radians(30)) * side
        fi1 = math.degrees(math.atan2(d[1], d[0]))
        sc = L.OScan(r0, 0.0, 30.0, 0.0, fi1, 8.0, 0.1, 20.0, -1.0, 400, 0)
        _, c = oracle.trace_scan(sc, L.models(1, 0), p, 200000)
        cuts[name] = c[2]
    assert cuts["west"] < stormer - 1.0 < stormer + 1.0 < cuts["east"], (cuts, stormer)
