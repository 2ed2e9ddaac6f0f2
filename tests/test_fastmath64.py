"""The FP64 e^x and sin/cos of the mixed tracer's TS05 harmonic sums
(cor-model_amd/csrc/gpu/fastmath64.hpp), compiled for the host from the same header the
GPU build uses (fma, rint, floor and ldexp give the same bits on both), against numpy:
within 2 ulp over the arguments those sums see and well beyond, NaN and the infinities
propagated, the quadrant boundaries of sin/cos exact to rounding."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

HDR = Path(__file__).resolve().parent.parent / "cor-model_amd" / "csrc" / "gpu"

SRC = r"""
#include "fastmath64.hpp"
extern "C" void fm_exp(const double *x, double *y, long n) {
    for (long i = 0; i < n; i++) y[i] = cor::fm64::exp(x[i]);
}
extern "C" void fm_sincos(const double *x, double *s, double *c, long n) {
    for (long i = 0; i < n; i++) cor::fm64::sincos(x[i], s + i, c + i);
}
"""


@pytest.fixture(scope="module")
def fm(tmp_path_factory):
    d = tmp_path_factory.mktemp("fm64")
    (d / "fm.cpp").write_text(SRC)
    so = d / "fm.so"
    subprocess.run(["g++", "-O2", "-ffp-contract=off", "-fPIC", "-shared", f"-I{HDR}", str(d / "fm.cpp"), "-o",
                    str(so)], check=True)
    L = C.CDLL(str(so))
    D = np.ctypeslib.ndpointer(np.float64, flags="C")
    L.fm_exp.argtypes = [D, D, C.c_long]
    L.fm_sincos.argtypes = [D, D, D, C.c_long]

    def exp(x):
        x = np.ascontiguousarray(x, np.float64)
        y = np.empty_like(x)
        L.fm_exp(x, y, len(x))
        return y

    def sincos(x):
        x = np.ascontiguousarray(x, np.float64)
        s, c = np.empty_like(x), np.empty_like(x)
        L.fm_sincos(x, s, c, len(x))
        return s, c
    return exp, sincos


def ulps(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


def test_exp_within_two_ulp(fm):
    exp, _ = fm
    rng = np.random.default_rng(5)
    x = np.concatenate([rng.uniform(-700, 700, 100000), rng.uniform(-60, 60, 100000),
                        rng.uniform(-1e-6, 1e-6, 1000), np.arange(-700, 701) * np.log(2.0)])
    u = ulps(exp(x), np.exp(x))
    print("exp: max", u.max(), "ulp, mean", u.mean())
    assert u.max() <= 2.0


def test_exp_special_values(fm):
    exp, _ = fm
    y = exp(np.array([np.nan, np.inf, -np.inf, 800.0, -800.0, 0.0]))
    assert np.isnan(y[0]) and y[1] == np.inf and y[2] == 0.0 and y[3] == np.inf and y[4] == 0.0 and y[5] == 1.0


def test_sincos_within_two_ulp(fm):
    _, sincos = fm
    rng = np.random.default_rng(6)
    x = np.concatenate([rng.uniform(-1000, 1000, 100000), rng.uniform(-4, 4, 100000), rng.uniform(-1e-8, 1e-8, 1000)])
    s, c = sincos(x)
    for name, got, ref in (("sin", s, np.sin(x)), ("cos", c, np.cos(x))):
        big = np.abs(ref) > 1e-3
        u = ulps(got[big], ref[big])
        print(name, "max", u.max(), "ulp; max abs error", np.abs(got - ref).max())
        assert u.max() <= 2.0
        assert np.abs(got - ref).max() <= 2.3e-16


def test_sincos_quadrants_and_specials(fm):
    _, sincos = fm
    k = np.arange(-400, 401)
    x = k * (np.pi / 2)
    s, c = sincos(x)
    # at the multiples of pi/2 one of the pair is +-1 and the other the rounding of x
    assert np.abs(s - np.sin(x)).max() <= 1e-15 and np.abs(c - np.cos(x)).max() <= 1e-15
    s, c = sincos(np.array([np.nan, np.inf]))
    assert np.isnan(s).all() and np.isnan(c).all()
