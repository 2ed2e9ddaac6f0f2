"""Writes t96_derived.h: constants derived from the T96 shield-field coefficients for
the FP32 path of t96_dev.hpp -- the reciprocal y and z scales of the Cartesian
harmonic shields (lib/t96/t96.c:501-617, birk1shld :1030-1194, birk2shl :1196-1263) and
the decay rates sqrt(1/p^2 + 1/r^2) of every term, each formed once in FP64 here
instead of by a square root and divisions per evaluation.

    python3 gen_t96_derived.py            (rewrites t96_derived.h beside this file)

The coefficients are read from t96_dev.hpp's own tables (k_shl, k_b1s_py, k_b1s_pz) and
from BIRK2SHL_AI / BIRK2SHL_AK below (the two y and z scales of birk2shl, as its
source writes them)."""
import math
import re
from pathlib import Path

HERE = Path(__file__).resolve().parent
BIRK2SHL_AI = (13.85650567, 14.90554500)
BIRK2SHL_AK = (10.21914434, 10.09021632)


def table(src, name):
    m = re.search(r"COR_CTAB\(%s, (\d+),(.*?)\)\n" % name, src, re.S)
    body = re.sub(r"//[^\n]*", "", m.group(2))
    vals = [eval(v.replace("(double)", "").strip()) for v in body.split(",") if v.strip()]
    assert len(vals) == int(m.group(1)), name
    return vals


def derived():
    src = (HERE / "t96_dev.hpp").read_text()
    shl, py, pz = table(src, "k_shl"), table(src, "k_b1s_py"), table(src, "k_b1s_pz")
    rp, rk, sq = [], [], []
    for s in range(3):
        p, r = shl[15 * s + 9:15 * s + 12], shl[15 * s + 12:15 * s + 15]
        rp += [1.0 / v for v in p]
        rk += [1.0 / v for v in r]
        sq += [math.sqrt(1.0 / (a * a) + 1.0 / (b * b)) for a in p for b in r]
    b1rp, b1rk = [1.0 / v for v in py], [1.0 / v for v in pz]
    b1sq = [math.sqrt(a * a + b * b) for a in b1rp for b in b1rk]
    b2rp, b2rk = [1.0 / v for v in BIRK2SHL_AI], [1.0 / v for v in BIRK2SHL_AK]
    b2sq = [math.sqrt(1.0 / (a * a) + 1.0 / (b * b)) for a in BIRK2SHL_AI for b in BIRK2SHL_AK]
    return [("k_shl_rp", rp), ("k_shl_rk", rk), ("k_shl_sq", sq), ("k_b1s_rp", b1rp), ("k_b1s_rk", b1rk),
            ("k_b1s_sq", b1sq), ("k_b2s_rp", b2rp), ("k_b2s_rk", b2rk), ("k_b2s_sq", b2sq)]


def render():
    out = ["// t96_derived.h -- generated by gen_t96_derived.py from the T96 shield coefficients",
           "// (reciprocal scales and per-term decay rates for the FP32 path); do not edit.",
           "#pragma once", ""]
    for name, vals in derived():
        items = ["%.17g" % v for v in vals]
        out.append("COR_FTAB(%s, %d," % (name, len(vals)))
        line = "        "
        lines = []
        for it in items:
            if len(line) + len(it) + 2 > 100:
                lines.append(line.rstrip())
                line = "        "
            line += it + ", "
        lines.append(line.rstrip().rstrip(","))
        out.append("\n".join(lines) + ")")
    return "\n".join(out) + "\n"


if __name__ == "__main__":
    (HERE / "t96_derived.h").write_text(render())
