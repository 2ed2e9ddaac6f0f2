"""Extract the CHAOS-7.16 spherical-harmonic coefficients (the crustal model
`chaos7.16` of the reference's registry, src/models.c:102-112) from the reference's
header into C data.

The reference keeps two [186][186] arrays, G_CHAOS_7_16 and H_CHAOS_7_16, row n,
column m, nT (lib/chaos/coefs/chaos_7_16.h, descriptor at :384-388, nm_max = 186);
entries with m > n are zero.  This script reads those numbers as data and writes the
triangle m <= n of each, row after row, for the host library, which expands them into
the [186][186] layout of the model set at load time.

    python cor-model_amd/csrc/host/gen_chaos_table.py [/root/reference] > cor-model_amd/csrc/host/chaos716_table.c
"""
import re
import sys
from pathlib import Path

NM = 186


def rows(text: str, name: str) -> list:
    m = re.search(r"\b" + re.escape(name) + r"\s*\[\s*\w*\s*\]\s*\[\s*\w*\s*\]\s*=\s*\{(.*?)\};", text, re.S)
    assert m, name
    out = []
    for body in re.findall(r"\{([^{}]*)\}", m.group(1)):
        vals = [v.strip() for v in body.split(",") if v.strip()]
        assert len(vals) == NM, (name, len(out), len(vals))
        out.append(vals)
    assert len(out) == NM, (name, len(out))
    return out


def main() -> None:
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    text = (ref / "lib" / "chaos" / "coefs" / "chaos_7_16.h").read_text()
    g, h = rows(text, "G_CHAOS_7_16"), rows(text, "H_CHAOS_7_16")
    for tab in (g, h):
        for n in range(NM):
            assert all(float(v) == 0.0 for v in tab[n][n + 1:]), n
    out = ["/* CHAOS-7.16 coefficients [nT] for the crustal model chaos7.16: rows n = 0..185,",
           " * columns m = 0..n, G then H (index n(n+1)/2 + m).  Data read from",
           " * lib/chaos/coefs/chaos_7_16.h by gen_chaos_table.py. */",
           '#include "cor_host.h"', "",
           "const double cor_chaos716_tri[2][COR_CHAOS716_NTRI] = {"]
    for name, tab in (("G", g), ("H", h)):
        out.append("    { /* %s */" % name)
        for n in range(NM):
            out.append("        " + ", ".join(tab[n][:n + 1]) + ",")
        out.append("    },")
    out.append("};")
    print("\n".join(out))


if __name__ == "__main__":
    main()
