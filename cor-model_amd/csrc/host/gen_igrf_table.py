"""Extract the IGRF-14 Gauss-coefficient chain from the reference headers into C data.

The reference keeps one ``const double G<year>_<gen>[105]`` / ``H...`` array pair per
5-year epoch spread over lib/internal/coefs/igrf_{9..14}_coefs.h and strings them
together in ``_IGRF_14[]`` (lib/internal/coefs/igrf_14_coefs.h:87), with the
descriptor at igrf_14_coefs.h:88-95 (54 rows, 5-year step, 1900..2030,
extrapolation on).  This script reads those numbers as data and writes them as one
row-major table for the host-side epoch setup.

    python cor-model_amd/csrc/host/gen_igrf_table.py [/root/reference] > cor-model_amd/csrc/host/igrf14_table.c
"""
import re
import sys
from pathlib import Path


def load_arrays(coef_dir: Path) -> dict:
    arrays = {}
    pat = re.compile(r"const double (\w+)\[INTERNAL_COEFS_SIZE\]\s*=\s*\{([^}]*)\}")
    for hdr in sorted(coef_dir.glob("*_coefs.h")):
        for name, body in pat.findall(hdr.read_text()):
            vals = [v.strip() for v in body.split(",") if v.strip()]
            arrays[name] = vals
    return arrays


def chain(coef_dir: Path, var: str) -> list:
    text = (coef_dir / "igrf_14_coefs.h").read_text()
    m = re.search(r"const double \*" + re.escape(var) + r"\[\]\s*=\s*\{([^}]*)\}", text)
    return [v.strip() for v in m.group(1).split(",") if v.strip()]


def main() -> None:
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    coef_dir = ref / "lib" / "internal" / "coefs"
    arrays = load_arrays(coef_dir)
    names = chain(coef_dir, "_IGRF_14")
    assert len(names) == 54, len(names)
    out = []
    out.append("/* IGRF-14 Gauss-coefficient chain used by the host epoch setup (igrf14_recalc).")
    out.append(" * Rows alternate G,H per 5-year epoch 1900..2025; the last pair is the")
    out.append(" * 2025-2030 secular variation.  105 entries per row (n <= 13, index")
    out.append(" * n(n+1)/2 + m).  Data read from lib/internal/coefs/igrf_{9..14}_coefs.h")
    out.append(" * (chain order: igrf_14_coefs.h:87) by gen_igrf_table.py. */")
    out.append('#include "cor_host.h"')
    out.append("")
    out.append("const double cor_igrf14_rows[COR_IGRF14_NROWS][INTERNAL_COEFS_SIZE] = {")
    for nm in names:
        vals = arrays[nm]
        assert len(vals) == 105, (nm, len(vals))
        out.append("    /* %s */ {%s}," % (nm, ", ".join(vals)))
    out.append("};")
    print("\n".join(out))


if __name__ == "__main__":
    main()
