"""Write the Gauss-coefficient tables of every internal field model as C data.

The reference keeps one ``const double <row>[INTERNAL_COEFS_SIZE]`` array per epoch
and coefficient kind (G or H) in lib/internal/coefs/*_coefs.h, strings them into one
row chain per model (e.g. ``_IGRF_14[]``, igrf_14_coefs.h:87) and describes each chain
with an ``InternalModelCoefs`` record (rows, year step, first and last year,
secular-variation extrapolation; e.g. igrf_14_coefs.h:88-95).  This script reads those
numbers as data and writes them as one pool of distinct rows, one chain of row
pointers per model and the eight descriptors under the reference's names, so callers
of internal_recalc find the same symbols.

    python cor-model_amd/csrc/host/gen_internal_tables.py [/root/reference] \
        > cor-model_amd/csrc/host/internal_tables.c
"""
import re
import sys
from pathlib import Path

# descriptor symbol -> (header, chain variable), in the reference's registry order
# (src/models.c:110)
MODELS = [
    ("IGRF_14_COEFS", "igrf_14_coefs.h", "_IGRF_14"),
    ("IGRF_13_COEFS", "igrf_13_coefs.h", "_IGRF_13"),
    ("IGRF_12_COEFS", "igrf_12_coefs.h", "_IGRF_12"),
    ("IGRF_11_COEFS", "igrf_11_coefs.h", "_IGRF_11"),
    ("IGRF_10_COEFS", "igrf_10_coefs.h", "_IGRF_10"),
    ("IGRF_9_COEFS", "igrf_9_coefs.h", "_IGRF_9"),
    ("HIST_COEFS", "hist_coefs.h", "_HIST"),
    ("CALS10K_2_COEFS", "cals10k_2_coefs.h", "_CALS10K_2"),
]
FIELDS = ("coefs_size", "years_step", "min_year", "max_year", "extrapolate")


def load_rows(coef_dir: Path) -> dict:
    rows = {}
    pat = re.compile(r"const double (\w+)\[INTERNAL_COEFS_SIZE\]\s*=\s*\{([^}]*)\}")
    for hdr in sorted(coef_dir.glob("*_coefs.h")):
        for name, body in pat.findall(hdr.read_text()):
            vals = [v.strip() for v in body.split(",") if v.strip()]
            assert len(vals) == 105, (hdr.name, name, len(vals))
            assert rows.get(name, vals) == vals, name
            rows[name] = vals
    return rows


def load_model(coef_dir: Path, header: str, chain: str, symbol: str):
    text = (coef_dir / header).read_text()
    m = re.search(r"const double \*" + re.escape(chain) + r"\[\]\s*=\s*\{([^}]*)\}", text)
    names = [v.strip() for v in m.group(1).split(",") if v.strip()]
    d = re.search(r"const InternalModelCoefs " + re.escape(symbol) + r"\s*=\s*\{([^}]*)\}", text)
    desc = {}
    for item in d.group(1).split(","):
        if "=" in item:
            k, v = item.split("=")
            desc[k.strip().lstrip(".")] = v.strip()
    assert int(desc["coefs_size"]) == len(names), (symbol, desc, len(names))
    return names, desc


def main() -> None:
    ref = Path(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
    coef_dir = ref / "lib" / "internal" / "coefs"
    rows = load_rows(coef_dir)
    models = [(sym, *load_model(coef_dir, hdr, chain, sym)) for sym, hdr, chain in MODELS]
    pool = []
    index = {}
    for _, names, _ in models:
        for nm in names:
            if nm not in index:
                index[nm] = len(pool)
                pool.append(nm)
    out = [
        "/* Gauss-coefficient tables of the internal field models (IGRF-9..14, the",
        " * historical 0-2000 CE model and CALS10k.2): a pool of distinct G and H rows",
        " * (105 entries each: n <= 13, index n(n+1)/2 + m, nT), one row chain per model,",
        " * G and H alternating per epoch with the secular-variation pair last where the",
        " * model extrapolates, and the models' descriptors under the reference's names.",
        " * Data read from the reference headers lib/internal/coefs/ by gen_internal_tables.py. */",
        '#include "cor_host.h"',
        "",
        f"static const double pool[{len(pool)}][INTERNAL_COEFS_SIZE] = {{",
    ]
    for nm in pool:
        out.append("    /* %s */ {%s}," % (nm, ", ".join(rows[nm])))
    out.append("};")
    out.append("")
    for sym, names, desc in models:
        cname = "chain_" + sym.lower()
        out.append(f"static const double *{cname}[{len(names)}] = {{")
        for k in range(0, len(names), 6):
            out.append("    " + " ".join(f"pool[{index[n]}]," for n in names[k:k + 6]))
        out.append("};")
        vals = ", ".join(f".{f} = {desc[f]}" for f in FIELDS)
        out.append(f"const InternalModelCoefs {sym} = {{.coefs = {cname}, {vals}}};")
        out.append("")
    print("\n".join(out).rstrip())


if __name__ == "__main__":
    main()
