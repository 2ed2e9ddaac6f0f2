"""Field evaluators on the GPU, one batch per precision: 2^20 seeded GSM points between
1.2 and 8 Re, each batch run twice.  Kernel times come from a kernel trace of this
script, e.g.

    rocprofv3 --kernel-trace --stats --output-format csv -d OUT -o run -- \\
        python3 cor-model_amd/field_bench.py [--models ts05]

Model sets:
  t96 (default)  IGRF-14 2020.0 + T96 (Pdyn 2, Dst -20, By 0, Bz -2): FP64 and mixed one
                 point per lane (cor::field_batch_kernel<0|1, 1, 1, 0>), and the
                 cooperative evaluator with 64, 32, 16 and 8 lanes per point
                 (cor::field_coop_kernel<1, 1, 1, 0, GS>), in that order;
  ts05           IGRF-14 2020.0 + TS05 (Pdyn 3, Dst -100, By 2, Bz -5, W1..W6 0.5): FP64
                 and mixed one point per lane (cor::field_batch_kernel<0|1, 1, 2, 0>)."""
import argparse
import sys
from pathlib import Path

import numpy as np

sys.path.insert(0, str(Path(__file__).resolve().parent))
import cor_amd as A  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--models", choices=("t96", "ts05"), default="t96")
    args = ap.parse_args()
    A.init(0)
    if args.models == "t96":
        p = A.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[2, -20, 0, -2])
        ms = A.model_set(A.INTERNAL_IGRF, A.EXTERNAL_T96)
        precs = (A.PREC_FP64, A.PREC_MIXED, A.PREC_MIXED_WAVE, A.PREC_MIXED_GROUP32, A.PREC_MIXED_GROUP16,
                 A.PREC_MIXED_GROUP)
    else:
        p = A.recalc(2020, 1, 1, 1, 0, 0, 0, parmod=[3, -100, 2, -5] + [0.5] * 6)
        ms = A.model_set(A.INTERNAL_IGRF, A.EXTERNAL_TS05)
        precs = (A.PREC_FP64, A.PREC_MIXED)
    rng = np.random.default_rng(1)
    n = 1 << 20
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pts = d * rng.uniform(1.2, 8.0, size=(n, 1)) * A.RE_M
    for prec in precs:
        for _ in range(2):
            A.field(ms, p, pts, prec)
        print("done precision", prec, flush=True)


if __name__ == "__main__":
    main()
