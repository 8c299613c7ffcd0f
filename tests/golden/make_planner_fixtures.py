"""Compact the planner measurements into the fixtures the planner tests read.

The measurements are made on an MI355X by caps_amd.planner_data (one npz
per part, written to its --out path, OUT/ below); this script concatenates
the parts of a split and keeps the columns with small dtypes:

    python tests/golden/make_planner_fixtures.py train.npz OUT/planner_train_1.npz
    python tests/golden/make_planner_fixtures.py test.npz OUT/planner_test_0.npz ...
    python tests/golden/make_planner_fixtures.py planner_runs_batch_train.npz OUT/planner_batch_train.npz

Rows: one per (shape, candidate): shape id, n, layout depth, exponent
spread, the candidate's feature vector, measured ms, normwise error against
the 16-modulus reference product, mode, moduli, cutoff and pattern.
"""
import sys

import numpy as np


def main(argv):
    if len(argv) < 2:
        print(__doc__)
        return 2
    dst, srcs = argv[0], argv[1:]
    parts = [dict(np.load(p)) for p in srcs]
    z = {k: np.concatenate([p[k] for p in parts]) for k in parts[0]}
    out = {"shape": z["shape"].astype(np.int32), "n": z["n"].astype(np.int32),
           "nrec": z["nrec"].astype(np.int8), "spread": z["spread"].astype(np.int8),
           "features": np.asarray(z["features"], np.float32), "ms": z["ms"].astype(np.float64),
           "err": z["err"].astype(np.float64), "mode": z["mode"].astype(np.int8),
           "moduli": z["moduli"].astype(np.int8), "cutoff": z["cutoff"].astype(np.int8),
           "pattern": z["pattern"].astype("U8")}
    np.savez_compressed(dst, **out)
    print(dst, len(out["ms"]), "rows", len(np.unique(out["shape"])), "shapes")
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv[1:]))
