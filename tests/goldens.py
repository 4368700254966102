"""Loader for the golden fixtures in tests/golden/ (written by make_golden.py)."""
import glob
import json
import os

import numpy as np

from hyperflow_amd.case import Case

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    """The block-Cartesian goldens (uns_*.npz belong to the unstructured solver)."""
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(GOLDEN, "*.npz"))
                  if not os.path.basename(p).startswith("uns_"))


def load(name):
    d = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    case = Case.from_json(json.loads(str(d["case_json"])))
    nb = int(d["nblocks"])
    out = {"case": case, "init": [d[f"init_{b}"] for b in range(nb)], "steps": [int(s) for s in d["steps"]]}
    for n in out["steps"]:
        out[n] = {"fields": [d[f"after{n}_{b}"] for b in range(nb)], "calls": d[f"calls{n}"],
                  "stepdt": d[f"stepdt{n}"]}
    return out


def same(a, b):
    """Bitwise equality up to the sign of zero, NaN equal to NaN."""
    a = np.asarray(a)
    b = np.asarray(b)
    return a.shape == b.shape and bool(np.all((a == b) | (np.isnan(a) & np.isnan(b))))
