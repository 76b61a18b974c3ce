"""Loader of the golden fixtures in tests/golden (not a test module)."""
import glob
import json
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def names():
    return sorted(os.path.basename(p)[:-4] for p in glob.glob(os.path.join(HERE, "*.npz")))


def load(name):
    d = np.load(os.path.join(HERE, name + ".npz"))  # allow_pickle stays False
    out = {k: d[k] for k in d.files if k != "settings"}
    out["settings"] = json.loads(str(d["settings"]))
    return out
