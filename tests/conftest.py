import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


def case_names(prefix=""):
    names = sorted(f[:-4] for f in os.listdir(os.path.join(GOLDEN, "cases")) if f.endswith(".npz"))
    return [n for n in names if n.startswith(prefix)]


def load_case(name):
    c = dict(np.load(os.path.join(GOLDEN, "cases", name + ".npz"), allow_pickle=False))
    ds = dict(np.load(os.path.join(GOLDEN, "datasets", str(c["dataset"]).replace(".hdf5", ".npz")),
                      allow_pickle=False))
    return c, ds
