"""Loader for the committed golden fixtures (written by tests/golden/make_golden.py)."""
from __future__ import annotations

import json
import os

import numpy as np

DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fixtures")


class Fixture:
    def __init__(self, name: str):
        self.name = name
        with np.load(os.path.join(DIR, name + ".npz"), allow_pickle=False) as z:
            self.params = json.loads(str(z["params"]))
            self.meta = json.loads(str(z["meta"]))
            self.inp = {k[3:]: z[k] for k in z.files if k.startswith("in_")}
            self.out = {k[4:]: z[k] for k in z.files if k.startswith("out_")}
            self.arrays = {k[5:]: z[k] for k in z.files if k.startswith("meta_")}


def names(prefix: str = ""):
    return sorted(f[:-4] for f in os.listdir(DIR) if f.endswith(".npz") and f.startswith(prefix))


def load(name: str) -> Fixture:
    return Fixture(name)
