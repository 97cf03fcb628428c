"""Loader for the golden fixtures under tests/golden/ (JSON index + float64 .bin)."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load(name):
    with open(os.path.join(GOLDEN, name + ".json")) as f:
        js = json.load(f)
    raw = np.fromfile(os.path.join(GOLDEN, name + ".bin"), dtype="<f8")
    arrays = {}
    for a in js["arrays"]:
        n = int(np.prod(a["shape"]))
        arrays[a["name"]] = raw[a["offset"]:a["offset"] + n].reshape(a["shape"])
    return js["params"], arrays
