"""Loading the golden fixtures of tests/golden/ (written by make_golden.py)."""
import hashlib
import json
import os

import numpy as np

from fd3d_tsn_amd import cases as C

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
OUTS = ("sliprateX", "sliprateZ", "shearstressX", "shearstressZ")
FINALS = ("ruptime", "rise", "sliptime", "slipX", "slipZ", "stressdropX", "stressdropZ")


def manifest():
    with open(os.path.join(GOLDEN, "manifest.json")) as f:
        return json.load(f)


def case_of(name):
    cfg = dict(manifest()[name]["case"])
    cfg["stations"] = [tuple(s) for s in cfg["stations"]]
    if cfg.get("dc_range") is not None:
        cfg["dc_range"] = tuple(cfg["dc_range"])
    return C.Case(**cfg)


def load(name):
    m = manifest()[name]
    path = os.path.join(GOLDEN, m["file"])
    assert hashlib.sha256(open(path, "rb").read()).hexdigest() == m["sha256"], "fixture corrupted"
    with np.load(path, allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def slab_hash(a32):
    return np.frombuffer(hashlib.sha256(np.ascontiguousarray(np.asarray(a32, np.float32).T).tobytes()).digest(),
                         dtype=np.uint8)


def first_hash_mismatch(series, hashes, o):
    """First step whose slab hash differs (None if all match)."""
    for t in range(series.shape[0]):
        if not np.array_equal(slab_hash(series[t]), hashes[t, o]):
            return t
    return None
