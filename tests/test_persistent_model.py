"""Host-thread model of the persistent batch form's queue/counter protocol
(tests/persistent_model.cpp, after fd3d_tsn_amd/csrc/batch.hip
k_batch_persistent), which the library refuses to launch.

With coherent counters every launch drains, no bounded wait runs out and
every item starts after all items of its member's earlier stages: the
item numbering, dependency targets and per-step reset cannot stall.  With
stale per-group copies (a polling group never sees the increments of other
groups) waits run out, but the bounded waits and the give-up flag still
drain each launch within about one wait bound.
"""
import os
import re
import subprocess
import tempfile

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def model():
    d = tempfile.mkdtemp(prefix="persistent_model_")
    exe = os.path.join(d, "persistent_model")
    subprocess.run(["g++", "-O2", "-std=c++17", "-pthread", "-o", exe, os.path.join(HERE, "persistent_model.cpp")],
                   check=True)
    return exe


def run(exe, *args):
    out = subprocess.run([exe, *map(str, args)], check=True, capture_output=True, text=True, timeout=120).stdout
    return {k: float(v) for k, v in re.findall(r"(\w+)=([\d.]+)", out)}


@pytest.mark.parametrize("members,workers,seed", [(5, 32, 1), (5, 8, 2), (5, 64, 3), (16, 32, 4), (64, 64, 5)])
def test_coherent_counters_drain(model, members, workers, seed):
    r = run(model, members, 19, workers, 4, seed, 250)
    assert r["drained"] == 1 and r["timeouts"] == 0 and r["violations"] == 0, r


def test_stale_counters_still_drain_with_bounded_waits(model):
    r = run(model, "--stale", 5, 19, 32, 2, 1, 20)
    assert r["drained"] == 1 and r["timeouts"] > 0, r
    assert r["max_launch_ms"] < 20 * 10, r
