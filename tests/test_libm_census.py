"""The libm census on a sample (tests/libm_census.py runs it on the full
workloads): the engine's split, faster sine for the poses a traversal only
tests and correctly rounded values for everything returned (oracle mode 2),
changes no output of the touch search or of uneven segments, and the host
C library (mode 1) changes no decision and moves returned taus by a few
ulp at most."""
import pytest

import libm_census as C
from puzzletunneldiscovery_amd import workloads as W


@pytest.fixture(scope="module")
def torus(oracle_mod, torus_meshes):
    env, rob = torus_meshes
    return oracle_mod.OracleWorld().setup(*env, *rob)


def test_touch_sample(oracle_mod, torus):
    starts = C.free_states(torus, 4096, 1000, 0)
    ends = W.segment_ends(starts, seed=2000)
    r = C.run(torus, lambda: torus.touch_batch(starts, ends, C.VD, 20), [0, 1, 2])
    for key, d in r["vs_mode_0"]["2"].items():
        assert d["differ"] == 0, key
    m1 = r["vs_mode_0"]["1"]
    assert m1["done"]["differ"] == 0 and m1["queries"]["differ"] == 0
    for key in ("last_tau", "tau", "lo", "hi"):
        assert m1[key]["max_ulp"] <= 4, key


def test_uneven_sample(oracle_mod, torus):
    starts = C.free_states(torus, 2048, 5000, 0)
    steps = W.lognormal_steps(len(starts), seed=5001, hi=1024)
    ends = W.segment_ends(starts, seed=5002, lengths=steps * C.VD)
    r = C.run(torus, lambda: torus.touch_batch(starts, ends, C.VD, 0), [0, 2])
    for key, d in r["vs_mode_0"]["2"].items():
        assert d["differ"] == 0, key
