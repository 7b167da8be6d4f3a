"""Seeded test cases that need the oracle to pick them (test infrastructure:
only the tests import this module)."""
import numpy as np

from puzzletunneldiscovery_amd import workloads as W

from . import oracle as O


def pairs_with_midpoint_hits(ow, seed, n_free=300, n_starts=60_000):
    """Free-to-free pairs at SE(3) lengths 0.2-0.35 in ow's scene: every pair
    whose midpoint state (the first check of isValidTransition at vd = d/2,
    lib/osr/unit_world.cc:477-486) collides, then n_free whose midpoint is
    free.  Returns (starts, ends, number of midpoint hits)."""
    qs = W.random_states(n_starts, seed=seed)
    a = qs[ow.valid_states(qs).astype(bool)]
    b = W.segment_ends(a, seed=seed + 1, lengths=np.random.default_rng(seed + 2).uniform(0.2, 0.35, len(a)))
    keep = ow.valid_states(b).astype(bool)
    a, b = a[keep], b[keep]
    mids = []
    for i in range(len(a)):
        d = O.distance(a[i], b[i])
        mids.append(O.interpolate(a[i], b[i], (d / 2) * (1.0 / d)))
    hit = ~ow.valid_states(np.array(mids)).astype(bool)
    idx = np.concatenate([np.nonzero(hit)[0], np.nonzero(~hit)[0][:n_free]])
    return np.ascontiguousarray(a[idx]), np.ascontiguousarray(b[idx]), int(hit.sum())
