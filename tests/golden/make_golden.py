"""Writes tests/golden/torus_small.npz: inputs and oracle outputs for the
collision-query path on the seeded two-torus scene, small enough to commit.

    python tests/golden/make_golden.py

The meshes are not stored; they are regenerated from their seeds by
puzzletunneldiscovery_amd.workloads.torus_scene(0).  The oracle runs with the
correctly rounded transcendental functions (se3_crmath.h, the values the
reference platform's C library returns; the GPU uses the same)."""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from puzzletunneldiscovery_amd import workloads as W  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "torus_small.npz")


def main():
    O.set_libm(False)
    env, rob = W.torus_scene(0)
    ow = O.OracleWorld().setup(*env, *rob)
    states = W.random_states(2048, seed=1001)
    valid = ow.valid_states(states)
    free = states[valid.astype(bool)]
    seg_from = free[:256]
    seg_to = W.segment_ends(seg_from, seed=1002, lmin=0.3, lmax=2.0)
    vd, levels = 1.0 / 1024, 20
    r = ow.touch_batch(seg_from, seg_to, vd, levels)
    lf = np.zeros((256, 7))
    st = np.zeros((256, 7))
    steps = np.zeros(256, np.int64)
    for i in range(256):
        lf[i], st[i], _, _, _, steps[i] = ow.transit(seg_from[i], seg_to[i], vd)
    vis = free[256:304]
    vis_vd = 1.0 / 256
    M = ow.visibility_matrix(vis, True, vis_vd)
    np.savez_compressed(
        OUT, states=states, valid=valid, seg_from=seg_from, seg_to=seg_to, vd=vd, levels=levels,
        done=r["done"], last_tau=r["last_tau"], tau=r["tau"], lo=r["lo"], hi=r["hi"],
        queries=r["queries"], last_free=lf, state=st, steps=steps, vis_states=vis, vis_vd=vis_vd, vis=M)
    print(OUT, os.path.getsize(OUT), "bytes;", int(valid.sum()), "free states,",
          int((~r["done"].astype(bool)).sum()), "segments hit")


if __name__ == "__main__":
    main()
