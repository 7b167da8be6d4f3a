"""Key-configuration clearance: the pipeline stage that writes
``DISTANCE_BATCH`` and the stage that reads it, on the GPU.

- ``estimate_keyconf_clearance`` (src/GP/pipeline/keyconf.py:276-317): for
  every oversampled key configuration (``KEYQ_OMPL``, OMPL convention), moved
  to the unit world, ``nsample`` touch queries toward random targets
  (touchq_util.py:8-27): the target is ``apply(q0, tr, aa)`` with ``tr`` a
  uniform point on the unit sphere and ``aa`` a uniform point in the ball of
  radius 2 pi (uw_random.py:20-44), and the query is
  transitStateToWithContact at the world's recommended_cres.  The stage
  stores, per key, the distance from q0 to each query's last free state
  (``multi_distance``) in ``clearance_batch-<task>.npz`` as
  ``DISTANCE_BATCH`` [keys of the task, nsample].
- ``least_visible_keyconf_fixed`` (solve2.py:104-125) keeps the K keys with
  the smallest mean squared distance, the first two (initial and goal
  states) excluded, and saves them as ``KEYQ_OMPL``.

The random targets come from numpy's legacy global stream (or a given
``RandomState``) in the reference's order: per sample, the sphere point's
rejection tries (2 draws each) and then the ball point's (3 draws each).
``touch_directions`` draws them in bulk, finds the accepted tries with array
operations, and rewinds the stream to exactly what the reference consumes, so
the same seed gives the same targets and leaves the stream where the
reference leaves it.  The touch queries of all keys go to the GPU in one
batch (``UnitWorld.transit_states_to_with_contact``).
"""
import itertools
import math
import os

import numpy as np

from . import pyosr

RDT_FOREST_INIT_AND_GOAL_RESERVATIONS = 2  # src/GP/pipeline/util.py:73
CLEARANCE_BATCH_FMT = "clearance_batch-{}.npz"  # keyconf.py:316
BALL_RADIUS = 2 * math.pi  # touchq_util.py:11

_NEVER = np.iinfo(np.int64).max


def _next_accepted(acc, period):
    """nxt[o]: the first o' >= o with o' = o (mod period) and acc[o'], or _NEVER."""
    n = len(acc)
    nxt = np.full(n + period, _NEVER, dtype=np.int64)
    for r in range(period):
        idx = np.arange(r, n, period, dtype=np.int64)
        if len(idx) == 0:
            continue
        v = np.where(acc[idx], idx, _NEVER)
        nxt[idx] = np.minimum.accumulate(v[::-1])[::-1]
    return nxt


def touch_directions(n, rs=None):
    """(tr [n, 3], aa [n, 3]): the translations and axis-angle rotations of n
    touch samples, drawn from ``rs`` (default: numpy's global legacy stream)
    exactly as uw_random.random_on_sphere(1.0) followed by
    random_within_sphere(2 pi) draws them per sample (uw_random.py:20-44):
    - sphere: x = 2 (u - 1/2) pairs until x1^2 + x2^2 < 1, then
      (2 x1 sqrt(1 - l2), 2 x2 sqrt(1 - l2), 1 - 2 l2) (Marsaglia, 1972);
    - ball: x = 2 (u - 1/2) triples until x1^2 + x2^2 + x3^2 <= 1, times 2 pi.
    """
    rs = np.random if rs is None else rs
    if not all(hasattr(rs, a) for a in ("get_state", "set_state", "random_sample")):
        raise TypeError("touch_directions draws from numpy's legacy stream: pass np.random or a "
                        "np.random.RandomState, as the reference's uw_random does")
    n = int(n)
    tr = np.empty((n, 3))
    aa = np.empty((n, 3))
    done = 0
    slack = 256
    while done < n:
        want = min(n - done, 1 << 16)
        state = rs.get_state()
        # 2 / (pi / 4) + 3 / (pi / 6) = 8.28 draws per sample on average
        m = 9 * want + slack
        x = 2.0 * (rs.random_sample(m) - 0.5)
        sq = x * x
        acc_s = np.zeros(m, dtype=bool)
        acc_s[:m - 1] = (sq[:m - 1] + sq[1:]) < 1
        acc_b = np.zeros(m, dtype=bool)
        acc_b[:m - 2] = ((sq[:m - 2] + sq[1:m - 1]) + sq[2:]) <= 1
        nxt_s = _next_accepted(acc_s, 2).tolist()
        nxt_b = _next_accepted(acc_b, 3).tolist()
        s_at, b_at = [], []
        o = 0
        while len(s_at) < want:
            s = nxt_s[o]
            if s > m - 2:
                break
            b = nxt_b[s + 2]
            if b > m - 3:
                break
            s_at.append(s)
            b_at.append(b)
            o = b + 3
        k = len(s_at)
        s_at = np.asarray(s_at, dtype=np.int64)
        b_at = np.asarray(b_at, dtype=np.int64)
        x1, x2 = x[s_at], x[s_at + 1]
        l2 = x1 * x1 + x2 * x2
        root = np.sqrt(1 - l2)
        tr[done:done + k, 0] = 2 * x1 * root
        tr[done:done + k, 1] = 2 * x2 * root
        tr[done:done + k, 2] = 1 - 2 * l2
        aa[done:done + k] = np.stack([x[b_at], x[b_at + 1], x[b_at + 2]], axis=1) * BALL_RADIUS
        # rewind, then consume exactly what the accepted samples used
        rs.set_state(state)
        if k == 0:
            slack *= 2  # too few draws for even one sample: draw more
            continue
        rs.random_sample(o)
        done += k
    return tr, aa


def touch_distances(uw, unit_keys, tr, aa):
    """DISTANCE_BATCH rows for unit-world keys [nkey, 7] and per-key sample
    directions tr, aa [nkey, nsample, 3]: multi_distance from each key to
    the last free state of transitStateToWithContact toward
    apply(key, tr, aa) at uw.recommended_cres (keyconf.py:309-315)."""
    q = np.ascontiguousarray(unit_keys, dtype=np.float64).reshape(-1, 7)
    nkey = len(q)
    tr = np.asarray(tr, dtype=np.float64).reshape(nkey, -1, 3)
    aa = np.asarray(aa, dtype=np.float64).reshape(nkey, -1, 3)
    ns = tr.shape[1]
    if nkey == 0 or ns == 0:
        return np.zeros((nkey, ns))
    frm = np.repeat(q, ns, axis=0)
    to = pyosr.apply_batch(frm, tr.reshape(-1, 3), aa.reshape(-1, 3))
    free = uw.transit_states_to_with_contact(frm, to, uw.recommended_cres)[0]
    out = np.empty((nkey, ns))
    for i in range(nkey):
        out[i] = pyosr.multi_distance(q[i], free[i * ns:(i + 1) * ns])
    return out


def estimate_clearance(uw, keys_ompl, nsample, rs=None):
    """DISTANCE_BATCH [nkey, nsample] for OMPL-convention keys, drawing the
    targets key by key in the reference's order (keyconf.py:309-315)."""
    keys = np.ascontiguousarray(keys_ompl, dtype=np.float64).reshape(-1, 7)
    tr, aa = touch_directions(len(keys) * int(nsample), rs)
    unit = uw.translate_ompl_to_unit(keys) if len(keys) else np.zeros((0, 7))
    return touch_distances(uw, unit, tr, aa)


def clearance_task(nkeys, total_chunks, task_id):
    """Key indices of one estimate_keyconf_clearance task
    (keyconf.py:284-305): partt's chunks of the keys, task 0 when there is a
    single chunk."""
    from .keyscreen import chunk_bounds
    task_id = 0 if total_chunks == 1 else int(task_id)
    b = chunk_bounds(int(nkeys), total_chunks)
    return np.arange(b[task_id], b[task_id + 1], dtype=np.int64)


def estimate_keyconf_clearance_task(uw, keys_ompl, nsample, total_chunks, task_id, outdir=None, rs=None):
    """One estimate_keyconf_clearance worker; with outdir, also writes
    outdir/clearance_batch-<task>.npz (keyconf.py:303-317)."""
    idx = clearance_task(len(keys_ompl), total_chunks, task_id)
    d = estimate_clearance(uw, np.asarray(keys_ompl)[idx], nsample, rs)
    if outdir is not None:
        tid = 0 if total_chunks == 1 else int(task_id)
        # the reference saves a list of rows: an empty task stores shape (0,)
        np.savez(os.path.join(outdir, CLEARANCE_BATCH_FMT.format(tid)),
                 DISTANCE_BATCH=d if len(d) else np.zeros(0))
    return d


def load_distance_batches(indir):
    """DISTANCE_BATCH of clearance_batch-0.npz, -1, ... up to the first
    missing file, concatenated without the empty ones (util.lsv2 and
    safe_concatente, util.py:634-679)."""
    arrs = []
    for i in itertools.count():
        fn = os.path.join(indir, CLEARANCE_BATCH_FMT.format(i))
        if not os.path.exists(fn):
            if i == 0:
                raise FileNotFoundError(f"no {CLEARANCE_BATCH_FMT.format(0)} under {indir}")
            break
        with np.load(fn, allow_pickle=False) as d:
            a = np.asarray(d["DISTANCE_BATCH"])
        if a.shape[0] != 0:
            arrs.append(a)
    return np.concatenate(arrs, axis=0) if arrs else np.array([])


def least_visible_keys(oskey, distance_batch, k):
    """The K keys with the least mean squared clearance, the first two
    excluded, in that order (solve2.py:116-124)."""
    osc = np.asarray(distance_batch)
    if osc.shape[0] != len(oskey):
        raise ValueError(f"DISTANCE_BATCH rows {osc.shape[0]} != keys {len(oskey)}")
    r = RDT_FOREST_INIT_AND_GOAL_RESERVATIONS
    mean = np.mean(osc ** 2, axis=1)[r:]
    top = mean.argsort()[:int(k)] + r
    return np.asarray(oskey)[top, :]
