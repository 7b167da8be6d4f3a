"""Key-configuration screening: the two pipeline stages around
``OmplDriver.validate_motion_pairs``, in the reference pipeline's on-disk
formats.

- ``screen_keyconf`` (src/GP/pipeline/solve2.py:164-231) checks the motion
  between every pair (i, j), j <= i, of the predicted key configurations in
  ``KEYQ_OMPL`` except the first two (the initial and goal states,
  util.py:73), one task chunk per worker, and writes the visible pairs of a
  chunk to ``edge_batch-<chunk>.npz`` as ``EDGE_FROM`` / ``EDGE_TO``.
  The chunking is partt.py:10-35 on the flat list of lower-triangle pairs
  (solve2.py:78-94): a chunk is a contiguous run of pairs in row-major order.
- ``assemble_roots`` (solve2.py:234-300) joins the keys that see each other
  into clusters, keeps from each cluster its lowest-numbered key that is not
  already disentangled (UnitWorld::isDisentangled on the key translated to
  the unit world), and writes the first two keys plus those to
  ``KEYQ_OMPL``.  Clusters come in the order of their lowest-numbered key,
  which is the order the reference's DisjointSet.get_cluster (a dict over
  the keys in ascending order, disjoint_set.py:51-59) produces whatever the
  union order was.

The motion checks run on the GPU through ``OmplDriver`` (whole pair lists
per call, not one chunk per process), the disentanglement checks through
``UnitWorld``.  Clustering is connected components over the edge list.
"""
import os

import numpy as np

RDT_FOREST_INIT_AND_GOAL_RESERVATIONS = 2  # src/GP/pipeline/util.py:73
RAW_KEY_PRED_SCHEMES = ("ge", "nt", "nn")  # src/GP/pipeline/file_locations.py:14
EDGE_BATCH_FMT = "edge_batch-{}.npz"        # solve2.py:222

# pairs per OmplDriver call: 2 x 56 bytes of states each on the host and the device
MOTION_BATCH = 1 << 22


def load_keys(fn):
    """``KEYQ_OMPL`` of an .npz key file (float64 [n, 7] OMPL states)."""
    with np.load(fn, allow_pickle=False) as d:
        return np.ascontiguousarray(d["KEYQ_OMPL"], dtype=np.float64)


def guess_chunk_number(shape, max_chunk_number, minimal_chunk_size):
    """partt.guess_chunk_number (partt.py:22-26): as many chunks as the quota
    allows while each keeps at least minimal_chunk_size tasks, and at least 1."""
    ntask = int(np.prod(shape))
    return max(1, min(int(max_chunk_number), int(ntask / float(minimal_chunk_size))))


def chunk_bounds(ntask, total_chunks):
    """Boundaries [total_chunks + 1] of partt.get_task_partition over a flat
    task list (partt.py:10-20, 28-35): chunk k is tasks [b[k], b[k + 1]).

    The reference cuts at int() of a running float sum of ntask / total_chunks
    and gives the last chunk whatever the sum leaves; the same float sum is
    kept here, so the cuts are the reference's to the task.  It fails (an
    assertion) where it cannot make total_chunks chunks, which includes an
    empty task list; so does this, with ValueError."""
    ntask = int(ntask)
    total_chunks = int(total_chunks)
    if total_chunks < 1:
        raise ValueError("total_chunks must be at least 1")
    step = ntask / float(total_chunks)
    bounds = [0]
    acc = 0.0
    for _ in range(total_chunks):
        if not acc < ntask:
            raise ValueError(f"{ntask} tasks cannot be cut into {total_chunks} chunks")
        acc += step
        bounds.append(min(int(acc), ntask))
    bounds[-1] = ntask
    return np.asarray(bounds, dtype=np.int64)


def tril_pairs(begin, end):
    """Rows and columns of the lower-triangle pairs number begin .. end - 1 in
    np.tril_indices order (row-major, diagonal included): pair t is (i, j)
    with i (i + 1) / 2 <= t < (i + 1) (i + 2) / 2 and j = t - i (i + 1) / 2."""
    t = np.arange(int(begin), int(end), dtype=np.int64)
    i = np.floor((np.sqrt(8.0 * t.astype(np.float64) + 1.0) - 1.0) / 2.0).astype(np.int64)
    # the float square root can land one row off near row starts
    i -= (i * (i + 1) // 2 > t)
    i += ((i + 1) * (i + 2) // 2 <= t)
    return i, t - i * (i + 1) // 2


def partition_screening(nkeys, total_chunks, index):
    """(from, to) key indices of chunk ``index`` (solve2.py:78-94): pairs of
    the lower triangle over keys R .. nkeys - 1, R = 2, from >= to."""
    r = RDT_FOREST_INIT_AND_GOAL_RESERVATIONS
    n = int(nkeys) - r
    if n < 0:
        raise ValueError("fewer keys than the initial and goal reservations")
    ntask = n * (n + 1) // 2
    b = chunk_bounds(ntask, total_chunks)
    if not 0 <= index < total_chunks:
        raise IndexError(f"chunk {index} out of {total_chunks}")
    i, j = tril_pairs(b[index], b[index + 1])
    return i + r, j + r


def screen_motions(driver, keys, from_idx, to_idx, batch=MOTION_BATCH):
    """EDGE_FROM / EDGE_TO: the (from, to) pairs whose OMPL motion
    keys[from] -> keys[to] is valid (solve2.py:219-224), int64, in input
    order.  A motion whose sample count reaches 2^31 raises, as
    OmplDriver.validate_motion_pairs does."""
    keys = np.ascontiguousarray(keys, dtype=np.float64)
    fi = np.asarray(from_idx, dtype=np.int64).reshape(-1)
    ti = np.asarray(to_idx, dtype=np.int64).reshape(-1)
    if fi.shape != ti.shape:
        raise ValueError("from and to index lists differ in length")
    if len(fi) and (min(fi.min(), ti.min()) < 0 or max(fi.max(), ti.max()) >= len(keys)):
        raise IndexError("key index out of range")
    keep = []
    for s in range(0, len(fi), int(batch)):
        f = fi[s:s + batch]
        t = ti[s:s + batch]
        ok = driver.validate_motion_pairs(keys[f], keys[t])
        keep.append(s + np.flatnonzero(ok))
    sel = np.concatenate(keep) if keep else np.zeros(0, np.int64)
    return fi[sel], ti[sel]


def save_edge_batch(fn, edge_from, edge_to):
    np.savez_compressed(fn, EDGE_FROM=np.asarray(edge_from), EDGE_TO=np.asarray(edge_to))


def screen_keyconf_chunk(driver, keys, total_chunks, index, outdir=None):
    """One screen_keyconf worker (solve2.py:211-224): chunk ``index`` of
    ``total_chunks``; with outdir, also writes outdir/edge_batch-<index>.npz."""
    fi, ti = partition_screening(len(keys), total_chunks, index)
    ef, et = screen_motions(driver, keys, fi, ti)
    if outdir is not None:
        save_edge_batch(os.path.join(outdir, EDGE_BATCH_FMT.format(index)), ef, et)
    return ef, et


def screen_keyconf(driver, keys, outdir=None, total_chunks=1):
    """Every chunk of the screening, run here instead of as separate jobs;
    returns the concatenated EDGE_FROM / EDGE_TO (chunk order)."""
    efs, ets = [], []
    for k in range(int(total_chunks)):
        ef, et = screen_keyconf_chunk(driver, keys, total_chunks, k, outdir)
        efs.append(ef)
        ets.append(et)
    return np.concatenate(efs), np.concatenate(ets)


def load_edge_batches(fns):
    """Concatenated EDGE_FROM / EDGE_TO of edge_batch files (solve2.py:251-260)."""
    efs, ets = [np.zeros(0, np.int64)], [np.zeros(0, np.int64)]
    for fn in fns:
        with np.load(fn, allow_pickle=False) as d:
            efs.append(np.asarray(d["EDGE_FROM"], dtype=np.int64).reshape(-1))
            ets.append(np.asarray(d["EDGE_TO"], dtype=np.int64).reshape(-1))
    ef, et = np.concatenate(efs), np.concatenate(ets)
    if ef.shape != et.shape:
        raise ValueError("EDGE_FROM and EDGE_TO differ in length")
    return ef, et


def clusters(lo, hi, edge_from, edge_to):
    """Connected components of keys lo .. hi - 1 under the edges with both
    ends in range (solve2.py:270-280).  Returns (label [hi - lo], first
    [ncluster]): label[v - lo] numbers v's cluster in the order of the
    clusters' lowest keys, and first[c] is cluster c's lowest key."""
    from scipy.sparse import coo_matrix
    from scipy.sparse.csgraph import connected_components

    lo, hi = int(lo), int(hi)
    n = max(hi - lo, 0)
    if n == 0:
        return np.zeros(0, np.int64), np.zeros(0, np.int64)
    ef = np.asarray(edge_from, dtype=np.int64).reshape(-1)
    et = np.asarray(edge_to, dtype=np.int64).reshape(-1)
    inr = (ef >= lo) & (ef < hi) & (et >= lo) & (et < hi)
    a, b = ef[inr] - lo, et[inr] - lo
    g = coo_matrix((np.ones(len(a), np.int8), (a, b)), shape=(n, n))
    _, lab = connected_components(g, directed=True, connection="weak")
    # renumber clusters by their lowest member
    _, first_local, inv = np.unique(lab, return_index=True, return_inverse=True)
    order = np.argsort(first_local, kind="stable")
    rank = np.empty_like(order)
    rank[order] = np.arange(len(order))
    return rank[inv].astype(np.int64), first_local[order].astype(np.int64) + lo


def select_roots(label, disentangled, lo):
    """Per cluster (in label order) its lowest key that is not disentangled;
    a cluster whose keys are all disentangled contributes nothing
    (solve2.py:285-293)."""
    label = np.asarray(label, dtype=np.int64)
    cand = np.flatnonzero(~np.asarray(disentangled, dtype=bool))
    if len(cand) == 0:
        return np.zeros(0, np.int64)
    cl, pos = np.unique(label[cand], return_index=True)  # first = lowest candidate
    return cand[pos[np.argsort(cl, kind="stable")]] + int(lo)


def assemble_roots(uw, keys, edge_from, edge_to, rtup=None):
    """Screened roots of one scheme (solve2.py:262-299): keys 0 and 1, then
    the chosen key of each cluster of keys rtup[0] .. rtup[1] - 1.  ``uw`` is
    the puzzle's UnitWorld (UnitWorld.is_disentangled of the keys in unit
    coordinates, solve2.py:285-290)."""
    keys = np.ascontiguousarray(keys, dtype=np.float64)
    r = RDT_FOREST_INIT_AND_GOAL_RESERVATIONS
    lo, hi = (r, len(keys)) if rtup is None else (int(rtup[0]), int(rtup[1]))
    label, _ = clusters(lo, hi, edge_from, edge_to)
    if len(label):
        unit = uw.translate_ompl_to_unit(keys[lo:hi])
        dis = uw.is_disentangled_states(unit)
        chosen = select_roots(label, dis, lo)
    else:
        chosen = np.zeros(0, np.int64)
    return keys[np.concatenate([np.arange(r, dtype=np.int64), chosen])]


def scheme_ranges(nkeys, bases):
    """Key ranges of the prediction schemes and of their combination
    (solve2.py:264-267): BASES_WITH_END[k] .. BASES_WITH_END[k + 1] for the
    schemes in order, and 2 .. nkeys for 'cmb'."""
    d = {}
    for k, s in enumerate(RAW_KEY_PRED_SCHEMES):
        d[s] = (int(bases[k]), int(bases[k + 1]))
    d["cmb"] = (RDT_FOREST_INIT_AND_GOAL_RESERVATIONS, int(nkeys))
    return d
