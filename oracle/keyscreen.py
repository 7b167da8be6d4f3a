"""CPU restatement of the key-configuration screening stages of the
reference pipeline, against which puzzletunneldiscovery_amd.keyscreen is
tested.

Test infrastructure only (tests/ use it); the shipped package never imports
it.  It follows the reference's own procedure step by step rather than the
product's vectorised one:

- task chunking: partt.chunk_it / get_task_partition over the list of task
  tuples (src/GP/pipeline/partt.py:10-35): slices at int() of a running float
  sum, the remainder appended to the last chunk;
- the pair list: np.tril_indices over the keys after the two reservations
  (src/GP/pipeline/solve2.py:78-94);
- the motion check per pair through a caller-given predicate (the oracle's
  OMPL motion validator, oracle.c);
- root assembly: a dict-based disjoint set with the reference's union rule
  (the second element's root is attached under the first's,
  src/GP/disjoint_set.py:17-49), clusters collected by walking the keys in
  ascending order (get_cluster, disjoint_set.py:51-59), and per cluster its
  members scanned in order until the first one that is not disentangled
  (solve2.py:283-297).
"""
import numpy as np

RESERVED = 2  # src/GP/pipeline/util.py:73


def chunk_list(seq, num):
    """partt.chunk_it (partt.py:10-20)."""
    step = len(seq) / float(num)
    out = []
    at = 0.0
    while at < len(seq) and len(out) < num:
        out.append(list(seq[int(at):int(at + step)]))
        at += step
    if not out or len(out) != num:
        raise ValueError("cannot cut into that many chunks")
    out[-1] += list(seq[int(at):])
    return out


def screening_chunk(nkeys, total_chunks, index):
    """(from, to) of chunk ``index`` (solve2.py:78-94)."""
    ti = np.tril_indices(nkeys - RESERVED)
    tasks = [e for e in np.ndindex(ti[0].shape)]
    chunk = np.array(chunk_list(tasks, total_chunks)[index], dtype=np.int64).reshape(-1)
    return ti[0][chunk] + RESERVED, ti[1][chunk] + RESERVED


class DisjointSet:
    """disjoint_set.DisjointSet's find / union / get_cluster semantics."""

    def __init__(self, items):
        self.parent = {v: v for v in items}

    def find(self, v):
        path = []
        while self.parent[v] != v:
            path.append(v)
            v = self.parent[v]
        for p in path:
            self.parent[p] = v
        return v

    def union(self, a, b):
        if a == b:
            return
        ra, rb = self.find(a), self.find(b)
        self.parent[rb] = ra

    def clusters(self):
        out = {}
        for v in self.parent:
            out.setdefault(self.find(v), []).append(v)
        return out


def assemble_roots(keys, edge_from, edge_to, lo, hi, is_disentangled):
    """Indices of the screened roots (solve2.py:262-297): 0, 1, then per
    cluster of keys lo .. hi - 1 its first member for which
    is_disentangled(index) is false."""
    ds = DisjointSet(range(lo, hi))
    for r, c in zip(np.asarray(edge_from).tolist(), np.asarray(edge_to).tolist()):
        if lo <= r < hi and lo <= c < hi:
            ds.union(r, c)
    chosen = []
    for members in ds.clusters().values():
        for m in members:
            if not is_disentangled(m):
                chosen.append(m)
                break
    return list(range(RESERVED)) + chosen
