"""Pins the oracle's triangle-triangle test (FCL's 17-axis separating-axis
test, called from lib/osr/cdmodel.cc:171-173) three ways:

  * against the same 17 axes evaluated in exact rational arithmetic, on
    triangles with small dyadic coordinates (where binary64 is exact, so the
    two must agree bit for bit) and on random binary64 triangles;
  * against an independent exact predicate (closed triangles intersect iff
    an edge of one meets the other), which a dropped axis, a wrong edge or a
    sign error in the SAT would contradict;
  * against the known answer of sancheck/fclcontacts/main.cc:17-64.
"""
from fractions import Fraction as Fr

import numpy as np


def sub(a, b):
    return [a[0] - b[0], a[1] - b[1], a[2] - b[2]]


def cross(a, b):
    return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]


def dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def sat17_exact(P, Q):
    """The 17 axes of FCL's intersect_Triangle, in exact arithmetic."""
    p = [sub(P[i], P[0]) for i in range(3)]
    q = [sub(Q[i], P[0]) for i in range(3)]
    e = [sub(p[1], p[0]), sub(p[2], p[1]), sub(p[0], p[2])]
    f = [sub(q[1], q[0]), sub(q[2], q[1]), sub(q[0], q[2])]
    n1 = cross(e[0], e[1])
    m1 = cross(f[0], f[1])
    axes = [n1, m1] + [cross(e[i], f[j]) for i in range(3) for j in range(3)]
    axes += [cross(e[i], n1) for i in range(3)] + [cross(f[j], m1) for j in range(3)]
    for ax in axes:
        ps = [dot(ax, v) for v in p]
        qs = [dot(ax, v) for v in q]
        if min(ps) > max(qs) or min(qs) > max(ps):
            return False
    return True


def orient(a, b, c, d):
    return dot(sub(b, a), cross(sub(c, a), sub(d, a)))


def seg_tri_exact(p, q, a, b, c):
    """Closed segment pq meets closed triangle abc (exact)."""
    o1, o2 = orient(a, b, c, p), orient(a, b, c, q)
    if o1 == 0 and o2 == 0:
        return coplanar_seg_tri(p, q, a, b, c)
    if (o1 > 0 and o2 > 0) or (o1 < 0 and o2 < 0):
        return False
    s1, s2, s3 = orient(p, q, a, b), orient(p, q, b, c), orient(p, q, c, a)
    return (s1 >= 0 and s2 >= 0 and s3 >= 0) or (s1 <= 0 and s2 <= 0 and s3 <= 0)


def _proj(pts, n):
    k = max(range(3), key=lambda i: abs(n[i]))
    keep = [i for i in range(3) if i != k]
    return [(v[keep[0]], v[keep[1]]) for v in pts]


def _o2(a, b, c):
    return (b[0] - a[0]) * (c[1] - a[1]) - (b[1] - a[1]) * (c[0] - a[0])


def _on_seg(a, b, p):
    return min(a[0], b[0]) <= p[0] <= max(a[0], b[0]) and min(a[1], b[1]) <= p[1] <= max(a[1], b[1])


def _seg_seg2(a, b, c, d):
    d1, d2, d3, d4 = _o2(c, d, a), _o2(c, d, b), _o2(a, b, c), _o2(a, b, d)
    if ((d1 > 0 and d2 < 0) or (d1 < 0 and d2 > 0)) and ((d3 > 0 and d4 < 0) or (d3 < 0 and d4 > 0)):
        return True
    return ((d1 == 0 and _on_seg(c, d, a)) or (d2 == 0 and _on_seg(c, d, b)) or
            (d3 == 0 and _on_seg(a, b, c)) or (d4 == 0 and _on_seg(a, b, d)))


def _in_tri2(p, a, b, c):
    s1, s2, s3 = _o2(a, b, p), _o2(b, c, p), _o2(c, a, p)
    return (s1 >= 0 and s2 >= 0 and s3 >= 0) or (s1 <= 0 and s2 <= 0 and s3 <= 0)


def coplanar_seg_tri(p, q, a, b, c):
    n = cross(sub(b, a), sub(c, a))
    P, Q, A, B, C = _proj([p, q, a, b, c], n)
    if _in_tri2(P, A, B, C) or _in_tri2(Q, A, B, C):
        return True
    return _seg_seg2(P, Q, A, B) or _seg_seg2(P, Q, B, C) or _seg_seg2(P, Q, C, A)


def tris_intersect_exact(P, Q):
    for i in range(3):
        if seg_tri_exact(P[i], P[(i + 1) % 3], *Q):
            return True
        if seg_tri_exact(Q[i], Q[(i + 1) % 3], *P):
            return True
    return False


def nondegenerate(T):
    return any(x != 0 for x in cross(sub(T[1], T[0]), sub(T[2], T[0])))


def as_frac(T):
    return [[Fr(float(x)) for x in v] for v in T]


def test_fclcontacts_known_answer(oracle_mod):
    # sancheck/fclcontacts/main.cc:17-64: vertical triangle placed by the
    # relative translation (0.49, 0.5, -0.5): "Collision (expect true)".
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    Q = np.array([[0, 0, 0], [1, 0, 0], [0, 0, 1]], float) + [0.49, 0.5, -0.5]
    assert oracle_mod.tri_intersect(P, Q)
    # moved 0.02 further along x the two are disjoint
    assert not oracle_mod.tri_intersect(P, Q + [0.02, 0, 0])


def test_dyadic_triangles_bit_exact(oracle_mod):
    """Small dyadic coordinates: every binary64 operation in the SAT is exact,
    so the oracle must equal the rational evaluation; and for non-degenerate
    triangles the SAT must equal the independent exact predicate."""
    rng = np.random.default_rng(11)
    hits = checked = 0
    for _ in range(3000):
        P = rng.integers(-4, 5, size=(3, 3)) / 4.0
        Q = rng.integers(-4, 5, size=(3, 3)) / 4.0
        got = oracle_mod.tri_intersect(P, Q)
        Pf, Qf = as_frac(P), as_frac(Q)
        assert got == sat17_exact(Pf, Qf)
        if nondegenerate(Pf) and nondegenerate(Qf):
            assert got == tris_intersect_exact(Pf, Qf), (P, Q)
            checked += 1
            hits += got
    assert checked > 2000 and 100 < hits < checked - 100


def test_touching_and_coplanar_cases(oracle_mod):
    """Closed-interval semantics: shared vertices, shared edges, touching
    faces and coplanar overlaps all count as intersecting."""
    A = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], float)
    cases = [
        (np.array([[0, 0, 0], [-1, 0, 0], [0, -1, 0]], float), True),   # shared vertex
        (np.array([[0, 0, 0], [1, 0, 0], [0, -1, 0]], float), True),    # shared edge
        (np.array([[0.2, 0.2, 0], [1, 1, 1], [0.5, 0, 1]], float), True),  # vertex on face
        (np.array([[0.1, 0.1, 0], [0.3, 0.1, 0], [0.1, 0.3, 0]], float), True),  # coplanar inside
        (np.array([[2, 0, 0], [3, 0, 0], [2, 1, 0]], float), False),    # coplanar apart
        (np.array([[0.2, 0.2, 1e-9], [1, 1, 1], [0.5, 0, 1]], float), False),  # just above
    ]
    for B, want in cases:
        assert oracle_mod.tri_intersect(A, B) == want, B
        assert oracle_mod.tri_intersect(B, A) == want, B
        assert tris_intersect_exact(as_frac(A), as_frac(B)) == want


def test_random_binary64_triangles(oracle_mod):
    """Random binary64 triangles: the rounded evaluation agrees with the
    exact evaluation of the same formula and with the exact predicate."""
    rng = np.random.default_rng(12)
    for _ in range(1500):
        c = rng.normal(size=3) * 0.05
        P = c + rng.normal(size=(3, 3)) * 0.1
        Q = c + rng.normal(size=(3, 3)) * 0.1
        got = oracle_mod.tri_intersect(P, Q)
        Pf, Qf = as_frac(P), as_frac(Q)
        assert got == sat17_exact(Pf, Qf)
        assert got == tris_intersect_exact(Pf, Qf)
