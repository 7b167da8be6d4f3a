"""Seeded synthetic scenes and query sets for the collision-query engine.

Meshes are returned as (V float32 [nv,3], F uint32 [nf,3]) in their raw
(pre-normalisation) coordinates; ``UnitWorld`` then applies the reference's
unit-world normalisation (lib/osr/unit_world.cc:123-161) exactly as it does to
meshes read from OBJ files.

Query generators follow the reference's callers:
  * states: uniform translation plus a Shoemake quaternion in the slot order of
    src/GP/pipeline/uw_random.py:8-17 (w, x, y, z = sqrt(1-u1) sin(2 pi u2),
    sqrt(1-u1) cos(2 pi u2), sqrt(u1) sin(2 pi u3), sqrt(u1) cos(2 pi u3)),
    rounded to binary32 as random_state does;
  * touch segments: a free start state and an end state reached by a
    translation plus an axis-angle rotation (src/GP/pipeline/touchq_util.py:8-13).
"""
import math

import numpy as np


def _rotation_from_quat(q):
    w, x, y, z = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
        [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
        [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)],
    ])


def _shoemake(rng, n):
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    return np.stack([
        np.sqrt(1 - u1) * np.sin(2 * np.pi * u2),
        np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
        np.sqrt(u1) * np.sin(2 * np.pi * u3),
        np.sqrt(u1) * np.cos(2 * np.pi * u3),
    ], axis=1)


def torus(R, r, nu, nv, center=(0.0, 0.0, 0.0), seed=None, noise=0.0):
    """Closed torus with 2*nu*nv triangles (nu around the ring, nv around the tube).

    With a seed, the torus gets a random orientation and radial jitter of
    ``noise * r`` per vertex.
    """
    th = 2 * np.pi * np.arange(nu) / nu
    ph = 2 * np.pi * np.arange(nv) / nv
    T, P = np.meshgrid(th, ph, indexing="ij")
    rr = np.full(T.shape, float(r))
    rot = np.eye(3)
    if seed is not None:
        rng = np.random.default_rng(seed)
        rr = rr * (1.0 + noise * (2 * rng.random(T.shape) - 1))
        rot = _rotation_from_quat(_shoemake(rng, 1)[0])
    X = (R + rr * np.cos(P)) * np.cos(T)
    Y = (R + rr * np.cos(P)) * np.sin(T)
    Z = rr * np.sin(P)
    V = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1) @ rot.T + np.asarray(center)
    i, j = np.meshgrid(np.arange(nu), np.arange(nv), indexing="ij")
    a = i * nv + j
    b = ((i + 1) % nu) * nv + j
    c = ((i + 1) % nu) * nv + (j + 1) % nv
    d = i * nv + (j + 1) % nv
    F = np.concatenate([np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)])
    return V.astype(np.float32), F.astype(np.uint32)


def torus_scene(seed=0):
    """Env torus R=0.3, r=0.08 (2,048 tris) centred in the unit box and a robot
    torus R=0.1, r=0.02 (1,024 tris)."""
    env = torus(0.3, 0.08, 32, 32, center=(0.5, 0.5, 0.5), seed=seed, noise=0.02)
    rob = torus(0.1, 0.02, 32, 16, seed=seed + 1, noise=0.02)
    return env, rob


def icosphere(level):
    t = (1.0 + 5 ** 0.5) / 2
    V = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t),
         (0, -1, -t), (0, 1, -t), (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    F = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4),
         (11, 10, 2), (10, 7, 6), (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8),
         (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10), (8, 6, 7), (9, 8, 1)]
    V = [np.array(v, float) / np.linalg.norm(v) for v in V]
    for _ in range(level):
        cache = {}
        nF = []

        def mid(a, b):
            key = (min(a, b), max(a, b))
            if key not in cache:
                m = V[a] + V[b]
                V.append(m / np.linalg.norm(m))
                cache[key] = len(V) - 1
            return cache[key]

        for a, b, c in F:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nF += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        F = nF
    return np.array(V), np.array(F, dtype=np.int64)


def shell_with_holes(level=6, r_out=0.45, r_in=0.40, n_holes=64, hole_r=0.05, seed=0,
                     noise=0.002, tube_seg=64, tube_rings=9):
    """Noisy icosphere shell (outer + inner surface) with cylindrical holes
    drilled radially through it; each hole is lined by a tube wall of
    tube_seg x (tube_rings - 1) quads.  The defaults give about 195k
    triangles."""
    rng = np.random.default_rng(seed)
    S, SF = icosphere(level)
    dirs = _shoemake(rng, n_holes)[:, 1:]
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    Vs, Fs, off = [], [], 0
    for radius, flip in ((r_out, False), (r_in, True)):
        P = S * radius
        P = P * (1.0 + noise * (2 * rng.random((len(P), 1)) - 1))
        cen = P[SF].mean(axis=1)
        cn = cen / np.linalg.norm(cen, axis=1, keepdims=True)
        keep = np.ones(len(SF), bool)
        for d in dirs:
            along = cen @ d
            perp = np.linalg.norm(cen - along[:, None] * d[None, :], axis=1)
            keep &= ~((along > 0) & (perp < hole_r))
        del cn
        F = SF[keep]
        if flip:
            F = F[:, ::-1]
        Vs.append(P)
        Fs.append(F + off)
        off += len(P)
    # tube walls
    ang = 2 * np.pi * np.arange(tube_seg) / tube_seg
    for d in dirs:
        u = np.cross(d, [1.0, 0, 0])
        if np.linalg.norm(u) < 0.1:
            u = np.cross(d, [0, 1.0, 0])
        u /= np.linalg.norm(u)
        v = np.cross(d, u)
        ring = np.cos(ang)[:, None] * u + np.sin(ang)[:, None] * v
        a0 = np.sqrt(max((r_in * 0.97) ** 2 - hole_r ** 2, 0.0))
        a1 = np.sqrt(max((r_out * 1.03) ** 2 - hole_r ** 2, 0.0))
        P = np.concatenate([along * d + hole_r * ring for along in np.linspace(a0, a1, tube_rings)])
        a = np.arange(tube_seg)
        b = (a + 1) % tube_seg
        quads = [np.concatenate([np.stack([a, b, b + tube_seg], 1), np.stack([a, b + tube_seg, a + tube_seg], 1)])
                 + k * tube_seg for k in range(tube_rings - 1)]
        F = np.concatenate(quads)
        Vs.append(P)
        Fs.append(F + off)
        off += len(P)
    V = np.concatenate(Vs) + 0.5
    F = np.concatenate(Fs)
    return V.astype(np.float32), F.astype(np.uint32)


def bent_rod(n_along=250, n_around=40, radius=0.012, length=0.3, bend=1.2, seed=0):
    """Tube of 2*n_along*n_around side triangles bent along a circular arc, with
    end caps."""
    rng = np.random.default_rng(seed)
    s = np.linspace(0.0, 1.0, n_along + 1)
    Rb = length / bend
    ang = s * bend - bend / 2
    cx, cy = Rb * np.sin(ang), Rb * (1 - np.cos(ang))
    tx, ty = np.cos(ang), np.sin(ang)
    nx, ny = -ty, tx
    ph = 2 * np.pi * np.arange(n_around) / n_around
    rad = radius * (1 + 0.02 * (2 * rng.random((n_along + 1, n_around)) - 1))
    X = cx[:, None] + rad * np.cos(ph)[None, :] * nx[:, None]
    Y = cy[:, None] + rad * np.cos(ph)[None, :] * ny[:, None]
    Z = rad * np.sin(ph)[None, :]
    V = np.stack([X.ravel(), Y.ravel(), Z.ravel()], 1)
    i, j = np.meshgrid(np.arange(n_along), np.arange(n_around), indexing="ij")
    a = i * n_around + j
    b = (i + 1) * n_around + j
    c = (i + 1) * n_around + (j + 1) % n_around
    d = i * n_around + (j + 1) % n_around
    F = [np.stack([a, b, c], -1).reshape(-1, 3), np.stack([a, c, d], -1).reshape(-1, 3)]
    nv = len(V)
    caps = np.array([[cx[0], cy[0], 0.0], [cx[-1], cy[-1], 0.0]])
    V = np.concatenate([V, caps])
    jj = np.arange(n_around)
    F.append(np.stack([np.full(n_around, nv), (jj + 1) % n_around, jj], 1))
    last = n_along * n_around
    F.append(np.stack([np.full(n_around, nv + 1), last + jj, last + (jj + 1) % n_around], 1))
    return V.astype(np.float32), np.concatenate(F).astype(np.uint32)


def shell_scene(seed=0):
    """About 200k-triangle (195,439) shell with 64 holes and a 20k-triangle
    (20,080) bent rod."""
    env = shell_with_holes(level=6, seed=seed)
    rob = bent_rod(seed=seed + 1)
    return env, rob


def random_states(n, seed=0, lo=0.0, hi=1.0):
    """n states with translation ~ U[lo,hi]^3 and a Shoemake quaternion, rounded
    to binary32 like uw_random.random_state, returned as binary64 [n,7]."""
    rng = np.random.default_rng(seed)
    t = lo + (hi - lo) * rng.random((n, 3))
    q = _shoemake(rng, n)
    return np.concatenate([t, q], axis=1).astype(np.float32).astype(np.float64)


def segment_ends(starts, seed=0, lmin=0.05, lmax=0.5, lengths=None):
    """End states for touch segments: a uniform direction on S^5 (translation and
    axis-angle parts) scaled so the two parts' magnitudes sum to a length
    L ~ U[lmin, lmax] (or L = lengths when given); the end state is the start
    moved by the translation and rotated by the axis-angle (the apply() of
    lib/osr/osr_state.cc:243-259)."""
    rng = np.random.default_rng(seed)
    n = starts.shape[0]
    d = rng.normal(size=(n, 6))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    L = lmin + (lmax - lmin) * rng.random(n)
    if lengths is not None:
        L = np.asarray(lengths, dtype=np.float64).reshape(n)
    mag = np.linalg.norm(d[:, :3], axis=1) + np.linalg.norm(d[:, 3:], axis=1)
    d *= (L / mag)[:, None]
    tr, aa = d[:, :3], d[:, 3:]
    ends = np.empty_like(starts)
    ends[:, :3] = starts[:, :3] + tr
    ang = np.linalg.norm(aa, axis=1)
    axis = aa / np.where(ang > 0, ang, 1.0)[:, None]
    axis[ang == 0] = [1.0, 0.0, 0.0]
    qa = np.concatenate([np.cos(ang / 2)[:, None], np.sin(ang / 2)[:, None] * axis], axis=1)
    q0 = starts[:, 3:] / np.linalg.norm(starts[:, 3:], axis=1, keepdims=True)
    w1, x1, y1, z1 = qa.T
    w2, x2, y2, z2 = q0.T
    q = np.stack([w1 * w2 - x1 * x2 - y1 * y2 - z1 * z2,
                  w1 * x2 + x1 * w2 + y1 * z2 - z1 * y2,
                  w1 * y2 - x1 * z2 + y1 * w2 + z1 * x2,
                  w1 * z2 + x1 * y2 - y1 * x2 + z1 * w2], axis=1)
    ends[:, 3:] = q / np.linalg.norm(q, axis=1, keepdims=True)
    return ends


def lognormal_steps(n, mu=5.0, sigma=1.2, lo=8, hi=4096, seed=0):
    rng = np.random.default_rng(seed)
    return np.clip(np.round(rng.lognormal(mu, sigma, n)), lo, hi).astype(np.int64)
