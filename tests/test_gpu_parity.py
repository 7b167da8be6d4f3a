"""GPU engine vs the CPU oracle on the same seeded inputs: every boolean, every
tau and every returned state must be bit-identical.  Calls go through the
C ABI (libse3cq.so via ctypes)."""
import numpy as np
import pytest

from puzzletunneldiscovery_amd import pyosr, workloads as W

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


def test_states_torus(gpu_world, torus_oracle):
    qs = W.random_states(200_000, seed=5)
    _flush("states: 200k torus")
    g = gpu_world.is_valid_states(qs)
    c = torus_oracle.valid_states(qs).astype(bool)
    assert 0.05 < (~c).mean() < 0.95
    assert int((g != c).sum()) == 0


def test_heavy_states_whole_wavefront(gpu_world, torus_oracle):
    """The states whose one-lane traversal runs longest are set aside and
    walked by a whole wavefront: the heaviest 2,048 of 400k random states by
    the oracle's own node-test count (so many per wavefront that some stay
    with their lane), on their own and spread among light ones, must match
    the oracle."""
    qs = W.random_states(400_000, seed=55)
    v, nbv, _ = torus_oracle.valid_states(qs, counters=True)
    order = np.argsort(-nbv)
    heavy = qs[order[:2048]]
    assert nbv[order[2047]] > 14 * 23  # above the set-aside threshold of the torus scene
    _flush("states: 2,048 heaviest")
    g = gpu_world.is_valid_states(heavy)
    assert np.array_equal(g, v[order[:2048]].astype(bool))
    assert 0 < g.sum() < len(g)
    mixed = np.concatenate([heavy, qs[order[-60_000:]]])[np.random.default_rng(56).permutation(62_048)]
    _flush("states: heaviest among light ones")
    g2 = gpu_world.is_valid_states(mixed)
    assert np.array_equal(g2, torus_oracle.valid_states(mixed).astype(bool))


def test_states_edge_cases(gpu_world, torus_oracle):
    rng = np.random.default_rng(7)
    qs = W.random_states(20_000, seed=8)
    # un-normalised quaternions (the reference does not normalise before
    # building the rotation), far-away and identity states
    qs[:5000, 3:] *= rng.uniform(0.5, 1.5, size=(5000, 1))
    qs[5000:6000, :3] += 10.0
    qs[6000:6100, 3:] = [1.0, 0.0, 0.0, 0.0]
    qs[6100:6200, 3:] = [-1.0, 0.0, 0.0, 0.0]
    _flush("states: edge cases")
    g = gpu_world.is_valid_states(qs)
    c = torus_oracle.valid_states(qs).astype(bool)
    assert int((g != c).sum()) == 0
    assert g[5000:6000].all()


def test_empty_batch(gpu_world):
    out = gpu_world.is_valid_states(np.zeros((0, 7)))
    assert out.shape == (0,)


def test_perturbation(torus_meshes, oracle_mod):
    env, rob = torus_meshes
    uw = pyosr.create_unit_world_from_arrays(env, rob)
    ow = oracle_mod.OracleWorld().setup(*env, *rob)
    pert = np.array([0.01, -0.02, 0.005, 0.9998, 0.01, -0.015, 0.003])
    uw.set_perturbation(pert)
    ow.set_perturbation(pert)
    qs = W.random_states(20_000, seed=9)
    _flush("states: perturbation")
    assert int((uw.is_valid_states(qs) != ow.valid_states(qs).astype(bool)).sum()) == 0


def _free_starts(uw, n, seed):
    qs = W.random_states(4 * n, seed=seed)
    v = uw.is_valid_states(qs)
    return qs[v][:n]


def test_transit_with_contact(gpu_world, torus_oracle):
    starts = _free_starts(gpu_world, 600, 21)
    ends = W.segment_ends(starts, seed=22, lmin=0.01, lmax=0.6)
    # a few identical endpoints (distance 0 -> early return) and short ones
    ends[:5] = starts[:5]
    ends[5:10] = starts[5:10] + np.array([1e-4, 0, 0, 0, 0, 0, 0])
    _flush("transit: 600 segments")
    for vd in (1.0 / 1024, 0.0123, 0.05):
        lf, st, done, lt, ta, steps = gpu_world.transit_states_to_with_contact(starts, ends, vd)
        for i in range(len(starts)):
            olf, ost, odone, olt, ota, ons = torus_oracle.transit(starts[i], ends[i], vd)
            assert done[i] == odone, (vd, i)
            assert lt[i] == olt and ta[i] == ota, (vd, i, lt[i], olt, ta[i], ota)
            assert steps[i] == ons, (vd, i)
            assert np.array_equal(lf[i], olf), (vd, i)
            assert np.array_equal(st[i], ost), (vd, i)


def test_transit_tuple_shape(gpu_world, torus_oracle):
    starts = _free_starts(gpu_world, 4, 31)
    ends = W.segment_ends(starts, seed=32)
    t = gpu_world.transit_state_to_with_contact(starts[0], ends[0], 1.0 / 1024)
    assert len(t) == 5 and t[0].shape == (7,) and isinstance(t[2], bool)
    o = torus_oracle.transit(starts[0], ends[0], 1.0 / 1024)
    assert t[3] == float(np.float32(o[3])) and t[4] == float(np.float32(o[4]))
    s = gpu_world.transit_state_to(starts[0], ends[0], 1.0 / 1024)
    assert len(s) == 3


def test_valid_transitions(gpu_world, torus_oracle):
    starts = _free_starts(gpu_world, 500, 41)
    ends = _free_starts(gpu_world, 500, 42)
    ends[:3] = starts[:3]  # distance 0 -> false
    _flush("transitions: 500 pairs")
    for ivd in (1.0 / 1024, 0.02):
        g = gpu_world.is_valid_transitions(starts, ends, ivd)
        c = np.array([torus_oracle.is_valid_transition(a, b, ivd) for a, b in zip(starts, ends)])
        assert np.array_equal(g, c)
        assert not g[:3].any()



def test_valid_transitions_clamp(gpu_world, torus_oracle):
    """Pairs shorter than 2 * initial_verify_delta walk at vd = min(d/2, ivd)
    (unit_world.cc:477-486): two checks, the midpoint and the end.  Pairs of
    SE(3) length 0.2-0.35, among them every free-to-free pair of a seeded pool
    whose midpoint collides, at verify steps that put all, some or none of
    them below 2 * ivd; answers and step counts must equal the oracle's."""
    import torch
    from oracle.cases import pairs_with_midpoint_hits
    a, b, n_hit = pairs_with_midpoint_hits(torus_oracle, 197)
    assert n_hit >= 5
    n = len(a)
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    out = torch.zeros(n, dtype=torch.uint8, device="cuda")
    steps = torch.zeros(n, dtype=torch.int32, device="cuda")
    for ivd in (0.2, 0.15, 0.125, 0.1, 1.0 / 256):
        _flush(f"transitions: {n} pairs shorter than 2 ivd, ivd {ivd}")
        torch.cuda.synchronize()
        gpu_world.valid_transitions_device(da.data_ptr(), db.data_ptr(), n, ivd, out.data_ptr(), steps.data_ptr())
        torch.cuda.synchronize()
        gpu_world.check_device()
        c = [torus_oracle.is_valid_transition(a[i], b[i], ivd, with_steps=True) for i in range(n)]
        assert np.array_equal(out.cpu().numpy().astype(bool), np.array([x for x, _ in c]))
        assert np.array_equal(steps.cpu().numpy(), np.array([k for _, k in c]))
        if ivd == 0.2:
            # every pair is below 2 * ivd: two checks, or one when the midpoint collides
            assert set(steps.cpu().numpy().tolist()) == {1, 2}
    g = gpu_world.is_valid_transitions(a, b, 0.2)
    assert int((~g).sum()) == n_hit

def test_touch_search(gpu_world, torus_oracle):
    starts = _free_starts(gpu_world, 2000, 51)
    ends = W.segment_ends(starts, seed=52)
    _flush("touch: 2000 segments")
    r = gpu_world.touch_search(starts, ends, 1.0 / 1024, 20)
    o = torus_oracle.touch_batch(starts, ends, 1.0 / 1024, 20)
    assert (~r["done"]).sum() > 50
    assert np.array_equal(r["done"], o["done"].astype(bool))
    assert np.array_equal(r["last_tau"], o["last_tau"])
    assert np.array_equal(r["tau"], o["tau"])
    assert np.array_equal(r["lo"], o["lo"])
    assert np.array_equal(r["hi"], o["hi"])
    assert np.array_equal(r["steps"], o["queries"])


def test_touch_levels(gpu_world, torus_oracle):
    starts = _free_starts(gpu_world, 300, 61)
    ends = W.segment_ends(starts, seed=62)
    for levels in (0, 1, 4, 5, 6, 11, 23):
        r = gpu_world.touch_search(starts, ends, 1.0 / 512, levels)
        o = torus_oracle.touch_batch(starts, ends, 1.0 / 512, levels)
        assert np.array_equal(r["lo"], o["lo"]) and np.array_equal(r["hi"], o["hi"]), levels
        assert np.array_equal(r["steps"], o["queries"]), levels


def test_visibility(gpu_world, torus_oracle):
    ms = _free_starts(gpu_world, 48, 71)
    _flush("visibility: 48 milestones")
    g = gpu_world.calculate_visibility_matrix(ms, True, 1.0 / 256)
    c = torus_oracle.visibility_matrix(ms, True, 1.0 / 256)
    assert np.array_equal(g, c)
    assert np.array_equal(g, g.T) and not np.diag(g).any()
    a, b = ms[:20], ms[20:45]
    g2 = gpu_world.calculate_visibility_matrix2(a, True, b, True, 1.0 / 256)
    assert np.array_equal(g2, torus_oracle.visibility_matrix2(a, True, b, True, 1.0 / 256))
    gp = gpu_world.calculate_visibility_pair(a, True, ms[25:45], True, 1.0 / 256)
    assert np.array_equal(gp.reshape(-1), torus_oracle.visibility_pair(a, True, ms[25:45], True, 1.0 / 256))


def test_visibility_non_unit(gpu_world, torus_oracle):
    ms = _free_starts(gpu_world, 16, 81)
    raw = np.array([gpu_world.translate_from_unit_state(q) for q in ms])
    g = gpu_world.calculate_visibility_matrix(raw, False, 1.0 / 256)
    c = torus_oracle.visibility_matrix(raw, False, 1.0 / 256)
    assert np.array_equal(g, c)


def test_visibility_bits(gpu_world, torus_oracle):
    import torch
    ms = _free_starts(gpu_world, 40, 91)
    n = ms.shape[0]
    dq = torch.from_numpy(ms).cuda()
    words = (n * n + 31) // 32
    bits = torch.zeros(words, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()
    gpu_world.visibility_bits_device(dq.data_ptr(), n, 1.0 / 256, bits.data_ptr())
    torch.cuda.synchronize()
    gpu_world.check_device()
    b = bits.cpu().numpy().view(np.uint32)
    flat = np.unpackbits(b.view(np.uint8), bitorder="little")[: n * n].reshape(n, n)
    c = torus_oracle.visibility_matrix(ms, True, 1.0 / 256)
    assert np.array_equal(flat.astype(np.int8), c)


def test_disentangled(torus_meshes, oracle_mod):
    """isDisentangled (unit_world.cc:249-264): bounding boxes of both models,
    bit-identical to the oracle, also with non-unit quaternions and a
    perturbation."""
    env, rob = torus_meshes
    uw = pyosr.create_unit_world_from_arrays(env, rob)
    ow = oracle_mod.OracleWorld().setup(*env, *rob)
    rng = np.random.default_rng(17)
    qs = W.random_states(50_000, seed=18, lo=-0.2, hi=1.2)
    qs[:5000, 3:] *= rng.uniform(0.5, 1.5, size=(5000, 1))
    _flush("disentangled: 50k states")
    g = uw.is_disentangled_states(qs)
    c = ow.disentangled_states(qs)
    assert 0.05 < c.mean() < 0.95
    assert np.array_equal(g, c)
    assert uw.is_disentangled(qs[0]) == bool(c[0])
    pert = np.array([0.03, -0.01, 0.02, 0.97, 0.1, -0.2, 0.05])
    uw.set_perturbation(pert)
    ow.set_perturbation(pert)
    assert np.array_equal(uw.is_disentangled_states(qs), ow.disentangled_states(qs))


def test_touch_search_in_hip_graph(gpu_world, torus_oracle):
    """The device entry points allocate nothing after the first call and
    clear with kernels, so a caller may capture them in a HIP graph: every
    replay (with eager work and synchronisation in between) must reproduce
    the eager result."""
    import torch
    starts = _free_starts(gpu_world, 512, 111)
    ends = W.segment_ends(starts, seed=112)
    n = len(starts)
    dev = torch.device("cuda")
    df, dt = torch.from_numpy(starts).to(dev), torch.from_numpy(ends).to(dev)
    outs = [torch.zeros(n, dtype=t, device=dev) for t in (torch.uint8, torch.float64, torch.float64,
                                                            torch.float64, torch.float64, torch.int32)]
    s = torch.cuda.Stream()

    def call():
        gpu_world.touch_search_device(df.data_ptr(), dt.data_ptr(), n, 1.0 / 1024, 20,
                                      *[o.data_ptr() for o in outs], stream=s.cuda_stream)

    with torch.cuda.stream(s):
        call()  # first call uploads the scene; after that nothing is allocated
    torch.cuda.synchronize()
    eager = [o.clone() for o in outs]
    o = torus_oracle.touch_batch(starts, ends, 1.0 / 1024, 20)
    assert np.array_equal(eager[3].cpu().numpy(), o["lo"]) and np.array_equal(eager[4].cpu().numpy(), o["hi"])
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        call()
    for rep in range(3):
        for x in outs:
            x.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(outs, eager):
            assert torch.equal(a, b), rep


def test_contacts(gpu_world, torus_oracle):
    """intersectingSegments (unit_world.cc:959-1012): the GPU lists the same
    intersecting triangle pairs as the oracle's brute force, with bit-equal
    segments and flags."""
    qs = W.random_states(300, seed=121)
    valid = torus_oracle.valid_states(qs).astype(bool)
    qs = qs[~valid][:60]
    _flush("contacts: 60 colliding states")
    counts, pairs, segs, flags = gpu_world.contacts(qs, cap=512)
    for i, q in enumerate(qs):
        p, sg, fl, m = torus_oracle.contacts(q)
        assert counts[i] == m and m > 0
        k = min(m, 512)
        assert np.array_equal(pairs[i, :k], p[:k])
        assert np.array_equal(segs[i, :k], sg[:k])
        assert np.array_equal(flags[i, :k], fl[:k])
    p0, p1, mag, fp = gpu_world.intersecting_segments(qs[0])
    assert fp.shape == (counts[0], 2) and p0.shape == (counts[0], 3)
    assert np.allclose(mag, np.linalg.norm(p1 - p0, axis=1))
    free = W.random_states(300, seed=121)[valid][:5]
    assert not gpu_world.contacts(free)[0].any()


def test_concurrent_streams_share_a_world(gpu_world, torus_oracle):
    """Persistent kernels take a fresh work counter per launch, so device
    calls in flight on different streams of one world do not interfere."""
    import torch
    starts = _free_starts(gpu_world, 1500, 131)
    ends = W.segment_ends(starts, seed=132)
    qs = W.random_states(50_000, seed=133)
    n, m = len(starts), len(qs)
    dev = torch.device("cuda")
    df, dt, dq = (torch.from_numpy(x).to(dev) for x in (starts, ends, qs))
    streams = [torch.cuda.Stream() for _ in range(3)]
    outs = []
    for rep in range(2):
        o = [torch.zeros(n, dtype=t, device=dev) for t in (torch.uint8, torch.float64, torch.float64,
                                                             torch.float64, torch.float64, torch.int32)]
        v = torch.zeros(m, dtype=torch.uint8, device=dev)
        tr = [torch.zeros(n, dtype=t, device=dev) for t in (torch.uint8, torch.float64, torch.float64, torch.int32)]
        torch.cuda.synchronize()
        gpu_world.touch_search_device(df.data_ptr(), dt.data_ptr(), n, 1.0 / 1024, 20, *[x.data_ptr() for x in o],
                                      stream=streams[0].cuda_stream)
        gpu_world.valid_states_device(dq.data_ptr(), m, v.data_ptr(), streams[1].cuda_stream)
        gpu_world.transit_device(df.data_ptr(), dt.data_ptr(), n, 1.0 / 1024, *[x.data_ptr() for x in tr],
                                 stream=streams[2].cuda_stream)
        torch.cuda.synchronize()
        gpu_world.check_device()
        outs.append((o, v, tr))
    ref = torus_oracle.touch_batch(starts, ends, 1.0 / 1024, 20)
    want_v = torus_oracle.valid_states(qs).astype(np.uint8)
    for o, v, tr in outs:
        assert np.array_equal(o[3].cpu().numpy(), ref["lo"]) and np.array_equal(o[4].cpu().numpy(), ref["hi"])
        assert np.array_equal(v.cpu().numpy(), want_v)
        assert np.array_equal(tr[0].cpu().numpy(), ref["done"].astype(np.uint8))


def test_persistent_entry_points_in_hip_graph(gpu_world, torus_oracle):
    """State validity (lane refill), OMPL motions and the bit-packed
    visibility matrix captured in one graph: every replay matches eager."""
    import torch
    from puzzletunneldiscovery_amd import pyse3ompl
    dev = torch.device("cuda")
    qs = W.random_states(20_000, seed=141)
    ms = _free_starts(gpu_world, 96, 142)
    drv = pyse3ompl.OmplDriver.from_unit_world(gpu_world)
    drv.set_bb([0, 0, 0], [1, 1, 1])
    drv.set_cdres(0.002)
    q0 = W.random_states(400, seed=143)[:, [0, 1, 2, 4, 5, 6, 3]].copy()
    q1 = q0.copy()
    q1[:, :3] += np.random.default_rng(144).normal(0, 0.1, (400, 3))
    dq, dms, d0, d1 = (torch.from_numpy(np.ascontiguousarray(x)).to(dev) for x in (qs, ms, q0, q1))
    v = torch.zeros(len(qs), dtype=torch.uint8, device=dev)
    words = (len(ms) * len(ms) + 31) // 32
    bits = torch.zeros(words, dtype=torch.int32, device=dev)
    mo = torch.zeros(len(q0), dtype=torch.int32, device=dev)
    s = torch.cuda.Stream()

    def call():
        gpu_world.valid_states_device(dq.data_ptr(), len(qs), v.data_ptr(), s.cuda_stream)
        gpu_world.visibility_bits_device(dms.data_ptr(), len(ms), 1.0 / 256, bits.data_ptr(), None, s.cuda_stream)
        drv.validate_motion_pairs_device(d0.data_ptr(), d1.data_ptr(), len(q0), mo.data_ptr(), None,
                                         stream=s.cuda_stream)

    with torch.cuda.stream(s):
        call()
    torch.cuda.synchronize()
    eager = [x.clone() for x in (v, bits, mo)]
    assert np.array_equal(eager[0].cpu().numpy(), torus_oracle.valid_states(qs))
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        call()
    for rep in range(3):
        for x in (v, bits, mo):
            x.fill_(7 if x.dtype != torch.uint8 else 3)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip((v, bits, mo), eager):
            assert torch.equal(a, b), rep
