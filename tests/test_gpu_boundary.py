"""The boundary's behaviour around the queries rather than their answers:
blocking calls from several threads on one world, device errors that are
reported once and then cleared, work counters of captured HIP graphs beside
many eager launches, the perturbation a captured graph keeps, the
longest-first segment order of large launches, and the reference's own
"Expecting True" states (src/RL/test_collide.py:40-43) on the engine."""
import os
import threading
import time

import numpy as np
import pytest

from puzzletunneldiscovery_amd import pyosr, workloads as W

pytestmark = pytest.mark.gpu
REF = os.path.join(os.path.dirname(__file__), "golden", "ref")
VD = 1.0 / 1024


def _flush(msg):
    print(msg, flush=True)


def _free_starts(uw, n, seed):
    qs = W.random_states(4 * n, seed=seed)
    return qs[uw.is_valid_states(qs)][:n]


def test_host_calls_from_two_threads_overlap(torus_meshes, torus_oracle):
    """module.cc:138-184 releases the GIL on the const query methods, so
    Python threads query one world at once.  Thread A walks 32 segments of 2^21
    steps each far outside the scene (each walk holds one wavefront for tens
    of milliseconds, the rest of the GPU stays free); thread B, started once A
    is inside its call, makes 20 state queries.  B's calls all end before A's
    does (a world-wide lock would hold B until A ends), and every answer equals
    the one the same call gives alone."""
    env, rob = torus_meshes
    uw = pyosr.create_unit_world_from_arrays(env, rob)
    far = W.random_states(32, seed=401)
    far[:, :3] = np.floor(far[:, :3] * 64) / 64 + 5.0   # dyadic, so the 0.25 move is exact
    ends = far.copy()
    ends[:, 0] += 0.25
    vd = 0.25 / 2 ** 21
    qs = [W.random_states(10_000, seed=410 + k) for k in range(20)]
    want_b = [torus_oracle.valid_states(q).astype(bool) for q in qs]
    _flush("threads: reference walk alone")
    alone = uw.transit_states_to_with_contact(far, ends, vd)
    t = {}
    got_b = []
    inside = threading.Event()

    def thread_a():
        t["a0"] = time.perf_counter()
        inside.set()
        t["ra"] = uw.transit_states_to_with_contact(far, ends, vd)
        t["a1"] = time.perf_counter()

    def thread_b():
        inside.wait()
        time.sleep(0.002)
        t["b0"] = time.perf_counter()
        for q in qs:
            got_b.append(uw.is_valid_states(q))
        t["b1"] = time.perf_counter()

    _flush("threads: A and B together")
    ta, tb = threading.Thread(target=thread_a), threading.Thread(target=thread_b)
    ta.start()
    tb.start()
    ta.join(60)
    tb.join(60)
    assert not ta.is_alive() and not tb.is_alive()
    for a, b in zip(t["ra"], alone):
        assert np.array_equal(a, b)
    assert t["ra"][2].all() and (t["ra"][5] >= 2 ** 21 - 1).all()
    for g, w_ in zip(got_b, want_b):
        assert np.array_equal(g, w_)
    assert t["b0"] < t["a1"] and t["b1"] < t["a1"], t


def test_device_error_is_reported_once(torus_meshes):
    """A walk that would need 2^31 or more steps fails its own call; the next
    call on the same world, and the next check after a failed _device call,
    are clean again (the error word is taken, not left set)."""
    import torch
    env, rob = torus_meshes
    uw = pyosr.create_unit_world_from_arrays(env, rob)
    starts = _free_starts(uw, 64, 421)
    ends = W.segment_ends(starts, seed=422)
    _flush("errors: host form")
    with pytest.raises(RuntimeError, match="2\\^31"):
        uw.transit_states_to_with_contact(starts, ends, 1e-12)
    ok = uw.transit_states_to_with_contact(starts, ends, VD)
    assert ok[2].any()
    _flush("errors: device form")
    n = len(starts)
    d_f, d_t = torch.from_numpy(starts).cuda(), torch.from_numpy(ends).cuda()
    outs = [torch.zeros(n, dtype=dt, device="cuda") for dt in (torch.uint8, torch.float64, torch.float64, torch.int32)]
    uw.transit_device(d_f.data_ptr(), d_t.data_ptr(), n, 1e-12, *[o.data_ptr() for o in outs])
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="2\\^31"):
        uw.check_device()
    uw.check_device()
    uw.transit_device(d_f.data_ptr(), d_t.data_ptr(), n, VD, *[o.data_ptr() for o in outs])
    torch.cuda.synchronize()
    uw.check_device()
    assert np.array_equal(outs[0].cpu().numpy().astype(bool), ok[2])


def test_captured_graph_beside_many_eager_launches(gpu_world):
    """A captured touch search keeps a work counter of its own: replayed on
    one stream while more than the 64 eager counters' worth of launches run on
    another, every replay and every eager launch gives the eager results."""
    import torch
    a = _free_starts(gpu_world, 3000, 431)
    b = W.segment_ends(a, seed=432)
    small_a, small_b = a[:96], b[:96]
    ref = gpu_world.touch_search(a, b, VD, 20)
    ref_small = gpu_world.touch_search(small_a, small_b, VD, 20)

    def bufs(x, y):
        n = len(x)
        ins = [torch.from_numpy(x).cuda(), torch.from_numpy(y).cuda()]
        outs = [torch.zeros(n, dtype=dt, device="cuda")
                for dt in (torch.uint8, torch.float64, torch.float64, torch.float64, torch.float64, torch.int32)]
        return ins, outs

    ins, outs = bufs(a, b)
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()

    def call(i, o, n, st):
        gpu_world.touch_search_device(i[0].data_ptr(), i[1].data_ptr(), n, VD, 20, *[x.data_ptr() for x in o],
                                      stream=st.cuda_stream)

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s1):
        call(ins, outs, len(a), s1)
    torch.cuda.synchronize()
    small = [bufs(small_a, small_b) for _ in range(100)]
    _flush("graph: 4 replays beside 100 eager launches")
    for rep in range(4):
        for x in outs:
            x.zero_()
        torch.cuda.synchronize()
        g.replay()
        for k in range(25):
            si, so = small[25 * rep + k]
            call(si, so, len(small_a), s2)
        torch.cuda.synchronize()
        gpu_world.check_device()
        for key, x in zip(("done", "last_tau", "tau", "lo", "hi", "steps"), outs):
            assert np.array_equal(x.cpu().numpy().astype(ref[key].dtype), ref[key]), (rep, key)
    for si, so in small:
        for key, x in zip(("done", "last_tau", "tau", "lo", "hi", "steps"), so):
            assert np.array_equal(x.cpu().numpy().astype(ref_small[key].dtype), ref_small[key]), key


def test_graph_keeps_its_perturbation(torus_meshes, oracle_mod):
    """set_perturbation (unit_world.cc:164-175) selects a new device block and
    never overwrites the old one: a graph captured before the change replays
    with the perturbation it was captured with, eager calls use the new one,
    and both equal the oracle under the matching perturbation."""
    import torch
    env, rob = torus_meshes
    uw = pyosr.create_unit_world_from_arrays(env, rob)
    ow = oracle_mod.OracleWorld().setup(*env, *rob)
    qs = W.random_states(30_000, seed=441)
    pert = np.array([0.03, -0.02, 0.01, 0.9, 0.2, -0.3, 0.1])
    base = ow.valid_states(qs).astype(bool)
    ow2 = oracle_mod.OracleWorld().setup(*env, *rob)
    ow2.set_perturbation(pert)
    moved = ow2.valid_states(qs).astype(bool)
    assert (base != moved).sum() > 100
    d_q = torch.from_numpy(qs).cuda()
    out = torch.zeros(len(qs), dtype=torch.uint8, device="cuda")
    s = torch.cuda.Stream()
    assert np.array_equal(uw.is_valid_states(qs), base)  # uploads the world
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        uw.valid_states_device(d_q.data_ptr(), len(qs), out.data_ptr(), s.cuda_stream)
    torch.cuda.synchronize()
    _flush("perturbation: replay after set_perturbation")
    uw.set_perturbation(pert)
    assert np.array_equal(uw.is_valid_states(qs), moved)
    for _ in range(2):
        out.zero_()
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert np.array_equal(out.cpu().numpy().astype(bool), base)
    uw.set_perturbation([0, 0, 0, 1, 0, 0, 0])
    assert np.array_equal(uw.is_valid_states(qs), base)
    uw.set_perturbation(pert)
    assert np.array_equal(uw.is_valid_states(qs), moved)


def test_longest_first_order_matches_oracle(gpu_world, torus_oracle):
    """Launches of more segments than the device keeps wavefronts resident
    walk their segments longest first; the answers stay the oracle's: 12,000
    uneven segments (transit), 12,000 touch searches and 12,000 transitions."""
    a = _free_starts(gpu_world, 12_000, 451)
    steps = W.lognormal_steps(len(a), seed=452, hi=1024)
    b = W.segment_ends(a, seed=453, lengths=steps * VD)
    b[:5] = a[:5]
    _flush("order: 12,000 uneven transits")
    lf, st, done, lt, ta, ns = gpu_world.transit_states_to_with_contact(a, b, VD)
    o = torus_oracle.touch_batch(a, b, VD, 0)
    assert np.array_equal(done, o["done"].astype(bool)) and np.array_equal(ns, o["queries"])
    assert np.array_equal(lt, o["last_tau"]) and np.array_equal(ta, o["tau"])
    _flush("order: 12,000 touch searches")
    e = W.segment_ends(a, seed=454)
    r = gpu_world.touch_search(a, e, VD, 20)
    o = torus_oracle.touch_batch(a, e, VD, 20)
    for k, ok in (("done", "done"), ("lo", "lo"), ("hi", "hi"), ("steps", "queries")):
        assert np.array_equal(r[k], o[ok].astype(r[k].dtype)), k
    _flush("order: 12,000 transitions")
    f = W.random_states(12_000, seed=455)
    g = gpu_world.is_valid_transitions(a, f, 1.0 / 256)
    c = np.array([torus_oracle.is_valid_transition(x, y, 1.0 / 256) for x, y in zip(a[:3000], f[:3000])])
    assert np.array_equal(g[:3000], c)


def test_reference_expecting_true_states():
    """src/RL/test_collide.py:40-43 loads res/alpha/env-1.2.obj and
    res/alpha/robot.obj and prints "Expecting True" for two unit states."""
    uw = pyosr.UnitWorld()
    uw.loadModelFromFile(os.path.join(REF, "env-1.2.obj"))
    uw.loadRobotFromFile(os.path.join(REF, "robot.obj"))
    uw.scaleToUnit()
    uw.angleModel(0.0, 0.0)
    for x in (0.04, 0.7):
        q = np.array([x, 0.0, 0.0, 0.5, 0.5, 0.5, 0.5], np.float32).astype(np.float64)
        assert uw.is_valid_state(q)
