"""GPU engine vs the oracle on the large scene: a noisy icosphere shell of
195,439 triangles with 64 lined holes and a bent rod of 20,080 triangles
(node ids beyond 16 bits, 22-level traversals).  Uneven segments with
lognormal step counts, touch searches and a visibility block."""
import numpy as np
import pytest

from puzzletunneldiscovery_amd import pyosr, workloads as W

pytestmark = pytest.mark.gpu


def _flush(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module")
def shell(oracle_mod):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    env, rob = W.shell_scene(0)
    uw = pyosr.create_unit_world_from_arrays(env, rob)
    ow = oracle_mod.OracleWorld().setup(*env, *rob)
    assert uw.bvh_info(0)[1] > 65536
    return uw, ow


def _free(uw, n, seed):
    qs = W.random_states(4 * n, seed=seed)
    return np.ascontiguousarray(qs[uw.is_valid_states(qs)][:n])


def test_shell_states(shell):
    uw, ow = shell
    qs = W.random_states(50_000, seed=301)
    _flush("shell: 50k states")
    g = uw.is_valid_states(qs)
    c = ow.valid_states(qs).astype(bool)
    assert 0.2 < c.mean() < 0.8
    assert int((g != c).sum()) == 0


def test_shell_uneven_segments(shell):
    uw, ow = shell
    vd = 1.0 / 1024
    starts = _free(uw, 600, 302)
    steps = W.lognormal_steps(len(starts), seed=303)
    ends = W.segment_ends(starts, seed=304, lengths=steps * vd)
    _flush("shell: 600 uneven segments")
    lf, st, done, lt, ta, ns = uw.transit_states_to_with_contact(starts, ends, vd)
    for i in range(len(starts)):
        olf, ost, odone, olt, ota, ons = ow.transit(starts[i], ends[i], vd)
        assert bool(done[i]) == bool(odone) and lt[i] == olt and ta[i] == ota and ns[i] == ons, i
        assert np.array_equal(lf[i], olf) and np.array_equal(st[i], ost), i
    assert 0 < done.sum() < len(done)


def test_shell_touch(shell):
    uw, ow = shell
    starts = _free(uw, 800, 305)
    ends = W.segment_ends(starts, seed=306)
    _flush("shell: 800 touch searches")
    r = uw.touch_search(starts, ends, 1.0 / 1024, 20)
    o = ow.touch_batch(starts, ends, 1.0 / 1024, 20)
    assert (~r["done"]).sum() > 50
    for k in ("last_tau", "tau", "lo", "hi"):
        assert np.array_equal(r[k], o[k]), k
    assert np.array_equal(r["done"], o["done"].astype(bool))
    assert np.array_equal(r["steps"], o["queries"])


def test_shell_visibility(shell):
    uw, ow = shell
    ms = _free(uw, 40, 307)
    _flush("shell: 40 milestones")
    g = uw.calculate_visibility_matrix(ms, True, 1.0 / 1024)
    c = ow.visibility_matrix(ms, True, 1.0 / 1024)
    assert np.array_equal(g, c)
    assert 0 < g.sum() < g.size - len(ms)


def test_shell_prescreen(shell):
    """The learned prescreen as a sampler would use it (src/GP/uw_random.py:46-52):
    trained on 2M exact labels of the shell scene, it selects among 200k new
    states and 2,000 lognormal segments, and the exact engine re-verifies
    every selection.  No state or segment is reported free unless the oracle
    finds it free, every selected free one is reported, the selection equals
    P(free) > threshold of the plain forward pass, and recall is at least 0.9."""
    import torch
    from puzzletunneldiscovery_amd import surrogate as S
    uw, ow = shell
    _flush("prescreen: 2M labels, training")
    tq = torch.from_numpy(W.random_states(2 * 1024 * 1024, seed=311)).cuda()
    ty = torch.zeros(tq.shape[0], dtype=torch.uint8, device="cuda")
    uw.valid_states_device(tq.data_ptr(), tq.shape[0], ty.data_ptr())
    torch.cuda.synchronize()
    uw.check_device()
    net = S.Surrogate(max_batch=16384, seed=312)
    net.fit(tq, ty, epochs=3, batch=16384, lr=2e-3)
    del tq, ty
    _flush("prescreen: 200k states")
    qs = W.random_states(200_000, seed=313)
    dq = torch.from_numpy(qs).cuda()
    pf = torch.zeros(len(qs), dtype=torch.float32, device="cuda")
    r = S.screen_states(uw, net, dq, 0.5, d_pfree=pf)
    torch.cuda.synchronize()
    uw.check_device()
    truth = ow.valid_states(qs).astype(bool)
    sel = S.selected_mask(r, len(qs)).cpu().numpy()
    free = r["free"].cpu().numpy().astype(bool)
    assert np.array_equal(free, sel & truth)
    assert np.array_equal(sel, pf.cpu().numpy() > 0.5)
    k = int(r["count"].item())
    assert len(np.unique(r["idx"][:k].cpu().numpy())) == k
    recall, precision = S.recall_precision(r, torch.from_numpy(truth).cuda())
    _flush(f"prescreen: states recall {recall:.4f} precision {precision:.4f}")
    assert recall >= 0.9 and precision > truth.mean()
    # a second screen reusing the first one's buffers (cleared on its stream)
    # selects the same states; a smaller batch reuses them too
    idx0 = np.sort(r["idx"][:k].cpu().numpy())
    r2 = S.screen_states(uw, net, dq, 0.5, out=r)
    torch.cuda.synchronize()
    assert int(r2["count"].item()) == k
    assert np.array_equal(np.sort(r2["idx"][:k].cpu().numpy()), idx0)
    assert np.array_equal(r2["free"].cpu().numpy().astype(bool), sel & truth)
    r3 = S.screen_states(uw, net, dq[:1000], 0.5, out=r2)
    torch.cuda.synchronize()
    assert int(r3["count"].item()) == int(sel[:1000].sum())
    assert np.array_equal(r3["free"].cpu().numpy().astype(bool), (sel & truth)[:1000])
    _flush("prescreen: 2,000 segments")
    a = _free(uw, 2000, 314)
    steps = W.lognormal_steps(len(a), seed=315)
    b = W.segment_ends(a, seed=316, lengths=steps * (1.0 / 1024))
    da, db = torch.from_numpy(a).cuda(), torch.from_numpy(b).cuda()
    rs = S.screen_segments(uw, net, da, db, 1.0 / 1024, 0.5, points=8)
    torch.cuda.synchronize()
    uw.check_device()
    tseg = np.array([ow.is_valid_transition(x, y, 1.0 / 1024) for x, y in zip(a, b)])
    ssel = S.selected_mask(rs, len(a)).cpu().numpy()
    sfree = rs["free"].cpu().numpy().astype(bool)
    assert np.array_equal(sfree, ssel & tseg)
    srecall, sprec = S.recall_precision(rs, torch.from_numpy(tseg).cuda())
    _flush(f"prescreen: segments recall {srecall:.4f} precision {sprec:.4f}")
    assert srecall >= 0.9 and 0.1 < tseg.mean() < 0.95


def test_subset_lists_from_the_caller(shell):
    """The subset entry points read their index list and count on the device:
    entries outside [0, n) are skipped, a count above n is cut to n, and
    entries of the output that are not listed keep their contents."""
    import torch
    uw, ow = shell
    qs = W.random_states(64, seed=321)
    dq = torch.from_numpy(qs).cuda()
    idx = torch.tensor([5, -1, 64, 7, 63, 1000] + [0] * 58, dtype=torch.int32, device="cuda")
    for cnt in (6, 500):
        count = torch.tensor([cnt], dtype=torch.int32, device="cuda")
        out = torch.full((64,), 9, dtype=torch.uint8, device="cuda")
        uw.valid_states_subset_device(dq.data_ptr(), 64, idx.data_ptr(), count.data_ptr(), out.data_ptr())
        torch.cuda.synchronize()
        uw.check_device()
        o = out.cpu().numpy()
        want = ow.valid_states(qs).astype(np.uint8)
        listed = [5, 7, 63] + ([0] if cnt > 6 else [])
        for i in range(64):
            assert o[i] == (want[i] if i in listed else 9), (cnt, i)
