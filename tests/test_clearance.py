"""Key-configuration clearance host logic (puzzletunneldiscovery_amd.clearance,
src/GP/pipeline/keyconf.py:276-317, solve2.py:104-125) against the oracle's
one-sample-at-a-time restatement; the oracle's direction sampler pinned by
what the sampling methods fix.  No GPU: the touch queries themselves are
checked in test_gpu_keyscreen.py."""
import math

import numpy as np
import pytest

from oracle import clearance as OC
from puzzletunneldiscovery_amd import clearance as C


# ---- the oracle's sampler, pinned ---------------------------------------------------
def test_oracle_directions_lie_where_they_should():
    rs = np.random.RandomState(1)
    tr, aa = OC.directions(20000, rs)
    assert np.allclose(np.linalg.norm(tr, axis=1), 1.0, rtol=0, atol=4e-16)
    r = np.linalg.norm(aa, axis=1)
    assert (r <= 2 * math.pi * (1 + 1e-15)).all()
    # uniform on the sphere: each coordinate has mean 0 and variance 1/3
    assert np.allclose(tr.mean(axis=0), 0, atol=0.02)
    assert np.allclose((tr ** 2).mean(axis=0), 1 / 3, atol=0.01)
    # uniform in the ball of radius R: P(|x| <= R / 2) = 1/8, E|x|^2 = 3 R^2 / 5
    assert abs((r <= math.pi).mean() - 0.125) < 0.01
    assert abs((r ** 2).mean() / (4 * math.pi ** 2) - 0.6) < 0.01


def test_oracle_sampler_consumes_whole_tries():
    # the first sample by hand: the first accepted pair, then the first accepted triple
    rs = np.random.RandomState(9)
    u = np.random.RandomState(9).rand(400)
    x = 2.0 * (u - 0.5)
    o = 0
    while not x[o] ** 2 + x[o + 1] ** 2 < 1:
        o += 2
    s = o
    o += 2
    while not x[o] ** 2 + x[o + 1] ** 2 + x[o + 2] ** 2 <= 1:
        o += 3
    tr, aa = OC.directions(1, rs)
    l2 = x[s] ** 2 + x[s + 1] ** 2
    assert np.allclose(tr[0], [2 * x[s] * math.sqrt(1 - l2), 2 * x[s + 1] * math.sqrt(1 - l2), 1 - 2 * l2],
                       rtol=0, atol=1e-15)
    assert np.array_equal(aa[0], x[o:o + 3] * 2 * math.pi)
    assert rs.rand() == u[o + 3]


# ---- the product against the oracle -----------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 7, 1000, 70000])
def test_touch_directions_match_oracle(n):
    rs = np.random.RandomState(n)
    tr, aa = C.touch_directions(n, rs)
    ref = np.random.RandomState(n)
    wtr, waa = OC.directions(n, ref)
    assert np.array_equal(tr, wtr) and np.array_equal(aa, waa)
    # the stream is left exactly where the reference leaves it
    assert np.array_equal(rs.rand(5), ref.rand(5))


def test_touch_directions_global_stream():
    np.random.seed(123)
    tr, aa = C.touch_directions(300)
    after = np.random.rand(3)
    np.random.seed(123)
    wtr, waa = OC.directions(300, np.random)
    assert np.array_equal(tr, wtr) and np.array_equal(aa, waa)
    assert np.array_equal(after, np.random.rand(3))
    assert C.touch_directions(0)[0].shape == (0, 3)
    # the reference draws from the legacy stream; a Generator has no such state
    with pytest.raises(TypeError):
        C.touch_directions(3, np.random.default_rng(0))


def test_clearance_tasks():
    # keyconf.py:284-305: chunks of the keys; task 0 when there is one chunk
    assert np.array_equal(C.clearance_task(10, 1, 5), np.arange(10))
    parts = [C.clearance_task(10, 3, k) for k in range(3)]
    assert [list(p) for p in parts] == [[0, 1, 2], [3, 4, 5], [6, 7, 8, 9]]


def test_distance_batch_files_and_downsampling(tmp_path):
    rng = np.random.default_rng(5)
    osc = rng.uniform(0, 1, (23, 16))
    oskey = rng.normal(size=(23, 7))
    np.savez(tmp_path / "clearance_batch-0.npz", DISTANCE_BATCH=osc[:10])
    np.savez(tmp_path / "clearance_batch-1.npz", DISTANCE_BATCH=np.zeros(0))
    np.savez(tmp_path / "clearance_batch-2.npz", DISTANCE_BATCH=osc[10:])
    np.savez(tmp_path / "clearance_batch-4.npz", DISTANCE_BATCH=osc[:1])  # after a gap: not read
    got = C.load_distance_batches(str(tmp_path))
    assert np.array_equal(got, osc)
    for k in (1, 5, 21, 40):
        assert np.array_equal(C.least_visible_keys(oskey, got, k), OC.least_visible(oskey, osc, k))
    # the two reservations never come back
    low = osc.copy()
    low[:2] = 0
    top = C.least_visible_keys(oskey, low, 3)
    assert not any((top == oskey[i]).all(axis=1).any() for i in range(2))
    with pytest.raises(ValueError):
        C.least_visible_keys(oskey[:5], got, 2)
    with pytest.raises(FileNotFoundError):
        C.load_distance_batches(str(tmp_path / "none"))
