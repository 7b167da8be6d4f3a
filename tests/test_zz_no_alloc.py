"""se3cq_touch_search_device_noalloc (no reference counterpart: an entry point
for callers that cannot allocate device memory) through the C ABI: bit-equal
to se3cq_touch_search_device on two calls in a row, the second with fewer
segments, both long enough to walk in the longest-first order that the
static scratch holds.  Sizes beyond SE3CQ_NOALLOC_MAX_SEGMENTS and a world
that was never uploaded are refused before anything is launched."""
import ctypes

import numpy as np
import pytest

from puzzletunneldiscovery_amd import _native as N, pyosr, workloads as W

pytestmark = pytest.mark.gpu
VD = 1.0 / 1024
LEVELS = 20
MAX_SEGMENTS = 1048576


def _outs(torch, n, fill):
    return [torch.full((n,), fill, dtype=dt, device="cuda")
            for dt in (torch.uint8, torch.float64, torch.float64, torch.float64, torch.float64, torch.int32)]


def _call(fn, w, d_a, d_b, n, outs, stream):
    return fn(w, ctypes.c_void_p(d_a.data_ptr()), ctypes.c_void_p(d_b.data_ptr()), ctypes.c_int64(n),
              ctypes.c_double(VD), ctypes.c_int(LEVELS), *[ctypes.c_void_p(o.data_ptr()) for o in outs],
              ctypes.c_void_p(stream))


def test_noalloc_equals_touch_search(torus_meshes):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    env, rob = torus_meshes
    uw = pyosr.create_unit_world_from_arrays(env, rob)
    qs = W.random_states(60_000, seed=501)
    a = np.ascontiguousarray(qs[uw.is_valid_states(qs)][:20_000])  # this query uploads the world
    b = W.segment_ends(a, seed=502)
    assert len(a) == 20_000
    L = N.lib()
    s = torch.cuda.Stream()
    for n in (20_000, 5_000):
        print(f"noalloc: {n} touch searches", flush=True)
        d_a = torch.from_numpy(a[:n].copy()).cuda()
        d_b = torch.from_numpy(b[:n].copy()).cuda()
        ref = _outs(torch, n, 0)
        got = _outs(torch, n, 7)
        torch.cuda.synchronize()
        assert _call(L.se3cq_touch_search_device, uw._w, d_a, d_b, n, ref, s.cuda_stream) == 0
        assert _call(L.se3cq_touch_search_device_noalloc, uw._w, d_a, d_b, n, got, s.cuda_stream) == 0, \
            L.se3cq_last_error()
        torch.cuda.synchronize()
        uw.check_device()
        for k, (x, y) in enumerate(zip(got, ref)):
            assert torch.equal(x, y), (n, k)
        assert (ref[0] == 0).sum().item() > 10  # some contacts, bisected


def test_noalloc_refuses_before_launching(torus_meshes):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    env, rob = torus_meshes
    L = N.lib()
    s = torch.cuda.Stream()
    d_a = torch.zeros((4, 7), dtype=torch.float64, device="cuda")
    d_a[:, 3] = 1.0
    outs = _outs(torch, 4, 0)
    fresh = pyosr.create_unit_world_from_arrays(env, rob)
    rc = _call(L.se3cq_touch_search_device_noalloc, fresh._w, d_a, d_a, 4, outs, s.cuda_stream)
    assert rc == -2 and b"not on the device" in L.se3cq_last_error()
    fresh.is_valid_states(np.array([[0.5, 0.5, 0.5, 1.0, 0.0, 0.0, 0.0]]))
    rc = _call(L.se3cq_touch_search_device_noalloc, fresh._w, d_a, d_a, MAX_SEGMENTS + 1, outs, s.cuda_stream)
    assert rc == -1 and b"at most 1048576" in L.se3cq_last_error()
    torch.cuda.synchronize()
    fresh.check_device()
