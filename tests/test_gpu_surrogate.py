"""The BF16-MFMA prescreen (include/se3mlp.h) against plain float64 NumPy
references of the same operations:

  * the matrix-core kernel on small integers, where bf16 inputs and fp32
    accumulation are exact, so C must equal A.B bit for bit (this pins the
    lane layout of v_mfma_f32_32x32x16_bf16);
  * the features, to the bit;
  * the forward pass against a float64 evaluation that rounds weights and
    activations to bf16 at the same points (tolerance below);
  * one Adam step against the float64 gradient (first-step Adam moves every
    weight by lr * sign(gradient));
  * training on exact labels of the torus scene.
"""
import numpy as np
import pytest

from puzzletunneldiscovery_amd import pyosr, workloads as W
from puzzletunneldiscovery_amd.surrogate import Surrogate, features_reference

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_cuda():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch


def bf16(x):
    """Round float32 values to bf16 (nearest even) and back, in float64."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bits(x):
    """bf16 bit patterns of float32 values, rounded to nearest even."""
    a = np.ascontiguousarray(x, np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    return ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)


def test_mfma_gemm_exact(torch_cuda):
    torch = torch_cuda
    from puzzletunneldiscovery_amd import _native as N
    rng = np.random.default_rng(0)
    for M, Nn, K in ((32, 32, 16), (96, 64, 48), (64, 32, 256), (256, 256, 256)):
        A = rng.integers(-4, 5, size=(M, K)).astype(np.float32)
        Bt = rng.integers(-4, 5, size=(Nn, K)).astype(np.float32)
        dA = torch.from_numpy(bits(A).view(np.int16)).cuda()
        dB = torch.from_numpy(bits(Bt).view(np.int16)).cuda()
        dC = torch.zeros((M, Nn), dtype=torch.float32, device="cuda")
        N.check_mlp(N.lib().se3mlp_gemm_device(dA.data_ptr(), K, dB.data_ptr(), K, M, Nn, K, dC.data_ptr(), Nn, None))
        torch.cuda.synchronize()
        assert np.array_equal(dC.cpu().numpy(), A @ Bt.T), (M, Nn, K)


def test_features_bit_exact(torch_cuda):
    torch = torch_cuda
    from puzzletunneldiscovery_amd import _native as N
    qs = W.random_states(1000, seed=1)
    qs[:100, 3:] *= 1.7  # non-unit quaternions are normalised first
    d = torch.from_numpy(qs).cuda()
    X = torch.zeros((1000, 32), dtype=torch.int16, device="cuda")
    N.check_mlp(N.lib().se3mlp_features_device(d.data_ptr(), 1000, X.data_ptr(), None))
    torch.cuda.synchronize()
    got = X.cpu().numpy().view(np.uint16)
    assert np.array_equal(got, bits(features_reference(qs).astype(np.float32)))


def reference_forward(params, states):
    """float64 evaluation with bf16 rounding of features, weights and hidden
    activations at the points the kernels round them."""
    h = bf16(features_reference(states).astype(np.float32))
    acts = [h]
    for l, (w, b) in enumerate(params):
        z = h @ bf16(w).T + b.astype(np.float64)
        if l < len(params) - 1:
            h = bf16(np.maximum(z, 0.0).astype(np.float32))
            acts.append(h)
        else:
            return z[:, :2], acts


def test_forward_matches_reference(torch_cuda):
    torch = torch_cuda
    net = Surrogate(max_batch=2048, seed=3)
    params = net.params()
    rng = np.random.default_rng(4)
    # non-trivial biases so every layer's bias path is exercised
    params = [(w, (0.1 * rng.standard_normal(len(b))).astype(np.float32) * (np.abs(w).sum(1) > 0))
              for w, b in params]
    net.set_params(params)
    qs = W.random_states(5000, seed=5)  # more than max_batch: streamed in chunks
    d = torch.from_numpy(qs).cuda()
    logits = torch.zeros((5000, 2), dtype=torch.float32, device="cuda")
    p = torch.zeros(5000, dtype=torch.float32, device="cuda")
    net.forward_device(d.data_ptr(), 5000, p.data_ptr(), None, logits.data_ptr())
    torch.cuda.synchronize()
    z = logits.cpu().numpy().astype(np.float64)
    ref, _ = reference_forward(params, qs)
    # fp32 accumulation in a different order can move a hidden activation by
    # one bf16 step (2^-8 relative); through four layers such moves shift a
    # logit by well under 2e-2 + 1% of the logit scale
    err = np.abs(z - ref)
    assert err.max() <= 2e-2 + 1e-2 * np.abs(ref).max(), err.max()
    assert err.mean() <= 2e-3
    pf = 1.0 / (1.0 + np.exp(-(ref[:, 0] - ref[:, 1])))
    assert np.abs(p.cpu().numpy() - pf).max() <= 1e-2


@pytest.mark.parametrize("n", [4096, 3001])
def test_adam_step_follows_gradient_sign(torch_cuda, n):
    # 3001 rows: a padded last block and a last slice of the weight-gradient
    # reduction that is not a multiple of its chunk
    torch = torch_cuda
    net = Surrogate(max_batch=4096, seed=7)
    params = net.params()
    qs = W.random_states(n, seed=8)
    env, rob = W.torus_scene(0)
    uw = pyosr.create_unit_world_from_arrays(env, rob)
    y = uw.is_valid_states(qs).astype(np.uint8)
    # float64 gradient of the mean cross-entropy (bf16-rounded forward)
    z, acts = reference_forward(params, qs)
    d = z[:, 1] - z[:, 0]
    p0 = 1.0 / (1.0 + np.exp(d))
    g0 = (p0 - y) / len(qs)
    loss_ref = np.mean(np.where(y == 1, np.logaddexp(0, d), np.logaddexp(0, -d)))
    G = np.zeros((len(qs), params[-1][0].shape[0]))
    G[:, 0], G[:, 1] = g0, -g0
    grads = [None] * len(params)
    bgrads = [None] * len(params)
    for l in range(len(params) - 1, -1, -1):
        grads[l] = G.T @ acts[l]
        bgrads[l] = G.sum(axis=0)
        if l:
            G = (G @ bf16(params[l][0])) * (acts[l] > 0)
    dq = torch.from_numpy(qs).cuda()
    dy = torch.from_numpy(y).cuda()
    lr = 1e-3
    loss = net.train_step_device(dq.data_ptr(), dy.data_ptr(), len(qs), lr, want_loss=True)
    assert abs(loss - loss_ref) <= 1e-3 * max(1.0, loss_ref)
    after = net.params()
    for l, ((w0, _), (w1, _)) in enumerate(zip(params, after)):
        g = grads[l]
        big = np.abs(g) > 1e-2 * np.abs(g).max()
        step = (w1 - w0)[big]
        assert np.all(np.abs(np.abs(step) - lr) <= 1e-5), l
        agree = np.mean(np.sign(step) == -np.sign(g[big]))
        assert agree >= 0.99, (l, agree)
        # padding stays zero
        r_used, c_used = net.shapes[l][2], net.shapes[l][3]
        assert not w1[r_used:].any() and not w1[:, c_used:].any()
    for l, ((_, b0), (_, b1)) in enumerate(zip(params, after)):
        g = bgrads[l]
        big = np.abs(g) > 1e-2 * np.abs(g).max()
        step = (b1 - b0)[big]
        assert np.all(np.abs(np.abs(step) - lr) <= 1e-5), l
        agree = np.mean(np.sign(step) == -np.sign(g[big]))
        assert agree >= 0.99, (l, agree)
        assert not b1[net.shapes[l][2]:].any()


def test_training_learns_torus_labels(torch_cuda):
    torch = torch_cuda
    env, rob = W.torus_scene(0)
    uw = pyosr.create_unit_world_from_arrays(env, rob)
    qs = W.random_states(300_000, seed=9)
    y = uw.is_valid_states(qs).astype(np.uint8)
    dq = torch.from_numpy(qs[:260_000]).cuda()
    dy = torch.from_numpy(y[:260_000]).cuda()
    net = Surrogate(max_batch=8192, seed=10)
    losses = net.fit(dq, dy, epochs=4, batch=8192, lr=2e-3)
    assert losses[-1] < losses[0]
    p, m = net.predict(qs[260_000:])
    acc = np.mean((p > 0.5) == (y[260_000:] == 1))
    assert acc > 0.85, acc
    assert np.array_equal(p > 0.5, m > 0)


def test_train_epoch_equals_consecutive_steps(torch_cuda):
    # the epoch entry point queues the same steps as calling the single-step
    # one per batch, so parameters and sampled losses agree to the bit
    torch = torch_cuda
    import ctypes
    from puzzletunneldiscovery_amd import _native as N
    qs = W.random_states(3 * 2048 + 100, seed=21)
    y = (np.random.default_rng(22).random(len(qs)) > 0.4).astype(np.uint8)
    dq = torch.from_numpy(qs).cuda()
    dy = torch.from_numpy(y).cuda()
    a = Surrogate(max_batch=2048, seed=23)
    b = Surrogate(max_batch=2048, seed=23)
    d_loss = torch.full((2,), -1.0, dtype=torch.float32, device="cuda")
    st = torch.cuda.current_stream().cuda_stream
    N.check_mlp(N.lib().se3mlp_train_epoch_device(a._h, dq.data_ptr(), dy.data_ptr(), None, len(qs), 2048,
                                                ctypes.c_float(2e-3), 2, d_loss.data_ptr(), st))
    want = []
    for k in range(3):
        l = b.train_step_device(dq[k * 2048:].data_ptr(), dy[k * 2048:].data_ptr(), 2048, 2e-3, k % 2 == 0, st)
        if k % 2 == 0:
            want.append(l)
    torch.cuda.synchronize()
    assert np.array_equal(d_loss.cpu().numpy(), np.array(want, np.float32))
    for (wa, ba), (wb, bb) in zip(a.params(), b.params()):
        assert np.array_equal(wa, wb) and np.array_equal(ba, bb)
    # a shuffled epoch read through the permutation equals steps over the
    # shuffled copy
    perm = torch.randperm(len(qs), device="cuda", generator=torch.Generator(device="cuda").manual_seed(24))
    c = Surrogate(max_batch=2048, seed=25)
    d = Surrogate(max_batch=2048, seed=25)
    N.check_mlp(N.lib().se3mlp_train_epoch_device(c._h, dq.data_ptr(), dy.data_ptr(), perm.data_ptr(), len(qs),
                                                2048, ctypes.c_float(2e-3), 2, d_loss.data_ptr(), st))
    gq, gy = dq[perm].contiguous(), dy[perm].contiguous()
    want = []
    for k in range(3):
        l = d.train_step_device(gq[k * 2048:].data_ptr(), gy[k * 2048:].data_ptr(), 2048, 2e-3, k % 2 == 0, st)
        if k % 2 == 0:
            want.append(l)
    torch.cuda.synchronize()
    assert np.array_equal(d_loss.cpu().numpy(), np.array(want, np.float32))
    for (wa, ba), (wb, bb) in zip(c.params(), d.params()):
        assert np.array_equal(wa, wb) and np.array_equal(ba, bb)
