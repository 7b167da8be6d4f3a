"""The learned prescreen (include/se3mlp.h): a BF16-MFMA multilayer perceptron
that estimates P(free) of SE(3) states so a sampler can spend exact checks on
likely-free candidates.  It never decides validity: everything it accepts is
re-verified by the exact engine (UnitWorld.is_valid_states and friends).

Device memory comes from torch (CUDA tensors on ROCm); the arithmetic is the
library's own HIP kernels.
"""
import ctypes

import numpy as np

from . import _native as N


class Surrogate:
    """12 -> 256 -> 256 -> 256 -> 256 -> 2 MLP; features are the rotation
    matrix of the normalised quaternion and the translation minus 0.5."""

    def __init__(self, max_batch=65536, seed=0):
        L = N.lib()
        h = ctypes.c_void_p()
        N.check_mlp(L.se3mlp_create(int(max_batch), ctypes.c_uint64(seed), ctypes.byref(h)))
        self._h = h
        self.max_batch = int(max_batch)
        sizes = (ctypes.c_int64 * 20)()
        self.layers = L.se3mlp_param_layout(sizes)
        self.shapes = [tuple(sizes[4 * l:4 * l + 4]) for l in range(self.layers)]

    def __del__(self):
        h = getattr(self, "_h", None)
        if h and N is not None and N._lib is not None:
            N._lib.se3mlp_destroy(h)
            self._h = None

    # ---- parameters -------------------------------------------------------
    def params(self):
        """[(W [rows, cols] fp32, b [rows] fp32)] per layer (padded storage)."""
        out = []
        for l, (r, c, _, _) in enumerate(self.shapes):
            w = np.zeros((r, c), np.float32)
            b = np.zeros(r, np.float32)
            N.check_mlp(N.lib().se3mlp_get_params(self._h, l, w.ctypes.data_as(N.c_float_p),
                                                 b.ctypes.data_as(N.c_float_p)))
            out.append((w, b))
        return out

    def set_params(self, params):
        for l, (w, b) in enumerate(params):
            w = np.ascontiguousarray(w, np.float32)
            b = np.ascontiguousarray(b, np.float32)
            N.check_mlp(N.lib().se3mlp_set_params(self._h, l, w.ctypes.data_as(N.c_float_p),
                                                 b.ctypes.data_as(N.c_float_p)))

    # ---- device-pointer calls ----------------------------------------------------
    def forward_device(self, d_states, n, d_pfree=None, d_margin=None, d_logits=None, stream=None):
        N.check_mlp(N.lib().se3mlp_forward_device(self._h, d_states, int(n), d_pfree, d_margin, d_logits, stream))

    def screen_states_device(self, d_states, n, threshold, d_idx, d_count, d_pfree=None, stream=None):
        N.check_mlp(N.lib().se3mlp_screen_states_device(self._h, d_states, int(n), ctypes.c_float(threshold), d_idx,
                                                       d_count, d_pfree, stream))

    def screen_segments_device(self, d_from, d_to, n, points, threshold, d_idx, d_count, stream=None):
        N.check_mlp(N.lib().se3mlp_screen_segments_device(self._h, d_from, d_to, int(n), int(points),
                                                         ctypes.c_float(threshold), d_idx, d_count, stream))

    def train_step_device(self, d_states, d_labels, n, lr, want_loss=False, stream=None):
        loss = ctypes.c_float(0.0)
        N.check_mlp(N.lib().se3mlp_train_step_device(self._h, d_states, d_labels, int(n), ctypes.c_float(lr),
                                                    ctypes.byref(loss) if want_loss else None, stream))
        return loss.value if want_loss else None

    # ---- torch conveniences --------------------------------------------------------
    def predict(self, states):
        """(P(free), margin) as numpy float32 arrays for host states [n, 7]."""
        import torch
        q = torch.from_numpy(np.ascontiguousarray(states, np.float64).reshape(-1, 7)).cuda()
        n = q.shape[0]
        p = torch.empty(n, dtype=torch.float32, device="cuda")
        m = torch.empty(n, dtype=torch.float32, device="cuda")
        if n:
            self.forward_device(q.data_ptr(), n, p.data_ptr(), m.data_ptr(),
                                stream=torch.cuda.current_stream().cuda_stream)
        torch.cuda.synchronize()
        return p.cpu().numpy(), m.cpu().numpy()

    def fit(self, d_states, d_labels, epochs=1, batch=8192, lr=1e-3, seed=0, log=None):
        """Adam over shuffled minibatches of device tensors (states float64
        [n, 7], labels uint8 [n], 1 = free).  Returns the mean loss per epoch."""
        import torch
        n = d_states.shape[0]
        batch = min(batch, self.max_batch, n)
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        stream = torch.cuda.current_stream().cuda_stream
        losses = []
        every = 16  # the loss of every 16th step enters the epoch's mean
        d_states = d_states.contiguous()
        d_labels = d_labels.contiguous()
        for ep in range(epochs):
            perm = torch.randperm(n, device="cuda", generator=g)
            steps = n // batch
            d_loss = torch.empty((steps + every - 1) // every, dtype=torch.float32, device="cuda")
            # the whole epoch is queued by the library, one host call; the
            # kernels read the states through the permutation
            N.check_mlp(N.lib().se3mlp_train_epoch_device(self._h, d_states.data_ptr(), d_labels.data_ptr(),
                                                        perm.data_ptr(), n, batch, ctypes.c_float(lr), every,
                                                        d_loss.data_ptr(), stream))
            losses.append(float(d_loss.mean().item()))
            if log:
                log(f"epoch {ep}: loss {losses[-1]:.4f}")
        torch.cuda.synchronize()
        return losses


class _one_stream:
    """The stream both halves of a screen run on, so the exact pass reads the
    list the inference kernel wrote.  A NULL handle means different streams
    to the two libraries (the legacy default stream to the prescreen, the
    world's own stream to the engine), so with torch's default stream current
    the screen runs on a side stream ordered after it, and the default stream
    waits for the screen."""

    def __init__(self, stream):
        self.stream = stream

    def __enter__(self):
        import torch
        if self.stream:
            return self.stream
        self.cur = torch.cuda.current_stream()
        if self.cur.cuda_stream:
            self.side = None
            return self.cur.cuda_stream
        self.side = torch.cuda.Stream()
        self.side.wait_stream(self.cur)
        return self.side.cuda_stream

    def __exit__(self, *exc):
        if not self.stream and self.side is not None:
            self.cur.wait_stream(self.side)
        return False


def _outputs(out, n, device, sp):
    """idx, count and free for n items, count and free cleared on stream sp:
    the caller's buffers (out, a result of an earlier call with at least n
    items) or new ones."""
    import torch
    with torch.cuda.stream(torch.cuda.ExternalStream(sp, device=device)):
        if out is None:
            return (torch.empty(max(n, 1), dtype=torch.int32, device=device),
                    torch.zeros(1, dtype=torch.int32, device=device),
                    torch.zeros(n, dtype=torch.uint8, device=device))
        idx, count, free = out["idx"], out["count"], out["free"]
        if idx.numel() < max(n, 1) or free.numel() < n:
            raise ValueError("out holds buffers for fewer items than this call")
        count.zero_()
        free = free[:n]
        free.zero_()
        return idx, count, free


def screen_states(uw, net, d_states, threshold=0.5, stream=None, d_pfree=None, out=None):
    """The prescreen of the sampler loops (src/GP/uw_random.py:46-52,
    src/GP/pipeline/keyconf.py:127-131): the network selects the states with
    P(free) > threshold in one kernel, and the exact engine re-verifies every
    selected state (UnitWorld::isValid, lib/osr/unit_world.cc:215-247), so
    nothing the network accepts is reported free unless it is.  Everything is
    enqueued on `stream` (default: torch's current stream) without a host
    synchronisation.  d_states: float64 CUDA tensor [n, 7].  Returns a dict of
    CUDA tensors: idx (int32 [n], the first count entries are the selected
    states, in no fixed order), count (uint32 [1]) and free (uint8 [n], 1 for
    a selected state that is exactly free, else 0).  A sampler loop passes
    the previous result as `out` to reuse its buffers."""
    n = int(d_states.shape[0])
    with _one_stream(stream) as sp:
        idx, count, free = _outputs(out, n, d_states.device, sp)
        net.screen_states_device(d_states.data_ptr(), n, threshold, idx.data_ptr(), count.data_ptr(),
                                 d_pfree.data_ptr() if d_pfree is not None else None, stream=sp)
        if n:
            uw.valid_states_subset_device(d_states.data_ptr(), n, idx.data_ptr(), count.data_ptr(),
                                          free.data_ptr(), sp)
    return {"idx": idx, "count": count, "free": free}


def screen_segments(uw, net, d_from, d_to, initial_verify_delta, threshold=0.5, points=8, stream=None, out=None):
    """The prescreen of candidate edges (the visibility checks of
    src/GP/pipeline/solve2.py:164-233): a segment is selected when the
    network gives P(free) > threshold at `points` states along it, and every
    selected segment is re-verified exactly (UnitWorld::isValidTransition,
    lib/osr/unit_world.cc:477-486).  Returns idx, count and free as
    screen_states does, per segment, and reuses `out` the same way."""
    n = int(d_from.shape[0])
    with _one_stream(stream) as sp:
        idx, count, free = _outputs(out, n, d_from.device, sp)
        net.screen_segments_device(d_from.data_ptr(), d_to.data_ptr(), n, points, threshold, idx.data_ptr(),
                                   count.data_ptr(), stream=sp)
        if n:
            uw.valid_transitions_subset_device(d_from.data_ptr(), d_to.data_ptr(), n, idx.data_ptr(),
                                               count.data_ptr(), initial_verify_delta, free.data_ptr(), None, sp)
    return {"idx": idx, "count": count, "free": free}


def selected_mask(result, n):
    """Bool CUDA tensor [n]: the items a screen selected (synchronises)."""
    import torch
    k = int(result["count"].item())
    m = torch.zeros(n, dtype=torch.bool, device=result["idx"].device)
    m[result["idx"][:k].long()] = True
    return m


def recall_precision(result, truth):
    """Against exact labels truth (bool CUDA tensor [n], True = free): recall
    = free items the screen kept / free items, precision = free items among
    the selected ones."""
    n = int(truth.shape[0])
    sel = selected_mask(result, n)
    kept = result["free"].bool()
    t = int(truth.sum().item())
    s = int(sel.sum().item())
    return int(kept.sum().item()) / max(t, 1), int((sel & truth).sum().item()) / max(s, 1)


def features_reference(states):
    """Float64 features the kernel rounds to bf16: row-major rotation of the
    normalised quaternion, translation - 0.5, zero padding to 32."""
    q = np.asarray(states, np.float64).reshape(-1, 7)
    w, x, y, z = q[:, 3], q[:, 4], q[:, 5], q[:, 6]
    nn = np.sqrt((x * x + z * z) + (y * y + w * w))
    w, x, y, z = w / nn, x / nn, y / nn, z / nn
    tx, ty, tz = 2 * x, 2 * y, 2 * z
    R = np.stack([1 - (ty * y + tz * z), tx * y - tz * w, tx * z + ty * w,
                  tx * y + tz * w, 1 - (tx * x + tz * z), ty * z - tx * w,
                  tx * z - ty * w, ty * z + tx * w, 1 - (tx * x + ty * y)], axis=1)
    f = np.zeros((len(q), 32))
    f[:, :9] = R
    f[:, 9:12] = q[:, :3] - 0.5
    return f
