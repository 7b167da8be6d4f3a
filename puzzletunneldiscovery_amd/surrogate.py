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
        batch = min(batch, self.max_batch)
        g = torch.Generator(device="cuda")
        g.manual_seed(seed)
        stream = torch.cuda.current_stream().cuda_stream
        losses = []
        for ep in range(epochs):
            perm = torch.randperm(n, device="cuda", generator=g)
            xs = d_states[perm].contiguous()
            ys = d_labels[perm].contiguous()
            tot, cnt = 0.0, 0
            for off in range(0, n - batch + 1, batch):
                want = (off // batch) % 16 == 0
                l = self.train_step_device(xs[off:].data_ptr(), ys[off:].data_ptr(), batch, lr, want, stream)
                if want:
                    tot += l
                    cnt += 1
            losses.append(tot / max(cnt, 1))
            if log:
                log(f"epoch {ep}: loss {losses[-1]:.4f}")
        torch.cuda.synchronize()
        return losses


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
