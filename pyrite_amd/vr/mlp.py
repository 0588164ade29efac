"""The importance network on the device: parameters, inference, training.

A 6 -> 64 -> 64 -> 1 SiLU MLP (4673 FP32 master parameters, BF16 shadows for
the MFMA operands) whose kernels live in libhiprite (``hr_mlp_*``).  It
predicts ``log(eps + yield-to-go)`` of an electron state
``x = (E/E0, ux, uy, uz, depth/thickness, tilt)``.  There is no counterpart in
the reference (PyRITE's variance reduction is detector-cone splitting,
``src/pyrite/_cone_splitting.py``); this module is the learned replacement's
network half.
"""

from __future__ import annotations

import ctypes as C

import numpy as np

from .. import _lib

N_IN = 6
HIDDEN = 64
N_PARAM = HIDDEN * N_IN + HIDDEN + HIDDEN * HIDDEN + HIDDEN + HIDDEN + 1


def _stream():
    import torch

    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


def init_params(rng: np.random.Generator) -> np.ndarray:
    """He-style initialisation of the FP32 master vector (host)."""
    p = np.zeros(N_PARAM, dtype=np.float32)
    o = 0
    W1 = rng.normal(0.0, np.sqrt(2.0 / N_IN), size=(HIDDEN, N_IN))
    p[o:o + W1.size] = W1.ravel(); o += W1.size
    o += HIDDEN  # b1 = 0
    W2 = rng.normal(0.0, np.sqrt(2.0 / HIDDEN), size=(HIDDEN, HIDDEN))
    p[o:o + W2.size] = W2.ravel(); o += W2.size
    o += HIDDEN  # b2 = 0
    w3 = rng.normal(0.0, np.sqrt(1.0 / HIDDEN), size=HIDDEN)
    p[o:o + HIDDEN] = w3; o += HIDDEN
    return p


class ImportanceNet:
    """Device-resident network with Adam state."""

    def __init__(self, params: np.ndarray, device):
        import torch

        lib = _lib.load()
        assert lib.hr_mlp_param_count() == N_PARAM
        self.device = device
        self.params = torch.as_tensor(np.ascontiguousarray(params, dtype=np.float32), device=device)
        self.w1b = torch.zeros(HIDDEN * 8, dtype=torch.int16, device=device)
        self.w2b = torch.zeros(HIDDEN * HIDDEN, dtype=torch.int16, device=device)
        self.m1 = torch.zeros(N_PARAM, dtype=torch.float32, device=device)
        self.m2 = torch.zeros(N_PARAM, dtype=torch.float32, device=device)
        self.grad = torch.zeros(N_PARAM, dtype=torch.float32, device=device)
        self.loss = torch.zeros(1, dtype=torch.float32, device=device)
        self.step_count = 0
        self._ws = None
        _lib.call("hr_mlp_shadow", self.params.data_ptr(), self.w1b.data_ptr(), self.w2b.data_ptr(), _stream())

    def _check(self, x, t=None):
        import torch

        if x.dim() != 2 or x.shape[1] != N_IN or x.dtype != torch.float32 or not x.is_contiguous():
            raise ValueError(f"x must be a contiguous float32 [n, {N_IN}] tensor, got {tuple(x.shape)} {x.dtype}")
        if x.device != self.params.device:
            raise ValueError("x is not on the network's device")
        if t is not None and (t.shape != (x.shape[0],) or t.dtype != torch.float32 or not t.is_contiguous()
                              or t.device != x.device):
            raise ValueError("t must be a contiguous float32 [n] tensor on the network's device")

    def _net(self):
        n = _lib.MlpNet()
        n.params, n.w1b, n.w2b = self.params.data_ptr(), self.w1b.data_ptr(), self.w2b.data_ptr()
        return n

    def forward(self, x, out=None):
        """``y = net(x)`` for a device ``[n, 6]`` FP32 tensor."""
        import torch

        self._check(x)
        n = int(x.shape[0])
        if out is None:
            out = torch.empty(n, dtype=torch.float32, device=self.device)
        net = self._net()
        _lib.call("hr_mlp_forward", C.byref(net), x.data_ptr(), n, out.data_ptr(), _stream())
        return out

    def gradient(self, x, t):
        """``(grad [4673], loss)`` of the MSE loss on a batch (device tensors)."""
        import torch

        self._check(x, t)
        n = int(x.shape[0])
        need = int(_lib.load().hr_mlp_workspace_bytes(n))
        if self._ws is None or self._ws.numel() * 4 < need:
            self._ws = torch.empty(need // 4 + 1, dtype=torch.float32, device=self.device)
        net = self._net()
        _lib.call("hr_mlp_grad", C.byref(net), x.data_ptr(), t.data_ptr(), n, self._ws.data_ptr(),
                  self.grad.data_ptr(), self.loss.data_ptr(), _stream())
        return self.grad, self.loss

    def train_step(self, x, t, lr=1e-3):
        self.gradient(x, t)
        self.step_count += 1
        net = self._net()
        _lib.call("hr_mlp_adam", C.byref(net), self.grad.data_ptr(), self.m1.data_ptr(), self.m2.data_ptr(),
                  float(lr), self.step_count, _stream())
        return self.loss

    def rescale_output(self, scale: float, shift: float):
        """Fold ``y -> scale * y + shift`` into the output layer (w3, b3)."""
        o3 = N_PARAM - HIDDEN - 1
        self.params[o3:o3 + HIDDEN] *= float(scale)
        self.params[N_PARAM - 1] = self.params[N_PARAM - 1] * float(scale) + float(shift)

    def fit(self, x, t, epochs=20, batch=8192, lr=2e-3, seed=0):
        """Mini-batch Adam over device tensors; returns the per-epoch mean loss."""
        import torch

        g = torch.Generator(device=self.device)
        g.manual_seed(seed)
        n = int(x.shape[0])
        hist = []
        for _ in range(epochs):
            perm = torch.randperm(n, generator=g, device=self.device)
            tot = 0.0
            nb = 0
            for b0 in range(0, n, batch):
                idx = perm[b0:b0 + batch]
                loss = self.train_step(x[idx].contiguous(), t[idx].contiguous(), lr)
                tot += float(loss.item())
                nb += 1
            hist.append(tot / max(nb, 1))
        return hist
