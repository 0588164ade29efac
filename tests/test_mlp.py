"""Importance-network oracle (CPU) and HIP kernels against it (GPU).

The oracle's backward is pinned by central finite differences of its own
forward in FP64 (a wrong index or a dropped SiLU' term fails); the bfloat16
rounding is pinned by known values.  The device forward must match the oracle
(which rounds exactly where the device does) to FP32-accumulation accuracy,
the device gradient to BF16-operand accuracy, and Adam training must reduce
the loss.
"""

import numpy as np
import pytest

from pyrite_amd.vr.mlp import N_PARAM, init_params


def _data(n, seed=3):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, size=(n, 6)).astype(np.float32)
    t = (np.sin(2 * x[:, 0]) + x[:, 1] * x[:, 2] - 0.5 * x[:, 4] ** 2).astype(np.float32)
    return x, t


def test_bf16_rounding(oracle):
    assert oracle.bf16(1.0) == 1.0
    assert oracle.bf16(1.0 + 2**-8) == 1.0  # tie -> even
    assert oracle.bf16(1.0 + 3 * 2**-8) == 1.0 + 2**-6
    assert oracle.bf16(3.14159265) == 3.140625


def test_oracle_gradient_by_finite_differences(oracle):
    # exact FP64 mode (no bfloat16 rounding) so the loss is smooth in every parameter
    lib = oracle.lib()
    lib.or_mlp_set_rounding.argtypes = [oracle.C.c_int]
    lib.or_mlp_set_rounding(0)
    try:
        rng = np.random.default_rng(0)
        p = init_params(rng).astype(np.float32)
        x, t = _data(16)
        loss, g = oracle.mlp_loss_grad(p, x, t)
        assert g.size == N_PARAM == 4673
        for i in list(range(0, N_PARAM, 97)) + [N_PARAM - 1]:
            h = np.float32(4e-3)
            pp, pm = p.copy(), p.copy()
            pp[i] += h
            pm[i] -= h
            lp, _ = oracle.mlp_loss_grad(pp, x, t)
            lm, _ = oracle.mlp_loss_grad(pm, x, t)
            fd = (lp - lm) / float(pp[i] - pm[i])
            assert g[i] == pytest.approx(fd, rel=2e-3, abs=2e-6), i
    finally:
        lib.or_mlp_set_rounding(1)


@pytest.mark.gpu
def test_device_forward_matches_oracle(oracle, device):
    import torch

    from pyrite_amd.vr.mlp import ImportanceNet

    p = init_params(np.random.default_rng(1))
    x, _ = _data(1000)
    net = ImportanceNet(p, device)
    y = net.forward(torch.as_tensor(x, device=device)).cpu().numpy()
    want = oracle.mlp_forward(p, x)
    np.testing.assert_allclose(y, want, rtol=2e-4, atol=2e-4 * np.abs(want).max())


@pytest.mark.gpu
def test_device_gradient_matches_oracle(oracle, device):
    import torch

    from pyrite_amd.vr.mlp import ImportanceNet

    p = init_params(np.random.default_rng(2))
    x, t = _data(777)
    net = ImportanceNet(p, device)
    g, loss = net.gradient(torch.as_tensor(x, device=device), torch.as_tensor(t, device=device))
    g = g.cpu().numpy()
    lo, go = oracle.mlp_loss_grad(p, x, t)
    assert float(loss.item()) == pytest.approx(lo, rel=1e-4)
    # dz2 enters the GEMMs as bfloat16 (8-bit mantissa): compare blockwise in norm
    for a, b in ((0, 384), (384, 448), (448, 4544), (4544, 4608), (4608, 4672), (4672, 4673)):
        err = np.linalg.norm(g[a:b] - go[a:b]) / max(np.linalg.norm(go[a:b]), 1e-12)
        assert err < 2e-2, (a, b, err)


@pytest.mark.gpu
def test_device_training_reduces_loss(device):
    import torch

    from pyrite_amd.vr.mlp import ImportanceNet

    x, t = _data(20000, seed=9)
    net = ImportanceNet(init_params(np.random.default_rng(4)), device)
    hist = net.fit(torch.as_tensor(x, device=device), torch.as_tensor(t, device=device),
                   epochs=15, batch=2048, lr=3e-3)
    assert hist[-1] < 0.2 * hist[0], hist
