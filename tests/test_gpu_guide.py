"""GPU learned-proposal kernels (MFMA bf16, fp32 accumulation) against the
oracle's bf16-rounded f32 network, through the C ABI."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
UNIT = np.array([0, 0, 0, 1, 1, 1], np.float32)   # bounds for already-normalised positions


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _records(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.random((n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    mu = np.array([0.2, 0.9, 0.3])
    wi = mu / np.linalg.norm(mu) + 0.4 * rng.normal(size=(n, 3))
    wi = (wi / np.linalg.norm(wi, axis=1, keepdims=True)).astype(np.float32)
    w = rng.exponential(size=n).astype(np.float32)
    w[rng.random(n) < 0.1] = 0.0
    return x, nrm, wi, w


def test_infer_matches_oracle(gpu, oracle):
    import torch
    from rs_illume_amd.guide import Guide, init_params
    _say("case guide infer")
    p = init_params(1, out_scale=1.0)
    g = Guide(p)
    assert np.array_equal(g.params(), p)
    x, nrm, _, _ = _records(1000, 2)
    inp = torch.from_numpy(np.concatenate([x, nrm], 1)).cuda()
    out = g.infer(inp).cpu().numpy()
    ref16 = oracle.guide_forward(p, x, nrm, UNIT, bf16=True)
    ref32 = oracle.guide_forward(p, x, nrm, UNIT)
    err16 = np.abs(out - ref16).max()
    err32 = np.abs(out - ref32).max()
    _say("max |gpu - oracle bf16|", err16, "max |gpu - oracle f32|", err32, "scale", np.abs(ref32).max())
    assert err16 < 2e-2 * (np.abs(ref32).max() + 1)
    assert np.corrcoef(out.ravel(), ref16.ravel())[0, 1] > 0.9999


def test_grad_matches_oracle(gpu, oracle):
    import torch
    from rs_illume_amd.guide import Guide, init_params
    _say("case guide grad")
    p = init_params(2, out_scale=1.0)
    g = Guide(p)
    x, nrm, wi, w = _records(3000, 3)
    rec = torch.from_numpy(np.concatenate([x, nrm, wi, w[:, None]], 1).astype(np.float32)).cuda()
    gg = g.grad(rec).cpu().numpy()
    _, g16 = oracle.guide_loss_grad(p, x, nrm, wi, w, UNIT, bf16=True)
    _, g64 = oracle.guide_loss_grad(p, x, nrm, wi, w, UNIT, f64=True)
    rel16 = np.linalg.norm(gg - g16) / np.linalg.norm(g16)
    rel64 = np.linalg.norm(gg - g64) / np.linalg.norm(g64)
    cos = gg @ g16 / (np.linalg.norm(gg) * np.linalg.norm(g16))
    _say("grad rel err vs oracle bf16", rel16, "vs f64", rel64, "cos", cos)
    # same bf16 roundings as the oracle's network: 3.6e-4 observed
    assert cos > 0.9999 and rel16 < 1e-2 and rel64 < 0.15
    # every layer's block of the gradient agrees in direction
    from rs_illume_amd.guide import offsets
    for w_off, b_off in offsets():
        a, b = gg[w_off:b_off], g16[w_off:b_off]
        assert a @ b / (np.linalg.norm(a) * np.linalg.norm(b)) > 0.99


def test_adam_training_lowers_loss_like_oracle(gpu, oracle):
    import torch
    from rs_illume_amd.guide import Guide, init_params
    _say("case guide train")
    p = init_params(3)
    g = Guide(p, lr=5e-3)
    x, nrm, wi, w = _records(4096, 4)
    rec = torch.from_numpy(np.concatenate([x, nrm, wi, w[:, None]], 1).astype(np.float32)).cuda()
    l0, _ = oracle.guide_loss_grad(p, x, nrm, wi, w, UNIT)
    po, m, v = p.copy(), np.zeros_like(p), np.zeros_like(p)
    for t in range(1, 21):
        g.train_step(rec)
        _, go = oracle.guide_loss_grad(po, x, nrm, wi, w, UNIT, bf16=True)
        oracle.adam(po, m, v, go, t, lr=5e-3)
    torch.cuda.synchronize()
    pg = g.params()
    lg, _ = oracle.guide_loss_grad(pg, x, nrm, wi, w, UNIT)
    lo, _ = oracle.guide_loss_grad(po, x, nrm, wi, w, UNIT)
    _say("loss start", l0, "gpu-trained", lg, "oracle-trained", lo)
    assert lg < l0 - 0.3
    assert abs(lg - lo) < 0.1 * abs(l0 - lo) + 0.05
