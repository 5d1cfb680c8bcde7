"""Pins the oracle's learned proposal: finite-difference gradients in f64, the
vMF mixture's normalisation and sampling against its own density, Adam's
first step, and the bf16-rounded forward against f32."""
import numpy as np
import pytest

BOUNDS = np.array([-1, -1, -1, 1, 1, 1], np.float32)


def init_params(seed=0, out_scale=0.1):
    from rs_illume_amd.guide import init_params as ip
    return ip(seed, out_scale)


def records(n, seed=1):
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    wi = rng.normal(size=(n, 3))
    wi = (wi / np.linalg.norm(wi, axis=1, keepdims=True)).astype(np.float32)
    w = rng.exponential(size=n).astype(np.float32)
    return x, nrm, wi, w


def test_param_count(oracle):
    assert oracle._guide_decl().orc_guide_num_params() == 20200


def test_encoding_is_sin_cos_of_power_of_two_octaves(oracle):
    """Feature 24 i + 2 j = sin(2^j pi v_i), + 1 = cos(2^j pi v_i), with v the
    position normalised by the bounds and the raw normal: checked against the
    f64 closed form (the oracle reduces the angle exactly, so the only error
    is the final rounding to f32)."""
    x, nrm, _, _ = records(500, seed=3)
    f = oracle.guide_encode(x, nrm, BOUNDS)
    # the normalised position as the oracle rounds it (f32), then f64 angles
    vp = (x - np.float32(-1)) / np.float32(2)
    v = np.concatenate([vp.astype(np.float64), nrm.astype(np.float64)], 1)
    for i in range(6):
        for j in range(12):
            ang = (2.0 ** j) * np.pi * v[:, i]
            assert np.abs(f[:, 24 * i + 2 * j] - np.sin(ang)).max() < 6e-8
            assert np.abs(f[:, 24 * i + 2 * j + 1] - np.cos(ang)).max() < 6e-8


def test_gradient_matches_finite_differences(oracle):
    p = init_params(3, out_scale=1.0)
    x, nrm, wi, w = records(16, 2)
    loss, g = oracle.guide_loss_grad(p, x, nrm, wi, w, BOUNDS, f64=True)
    rng = np.random.default_rng(0)
    # probe parameters of every layer: weights and biases
    idx = np.concatenate([rng.integers(0, 9216, 12), rng.integers(9216, 9280, 4), rng.integers(9280, 17600, 12),
                          rng.integers(17600, 20200, 16)])
    p64 = p.astype(np.float64)
    for i in idx:
        h = 1e-4 * max(1.0, abs(p64[i]))
        pp, pm = p64.copy(), p64.copy()
        pp[i] += h
        pm[i] -= h
        fd = (oracle.guide_loss_f64(pp, x, nrm, wi, w, BOUNDS) - oracle.guide_loss_f64(pm, x, nrm, wi, w, BOUNDS)) / (2 * h)
        assert abs(fd - g[i]) <= 1e-4 * max(1.0, abs(fd)) + 1e-6, (i, fd, g[i])


def test_f32_and_bf16_forward_close_to_f64_grad(oracle):
    p = init_params(4, out_scale=1.0)
    x, nrm, wi, w = records(64, 5)
    _, g64 = oracle.guide_loss_grad(p, x, nrm, wi, w, BOUNDS, f64=True)
    _, g32 = oracle.guide_loss_grad(p, x, nrm, wi, w, BOUNDS)
    _, g16 = oracle.guide_loss_grad(p, x, nrm, wi, w, BOUNDS, bf16=True)
    rel = lambda a, b: np.linalg.norm(a - b) / np.linalg.norm(b)
    assert rel(g32, g64) < 1e-3   # f32 sin of 2^11*pi*x arguments
    assert rel(g16, g64) < 1e-1   # bf16 weights and activations (8-bit mantissa)
    o32 = oracle.guide_forward(p, x, nrm, BOUNDS)
    o16 = oracle.guide_forward(p, x, nrm, BOUNDS, bf16=True)
    assert np.abs(o16 - o32).max() < 0.05 * (np.abs(o32).max() + 1)


def test_vmf_mixture_normalised_and_sampled_by_its_density(oracle):
    rng = np.random.default_rng(7)
    outs = (rng.normal(size=(1, 40)) * 1.5).astype(np.float32)
    outs[0, 4::5] = rng.uniform(0, 4, 8)   # concentrations e^0 .. e^4
    # normalisation: quadrature over the sphere
    n = 400
    th = (np.arange(n) + 0.5) / n * np.pi
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    pdf = oracle.guide_pdf(np.repeat(outs, len(d), 0), d)
    integral = np.sum(pdf * np.sin(T).reshape(-1)) * (np.pi / n) * (2 * np.pi / (2 * n))
    assert abs(integral - 1) < 2e-3
    # sampling: E_sample[f] == integral f p dw for f = direction and f = indicator of a cap
    m = 200000
    u = rng.random((m, 3)).astype(np.float32)
    wi, spdf = oracle.guide_sample(np.repeat(outs, m, 0), u)
    assert np.allclose(np.linalg.norm(wi, axis=1), 1, atol=1e-5)
    assert np.allclose(spdf, oracle.guide_pdf(np.repeat(outs, m, 0), wi), rtol=1e-4)
    mean_q = (pdf[:, None] * d * np.sin(T).reshape(-1, 1)).sum(0) * (np.pi / n) * (2 * np.pi / (2 * n))
    assert np.allclose(wi.mean(0), mean_q, atol=1e-2)
    cap = d[:, 2] > 0.5
    p_cap = np.sum((pdf * np.sin(T).reshape(-1))[cap]) * (np.pi / n) * (2 * np.pi / (2 * n))
    assert abs((wi[:, 2] > 0.5).mean() - p_cap) < 5e-3


def test_adam_first_step_is_signed_lr(oracle):
    rng = np.random.default_rng(8)
    p = rng.normal(size=100).astype(np.float32)
    g = rng.normal(size=100).astype(np.float32)
    m = np.zeros(100, np.float32)
    v = np.zeros(100, np.float32)
    p0 = p.copy()
    oracle.adam(p, m, v, g, 1, lr=1e-3)
    assert np.allclose(p, p0 - 1e-3 * np.sign(g), atol=1e-7)
    # second step follows the bias-corrected moments
    g2 = rng.normal(size=100).astype(np.float32)
    p1 = p.copy()
    oracle.adam(p, m, v, g2, 2, lr=1e-3)
    mh = (0.9 * 0.1 * g + 0.1 * g2) / (1 - 0.81)
    vh = (0.999 * 0.001 * g ** 2 + 0.001 * g2 ** 2) / (1 - 0.999 ** 2)
    assert np.allclose(p, p1 - 1e-3 * mh / (np.sqrt(vh) + 1e-8), atol=1e-6)


def test_training_reduces_loss_on_a_fixed_distribution(oracle):
    """Fitting samples drawn from one direction lobe lowers the NLL."""
    p = init_params(5)
    rng = np.random.default_rng(9)
    n = 256
    x, nrm, _, _ = records(n, 10)
    mu = np.array([0.3, 0.8, 0.5]) / np.linalg.norm([0.3, 0.8, 0.5])
    wi = mu + 0.2 * rng.normal(size=(n, 3))
    wi = (wi / np.linalg.norm(wi, axis=1, keepdims=True)).astype(np.float32)
    w = np.ones(n, np.float32)
    m = np.zeros_like(p)
    v = np.zeros_like(p)
    l0, _ = oracle.guide_loss_grad(p, x, nrm, wi, w, BOUNDS)
    for t in range(1, 31):
        _, g = oracle.guide_loss_grad(p, x, nrm, wi, w, BOUNDS)
        oracle.adam(p, m, v, g, t, lr=5e-3)
    l1, _ = oracle.guide_loss_grad(p, x, nrm, wi, w, BOUNDS)
    assert l1 < l0 - 1.0, (l0, l1)


def _sphere_grid(n=400):
    th = (np.arange(n) + 0.5) / n * np.pi
    ph = (np.arange(2 * n) + 0.5) / (2 * n) * 2 * np.pi
    T, P = np.meshgrid(th, ph, indexing="ij")
    d = np.stack([np.sin(T) * np.cos(P), np.sin(T) * np.sin(P), np.cos(T)], -1).reshape(-1, 3)
    dw = (np.sin(T) * (np.pi / n) * (2 * np.pi / (2 * n))).reshape(-1)
    return d, dw


def test_product_with_a_bsdf_lobe(oracle):
    """The mixture times a vMF lobe (the BSDF-aware proposal): its pdf
    integrates to 1, is proportional to mixture pdf x lobe density at every
    direction (a wrong product scale breaks the proportionality, a wrong
    product mean the normalisation), and its sampler follows it."""
    rng = np.random.default_rng(17)
    outs = (rng.normal(size=(1, 40)) * 1.5).astype(np.float32)
    outs[0, 4::5] = rng.uniform(0, 4, 8)
    d, dw = _sphere_grid()
    for mu_b, k_b in (((0.0, 0.0, 1.0), 1.4), ((0.6, 0.0, 0.8), 9.0), ((-0.3, 0.9, 0.316), 40.0)):
        mu_b = np.array(mu_b, np.float32) / np.linalg.norm(mu_b)
        rows = np.repeat(outs, len(d), 0)
        mus, ks = np.repeat(mu_b[None], len(d), 0), np.full(len(d), k_b, np.float32)
        prod = oracle.guide_product_pdf(rows, mus, ks, d).astype(np.float64)
        assert abs(np.sum(prod * dw) - 1) < 3e-3, (k_b, np.sum(prod * dw))
        lobe = k_b / (2 * np.pi * (1 - np.exp(-2 * k_b))) * np.exp(k_b * (d @ mu_b - 1))
        base = oracle.guide_pdf(rows, d).astype(np.float64) * lobe
        big = base > 1e-3 * base.max()
        ratio = prod[big] / base[big]
        assert np.allclose(ratio, ratio.mean(), rtol=1e-4), (k_b, ratio.min(), ratio.max())
        m = 200000
        wi, spdf = oracle.guide_product_sample(np.repeat(outs, m, 0), np.repeat(mu_b[None], m, 0),
                                               np.full(m, k_b, np.float32), rng.random((m, 3)).astype(np.float32))
        assert np.allclose(np.linalg.norm(wi, axis=1), 1, atol=1e-5)
        assert np.allclose(spdf, oracle.guide_product_pdf(np.repeat(outs, m, 0), np.repeat(mu_b[None], m, 0),
                                                          np.full(m, k_b, np.float32), wi), rtol=1e-3)
        mean_q = (prod[:, None] * d * dw[:, None]).sum(0)
        assert np.allclose(wi.mean(0), mean_q, atol=1e-2), (wi.mean(0), mean_q)
        cap = d @ mu_b > 0.9
        assert abs((wi @ mu_b > 0.9).mean() - np.sum(prod[cap] * dw[cap])) < 5e-3
