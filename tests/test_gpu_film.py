"""Reconstruction filters on the GPU film (illume_filter_*, illume_render_filtered,
illume_film_resolve) against the oracle: identical weight tables, filtered
films at equal sampler streams, the box filter of radius 0.5 against the
unfiltered film, and a constant-radiance image resolving to 1."""
import sys

import numpy as np
import pytest

from rs_illume_amd import scenes as S
from rs_illume_amd.film import Filter, filter_desc, render_filtered, resolve

pytestmark = pytest.mark.gpu
REL_MSE_BAR = 1e-4
# per-pixel sums of a few float atomics arrive in varying order
SUM_RTOL = 2e-6
FILTERS = [("box", {}), ("triangle", {}), ("gaussian", {}), ("mitchell", {}), ("sinc", {}),
           ("gaussian", dict(radius=(1.0, 3.0), alpha=0.7)), ("mitchell", dict(radius=(2.0, 1.5), b=0.2, c=0.4)),
           ("triangle", dict(radius=0.75))]


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("kind,kw", FILTERS)
def test_tables_match_oracle(gpu, oracle, kind, kw):
    f = Filter(kind, **kw)
    assert np.array_equal(f.table(), oracle.filter_table(filter_desc(kind, **kw)))


def _gpu_rgbw(desc, cam, filt, spp, depth, seed):
    import torch
    from rs_illume_amd.render import PathIntegrator, Scene
    pi = PathIntegrator(cam, spp=spp, max_depth=depth, seed=seed)
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    rgbw = render_filtered(pi, Scene(desc), filt, ray_counts=counts)
    torch.cuda.synchronize()
    return rgbw, counts.cpu().numpy()


@pytest.mark.parametrize("kind,kw", FILTERS[1:6])
def test_filtered_render_matches_oracle(gpu, oracle, kind, kw):
    desc, cam = S.cornell_box(0, glossy=True, light_size=0.2, light_L=26.5625), S.cornell_camera(40, 32)
    _say("case filtered", kind, kw)
    rgbw, gc = _gpu_rgbw(desc, cam, Filter(kind, **kw), 4, 6, 2)
    o, oc = oracle.OracleRenderScene(desc.arrays()).render_filtered(cam, filter_desc(kind, **kw), 4, 6, seed=2)
    g = rgbw.cpu().numpy()
    # weights are the same per sample; only the order of the additions differs
    assert np.allclose(g[..., 3], o[..., 3], rtol=1e-5, atol=1e-5)
    gi = resolve(rgbw).cpu().numpy()
    w = o[..., 3:4]
    oi = np.where(w != 0, np.maximum(o[..., :3] * (np.float32(1) / np.where(w != 0, w, 1)), 0), o[..., :3])
    rm = S.rel_mse(gi, oi)
    _say(kind, "relMSE", rm, "rays", gc, oc)
    assert np.isfinite(g).all() and rm <= REL_MSE_BAR, rm
    assert abs(int(gc[0]) - int(oc[0])) <= 0.002 * oc[0] + 10


def test_box_half_pixel_equals_unfiltered(gpu):
    import torch
    from rs_illume_amd.render import PathIntegrator, Scene
    desc, cam = S.cornell_box(0), S.cornell_camera(48, 40)
    _say("case box 0.5")
    rgbw, _ = _gpu_rgbw(desc, cam, Filter("box"), 8, 5, 4)
    rgb = PathIntegrator(cam, spp=8, max_depth=5, seed=4).render_sum(Scene(desc))
    torch.cuda.synchronize()
    g = rgbw.cpu().numpy()
    assert np.all(g[..., 3] == 8)
    assert np.allclose(g[..., :3], rgb.cpu().numpy(), rtol=SUM_RTOL, atol=0)


@pytest.mark.parametrize("kind", ["box", "gaussian", "triangle"])
def test_constant_radiance_resolves_to_one(gpu, kind):
    cam = S.camera((0.1, -0.2, 0.05), (0.3, 0.1, -1), (0, 1, 0), 70, 33, 21)
    _say("case constant", kind)
    rgbw, _ = _gpu_rgbw(S.furnace_box(0.5, 1.0), cam, Filter(kind), 3, 0, 1)
    img = resolve(rgbw).cpu().numpy()
    assert np.allclose(img, 1.0, rtol=4e-6, atol=0)


def test_resolve(gpu):
    import torch
    rng = np.random.default_rng(0)
    a = rng.normal(size=(17, 23, 4)).astype(np.float32)
    a[::3, ::2, 3] = 0
    out = resolve(torch.from_numpy(a).cuda(), scale=2.5).cpu().numpy()
    w = a[..., 3:4]
    inv = np.float32(1) / np.where(w != 0, w, np.float32(1))
    expect = np.where(w != 0, np.maximum(a[..., :3] * inv, np.float32(0)), a[..., :3]) * np.float32(2.5)
    assert np.array_equal(out, expect.astype(np.float32))


def test_invalid_filters_raise(gpu):
    from rs_illume_amd._lib import IllumeError
    for kind, kw in (("box", dict(radius=0.0)), ("gaussian", dict(radius=33.0)), ("sinc", dict(tau=0.0)),
                     ("mitchell", dict(b=float("nan")))):
        with pytest.raises(IllumeError):
            Filter(kind, **kw)


def test_gathered_film_is_deterministic_and_chunked(gpu):
    """The megakernel's filtered film is gathered per pixel in a fixed order:
    two renders give the same bits, and a render of more than 2^24 samples
    (filmed in chunks of whole samples per pixel) gives the bits of the same
    samples rendered in two calls."""
    import torch
    from rs_illume_amd.render import PathIntegrator, Scene
    desc, cam = S.cornell_box(0), S.cornell_camera(250, 250)
    sc = Scene(desc)
    f = Filter("gaussian")
    pi = PathIntegrator(cam, spp=300, max_depth=3, seed=6)
    n_pix = 254 * 254   # sample bounds of the radius-2 filter: [-2, 252) on both axes
    chunk = (1 << 24) // n_pix   # whole samples per pixel in one chunk
    assert chunk < 300
    _say("case gathered film, chunks of", chunk)
    a = render_filtered(pi, sc, f)
    b = render_filtered(pi, sc, f)
    c = render_filtered(pi, sc, f, spp=chunk)
    render_filtered(pi, sc, f, rgbw=c, spp=300 - chunk, sample_offset=chunk)
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert torch.equal(a, c)
    small = render_filtered(PathIntegrator(cam, spp=2, max_depth=3, seed=6), sc, f)
    torch.cuda.synchronize()
    assert torch.equal(small, render_filtered(PathIntegrator(cam, spp=2, max_depth=3, seed=6), sc, f))
