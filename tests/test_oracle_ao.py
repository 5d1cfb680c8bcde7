"""The oracle's AOIntegrator (integrators/ao.rs:45-93, the reference's only
implemented Li) pinned by closed forms. The estimator sums
(wi . n) / (pdf * n_samples) over unoccluded hemisphere directions, so it
estimates the cosine-weighted visible solid angle: pi for an unoccluded point.

* an unoccluded plane: every cosine-sampled direction contributes exactly
  pi / n_samples, so every pixel is pi (to rounding); uniform sampling has the
  same mean;
* inside a closed box every shadow ray is occluded: 0;
* under a sphere of radius r whose centre is at height h above the point, the
  sphere hides a cone of half-angle asin(r / h) about the normal, so the
  value is pi (1 - r^2 / h^2);
* beside an infinite wall the whole half of the hemisphere facing it is
  hidden: pi / 2;
* cosine and uniform sampling estimate the same integral on the Cornell box.
"""
import numpy as np

from rs_illume_amd import scenes as S


def _plane(size=500.0):
    s = S.SceneDesc()
    m = s.add_material(S.matte((0.5, 0.5, 0.5)))
    s.add_quad((-size, 0, -size), (size, 0, -size), (size, 0, size), (-size, 0, size), m, normal=(0, 1, 0))
    return s, m


def _grazing_camera(w=4, h=4):
    # a camera low over the plane, looking at the origin with a tiny field of
    # view: every pixel sees the plane within 1e-3 of the origin
    return S.camera((5.0, 0.02, 0.0), (0, 0, 0), (0, 1, 0), 0.01, w, h)


def test_unoccluded_plane_is_pi(oracle):
    s, _ = _plane()
    cam = S.camera((0.3, 2.0, 0.1), (0, 0, 0), (0, 0, 1), 60.0, 16, 16)
    osc = oracle.OracleRenderScene(s.arrays())
    img, st = osc.render_ao(cam, 4, 8, cos_sample=True, seed=1)
    assert np.allclose(img / 4, np.pi, rtol=2e-6), (img.min() / 4, img.max() / 4)
    assert st[0] == 16 * 16 * 4 and st[1] == 16 * 16 * 4 * 8
    img_u, _ = osc.render_ao(cam, 64, 16, cos_sample=False, seed=2)
    assert abs(img_u.mean() / 64 - np.pi) < 0.01 * np.pi


def test_closed_box_is_zero(oracle):
    cam = S.camera((0.1, -0.2, 0.05), (0.3, 0.1, -1), (0, 1, 0), 70, 16, 16)
    osc = oracle.OracleRenderScene(S.furnace_box(0.5, 0.0).arrays())
    img, _ = osc.render_ao(cam, 4, 16, cos_sample=True, seed=3)
    assert np.all(img == 0.0)


def test_under_a_sphere(oracle):
    r, h = 0.5, 1.0
    s, m = _plane()
    s.add_sphere((0.0, h, 0.0), r, m)
    osc = oracle.OracleRenderScene(s.arrays())
    spp, n = 64, 64
    img, _ = osc.render_ao(_grazing_camera(), spp, n, cos_sample=True, seed=4)
    est = img.astype(np.float64).mean() / spp
    expect = np.pi * (1 - r * r / (h * h))
    # binomial noise over 4 x 4 x 64 x 64 directions: 0.23% of the value
    assert abs(est - expect) < 0.01 * expect, (est, expect)


def test_beside_an_infinite_wall(oracle):
    s, m = _plane()
    # a wall at x = -1 spanning y in [0, 500], z in [-500, 500], facing +x
    s.add_quad((-1, 0, -500), (-1, 500, -500), (-1, 500, 500), (-1, 0, 500), m, normal=(1, 0, 0))
    osc = oracle.OracleRenderScene(s.arrays())
    spp, n = 64, 64
    img, _ = osc.render_ao(_grazing_camera(), spp, n, cos_sample=True, seed=5)
    est = img.astype(np.float64).mean() / spp
    assert abs(est - np.pi / 2) < 0.01 * np.pi / 2, est


def test_cosine_and_uniform_agree_on_cornell(oracle):
    cam = S.cornell_camera(32, 32)
    osc = oracle.OracleRenderScene(S.cornell_box(0).arrays())
    a, _ = osc.render_ao(cam, 64, 16, cos_sample=True, seed=6)
    b, _ = osc.render_ao(cam, 64, 16, cos_sample=False, seed=7)
    ma, mb = a.mean() / 64, b.mean() / 64
    assert 0.5 < ma < np.pi
    assert abs(ma - mb) < 0.01 * ma, (ma, mb)
    # AO is grey
    assert np.array_equal(a[..., 0], a[..., 1]) and np.array_equal(a[..., 0], a[..., 2])
