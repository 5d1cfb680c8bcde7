"""AOIntegrator (integrators/ao.rs:45-93) on the GPU through illume_render_ao:
equal to the oracle at equal sampler streams (relMSE <= 1e-4, ray counts
within 0.2%), and the closed forms of tests/test_oracle_ao.py on the GPU."""
import sys

import numpy as np
import pytest

from rs_illume_amd import scenes as S

pytestmark = pytest.mark.gpu
REL_MSE_BAR = 1e-4


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _gpu_ao(desc, cam, spp, n, cos_sample, seed, sampler="random"):
    import torch
    from rs_illume_amd.render import AOIntegrator, Scene
    sc = Scene(desc)
    ao = AOIntegrator(cam, spp=spp, n_samples=n, cos_sample=cos_sample, seed=seed, sampler=sampler)
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    img = ao.render_sum(sc, ray_counts=counts)
    torch.cuda.synchronize()
    out = img.cpu().numpy(), counts.cpu().numpy()
    sc.close()
    return out


def _glass_box():
    desc = S.cornell_box(3)
    g = desc.add_material(S.glass(1.5))
    desc.sphere_material = [g, g]
    return desc


@pytest.mark.parametrize("cos_sample", [True, False])
@pytest.mark.parametrize("name", ["cornell", "glass", "terrain", "halton"])
def test_ao_matches_oracle(gpu, oracle, name, cos_sample):
    sampler = "random"
    if name == "cornell":
        desc, cam = S.cornell_box(0), S.cornell_camera(64, 64)
    elif name == "glass":
        desc, cam = _glass_box(), S.cornell_camera(48, 48)
    elif name == "terrain":
        desc = S.terrain_scene(0, n=24, n_spheres=40, height=0.35)
        cam = S.camera((1.6, 1.0, 1.2), (0, 0.1, 0), (0, 1, 0), 45.0, 48, 40)
    else:
        desc, cam, sampler = S.cornell_box(1), S.cornell_camera(40, 40), "halton"
    spp, n, seed = 4, 16, 11
    _say("case AO", name, "cos" if cos_sample else "uniform")
    g, gc = _gpu_ao(desc, cam, spp, n, cos_sample, seed, sampler)
    o, oc = oracle.OracleRenderScene(desc.arrays()).render_ao(cam, spp, n, cos_sample, seed=seed, sampler=sampler)
    rm = S.rel_mse(g / spp, o / spp)
    _say(name, "relMSE", rm, "gpu rays", gc, "oracle rays", oc)
    assert np.isfinite(g).all()
    assert rm <= REL_MSE_BAR, rm
    assert abs(int(gc[0]) - int(oc[0])) <= 0.002 * oc[0] + 10
    assert abs(int(gc[1]) - int(oc[1])) <= 0.002 * oc[1] + 10


def test_ao_closed_forms_gpu(gpu):
    size = 500.0
    s = S.SceneDesc()
    m = s.add_material(S.matte((0.5, 0.5, 0.5)))
    s.add_quad((-size, 0, -size), (size, 0, -size), (size, 0, size), (-size, 0, size), m, normal=(0, 1, 0))
    cam = S.camera((0.3, 2.0, 0.1), (0, 0, 0), (0, 0, 1), 60.0, 16, 16)
    img, gc = _gpu_ao(s, cam, 4, 8, True, 1)
    assert np.allclose(img / 4, np.pi, rtol=2e-6), (img.min() / 4, img.max() / 4)
    assert gc[0] == 16 * 16 * 4 and gc[1] == 16 * 16 * 4 * 8
    # closed box: every shadow ray occluded
    cam2 = S.camera((0.1, -0.2, 0.05), (0.3, 0.1, -1), (0, 1, 0), 70, 16, 16)
    img2, _ = _gpu_ao(S.furnace_box(0.5, 0.0), cam2, 4, 16, True, 3)
    assert np.all(img2 == 0.0)
    # under a sphere of radius r at height h: pi (1 - r^2 / h^2)
    r, h = 0.5, 1.0
    s.add_sphere((0.0, h, 0.0), r, m)
    cam3 = S.camera((5.0, 0.02, 0.0), (0, 0, 0), (0, 1, 0), 0.01, 4, 4)
    img3, _ = _gpu_ao(s, cam3, 64, 64, True, 4)
    est = img3.astype(np.float64).mean() / 64
    expect = np.pi * (1 - r * r / (h * h))
    assert abs(est - expect) < 0.01 * expect, (est, expect)
