"""Wavefront path tracer and the learned proposal on the GPU.

* the wavefront pipeline without a proposal reproduces the megakernel image;
* a proposal whose every mixture is invalid (NaN weights) falls back to BSDF
  sampling per vertex and reproduces the plain image at equal streams;
* a guided render is unbiased: its mean matches the plain render's;
* online training on the glossy small-light box lowers the error at equal
  sample count (the gain of the proposal over the method without it).
"""
import sys

import numpy as np
import pytest

from rs_illume_amd import scenes as S

pytestmark = pytest.mark.gpu


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _glossy():
    return S.cornell_box(0, glossy=True, light_size=0.2, light_L=26.5625)


def _terrain_camera(w, h):
    """Looks down onto the terrain scenes' [-1, 1]^2 heightfield."""
    return S.camera((1.6, 1.0, 1.2), (0, 0.1, 0), (0, 1, 0), 45.0, w, h)


def test_wavefront_equals_megakernel(gpu):
    import torch
    from rs_illume_amd.render import PathIntegrator, Scene, WavefrontPathIntegrator
    _say("case wavefront == megakernel")
    # the last scene has 2 * 47^2 + 200 = 4,618 primitives: at 4,096 or more the
    # extend stage runs its lane-refilling kernel instead of one path per thread
    big = S.terrain_scene(1, n=48, n_spheres=200)
    assert len(big.idx) + len(big.spheres) >= 4096
    box_cam, terrain_cam = S.cornell_camera(48, 48), _terrain_camera(48, 48)
    for desc, depth, cam in ((S.cornell_box(0), 5, box_cam), (_glossy(), 8, box_cam),
                             (S.terrain_scene(0, n=24, n_spheres=40), 6, terrain_cam), (big, 6, terrain_cam)):
        _say("scene", len(desc.idx), "triangles", len(desc.spheres), "spheres, depth", depth)
        sc = Scene(desc)
        mk_sum = PathIntegrator(cam, spp=8, max_depth=depth, seed=5).render_sum(sc)
        wf = WavefrontPathIntegrator(sc, cam, spp=8, max_depth=depth, seed=5, max_paths=5000)
        c = torch.zeros(2, dtype=torch.int64, device="cuda")
        wf_sum = wf.render_sum(ray_counts=c)
        torch.cuda.synchronize()
        a, b = mk_sum.cpu().numpy(), wf_sum.cpu().numpy()
        _say("max abs diff", np.abs(a - b).max(), "rays", c.cpu().numpy())
        assert np.allclose(a, b, rtol=1e-5, atol=1e-5)


def test_batched_jobs_equal_individual_renders(gpu):
    """Scenes and cameras of different sizes in one pool (path regeneration
    across jobs) give each job the image its own render gives; graph replay
    and direct launches agree."""
    from rs_illume_amd.render import PathIntegrator, Scene, Wavefront
    _say("case batch")
    descs = [S.cornell_box(0), S.cornell_box(1, glossy=True, light_size=0.2, light_L=26.5625), S.cornell_box(2)]
    g = descs[2].add_material(S.glass(1.5))
    descs[2].sphere_material = [g, g]
    scenes = [Scene(d) for d in descs]
    cams = [S.cornell_camera(40, 40), S.cornell_camera(24, 56), S.cornell_camera(64, 32)]
    jobs = [(scenes[0], cams[0], 6, 5, 1), (scenes[1], cams[1], 3, 8, 2), (scenes[2], cams[2], 9, 8, 3)]
    # with a job of 4,096 or more primitives the whole batch extends through
    # the lane-refilling kernel
    big = Scene(S.terrain_scene(1, n=48, n_spheres=200))
    jobs_big = jobs + [(big, _terrain_camera(32, 40), 4, 6, 4)]
    for graph_iters, block, regen_min, js in ((8, 128, None, jobs), (0, 64, 700, jobs), (3, 256, 1, jobs),
                                              (8, 128, None, jobs_big), (0, 64, 1, jobs_big)):
        wf = Wavefront(max_paths=4096, max_depth=8, block=block, regen_min=regen_min, graph_iters=graph_iters)
        film, offs = wf.render_batch(js)
        film = film.cpu().numpy()
        _say("graph_iters", graph_iters, "block", block, "jobs", len(js), "iterations", wf.last_iterations())
        for (sc, cam, spp, depth, seed), off in zip(js, offs):
            w, h = int(cam["width"][0]), int(cam["height"][0])
            ref = PathIntegrator(cam, spp=spp, max_depth=depth, seed=seed).render_sum(sc).cpu().numpy()
            got = film[off:off + w * h * 3].reshape(h, w, 3)
            assert np.allclose(ref, got, rtol=1e-5, atol=1e-5), np.abs(ref - got).max()


@pytest.mark.parametrize("product", [False, True])
@pytest.mark.parametrize("scene", ["box", "terrain"])
def test_nan_proposal_falls_back_to_plain(gpu, product, scene):
    """The terrain scene (4,618 primitives) takes the lane-refilling extend
    kernel, which forms the guided vertices' proposal inputs in its second
    pass."""
    from rs_illume_amd.guide import Guide, N_PARAMS
    from rs_illume_amd.render import Scene, WavefrontPathIntegrator
    _say("case NaN proposal, product", product, "scene", scene)
    if scene == "box":
        sc, cam = Scene(_glossy()), S.cornell_camera(32, 32)
    else:
        sc, cam = Scene(S.terrain_scene(1, n=48, n_spheres=200)), _terrain_camera(32, 32)
    plain = WavefrontPathIntegrator(sc, cam, spp=4, max_depth=8, seed=2).render_sum().cpu().numpy()
    g = Guide(np.full(N_PARAMS, np.nan, np.float32))
    guided = WavefrontPathIntegrator(sc, cam, spp=4, max_depth=8, seed=2, guide=g, guide_frac=0.5,
                                     product=product).render_sum()
    guided = guided.cpu().numpy()
    assert np.allclose(plain, guided, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("product", [False, True])
@pytest.mark.parametrize("scene", ["box", "terrain"])
def test_guided_render_is_unbiased(gpu, product, scene):
    """An untrained proposal (an arbitrary mixture per vertex), with or
    without the product with the BSDF lobe, at fraction 0.5: the image mean
    equals the plain tracer's, so the pdf the estimator divides by is the
    density the directions were drawn from."""
    from rs_illume_amd.guide import Guide
    from rs_illume_amd.render import Scene, WavefrontPathIntegrator
    _say("case guided unbiased, product", product, "scene", scene)
    if scene == "box":
        sc, cam = Scene(_glossy()), S.cornell_camera(32, 32)
    else:
        sc, cam = Scene(S.terrain_scene(1, n=48, n_spheres=200)), _terrain_camera(32, 32)
    spp = 1024
    plain = WavefrontPathIntegrator(sc, cam, spp=spp, max_depth=8, seed=11).render().cpu().numpy()
    g = Guide(seed=4)
    guided = WavefrontPathIntegrator(sc, cam, spp=spp, max_depth=8, seed=12, guide=g, guide_frac=0.5,
                                     product=product).render()
    guided = guided.cpu().numpy()
    m_p, m_g = plain.mean(), guided.mean()
    _say("mean plain", m_p, "mean guided (untrained proposal)", m_g, "relMSE between", S.rel_mse(guided, plain))
    assert abs(m_g - m_p) < 0.02 * m_p
    assert np.isfinite(guided).all()


def test_training_gains_over_plain_at_equal_spp(gpu):
    """What the trained proposal buys: a room lit only indirectly (the light
    faces the ceiling), equal spp, relMSE against an 8192-spp plain render.
    The proposal is trained online by the renderer itself (16 Adam steps per
    graph replay, batch 16384), then frozen for the measured render."""
    from rs_illume_amd.guide import Guide
    from rs_illume_amd.render import Scene, WavefrontPathIntegrator
    _say("case guided gain")
    sc = Scene(S.cornell_box(0, light_up=True, light_L=60.0))
    cam = S.cornell_camera(64, 64)
    ref = WavefrontPathIntegrator(sc, cam, spp=8192, max_depth=8, seed=100, max_paths=1 << 22).render()
    ref = ref.cpu().numpy()
    spp = 64
    plain = WavefrontPathIntegrator(sc, cam, spp=spp, max_depth=8, seed=7).render().cpu().numpy()
    g = Guide(seed=1, lr=1e-3)
    # training passes (samples discarded), then a guided render at equal spp
    trainer = WavefrontPathIntegrator(sc, cam, spp=16, max_depth=8, seed=50, max_paths=1 << 16, guide=g,
                                      guide_frac=0.5, train=16, train_batch=16384)
    for k in range(64):
        trainer.render_sum(sample_offset=16 * k)
    steps, loss = g.stats()
    guided = WavefrontPathIntegrator(sc, cam, spp=spp, max_depth=8, seed=7, guide=g, guide_frac=0.5).render()
    guided = guided.cpu().numpy()
    e_p, e_g = S.rel_mse(plain, ref), S.rel_mse(guided, ref)
    _say("adam steps", steps, "last loss", loss, "relMSE plain", e_p, "guided (trained)", e_g, "ratio", e_g / e_p)
    assert steps > 100
    assert np.isfinite(guided).all()
    # Sanity bound on the mean, not the unbiasedness test: at 64 spp this
    # indirectly lit room's estimates are heavy-tailed, so the plain render's
    # mean sits 3.0% below the 8192-spp reference and the guided render's
    # 2.0-2.7% below it across training runs (the order of the GPU's training
    # updates is not deterministic).
    assert abs(guided.mean() - ref.mean()) < 0.05 * ref.mean()
    assert e_g < 0.5 * e_p


def test_trained_product_proposal_unbiased(gpu):
    """A proposal trained online on the glossy small-light box with the
    BSDF-lobe product, then frozen: at high spp its image mean equals the
    plain tracer's within 1%."""
    from rs_illume_amd.guide import Guide
    from rs_illume_amd.render import Scene, WavefrontPathIntegrator
    _say("case trained product proposal unbiased")
    sc = Scene(_glossy())
    cam = S.cornell_camera(48, 48)
    g = Guide(seed=3, lr=1e-3)
    trainer = WavefrontPathIntegrator(sc, cam, spp=16, max_depth=8, seed=60, max_paths=1 << 16, guide=g,
                                      guide_frac=0.5, train=16, train_batch=16384, product=True)
    for k in range(48):
        trainer.render_sum(sample_offset=16 * k)
    steps, _ = g.stats()
    assert steps > 100
    spp = 2048
    plain = WavefrontPathIntegrator(sc, cam, spp=spp, max_depth=8, seed=21, max_paths=1 << 20).render()
    guided = WavefrontPathIntegrator(sc, cam, spp=spp, max_depth=8, seed=22, max_paths=1 << 20, guide=g,
                                     guide_frac=0.5, product=True).render()
    plain, guided = plain.cpu().numpy(), guided.cpu().numpy()
    _say("adam steps", steps, "mean plain", plain.mean(), "guided", guided.mean(), "relMSE between",
         S.rel_mse(guided, plain))
    assert np.isfinite(guided).all()
    assert abs(guided.mean() - plain.mean()) < 0.01 * plain.mean()


def _three_jobs():
    from rs_illume_amd.render import Scene
    descs = [S.cornell_box(0), S.cornell_box(1, glossy=True, light_size=0.2, light_L=26.5625), S.cornell_box(2)]
    scenes = [Scene(d) for d in descs]
    cams = [S.cornell_camera(40, 40), S.cornell_camera(24, 56), S.cornell_camera(64, 32)]
    return scenes, [(scenes[0], cams[0], 6, 8, 1), (scenes[1], cams[1], 3, 8, 2), (scenes[2], cams[2], 9, 8, 3)]


def test_per_job_networks_route_each_job_to_its_network(gpu):
    """Frozen networks A and B on jobs 0 and 1, none on job 2, in one batch:
    each job's image equals its render alone with its own network (or plain);
    every job naming one network equals the shared-network call."""
    from rs_illume_amd.guide import Guide
    from rs_illume_amd.render import GuideCfg, Wavefront
    _say("case per-job networks")
    scenes, jobs = _three_jobs()
    a, b = Guide(seed=4), Guide(seed=5)
    frozen = GuideCfg(0.5, 0, 0, 0.0, 1)
    wf = Wavefront(max_paths=4096, max_depth=8, block=128, graph_iters=4)
    film, offs = wf.render_batch(jobs, gcfg=frozen, guides=[a, b, None])
    film = film.cpu().numpy()
    for j, g in enumerate((a, b, None)):
        one, _ = wf.render_batch([jobs[j]], guide=g, gcfg=frozen)
        one = one.cpu().numpy()
        got = film[offs[j]:offs[j] + one.size]
        assert np.allclose(got, one, rtol=1e-5, atol=1e-5), (j, np.abs(got - one).max())
    shared, _ = wf.render_batch(jobs, guide=a, gcfg=frozen)
    same, _ = wf.render_batch(jobs, gcfg=frozen, guides=[a, a, a])
    assert np.allclose(shared.cpu().numpy(), same.cpu().numpy(), rtol=1e-5, atol=1e-5)
    with pytest.raises(Exception):
        many = [Guide(seed=100 + k) for k in range(17)]
        wf.render_batch([jobs[0]] * 17, gcfg=frozen, guides=many)
    wf.close()


def test_per_job_networks_train_on_their_own_samples(gpu):
    """Training with a network per job: every network takes Adam steps from
    its job's samples; a job without a network trains nothing; the trained
    networks differ from each other and from their initial weights."""
    from rs_illume_amd.guide import Guide
    from rs_illume_amd.render import GuideCfg, Wavefront
    _say("case per-job training")
    scenes, jobs = _three_jobs()
    a, b = Guide(seed=6), Guide(seed=6)  # equal initial weights
    p0 = a.params()
    gc = GuideCfg(0.5, 8, 4096, 0.0, 1)
    wf = Wavefront(max_paths=1 << 14, max_depth=8, block=128, graph_iters=4)
    for k in range(6):
        wf.render_batch([(sc, cam, spp, d, seed, spp * k) for sc, cam, spp, d, seed in jobs], gcfg=gc,
                        guides=[a, b, None])
    sa, sb = a.stats()[0], b.stats()[0]
    _say("adam steps", sa, sb)
    assert sa > 0 and sb > 0
    pa, pb = a.params(), b.params()
    assert np.isfinite(pa).all() and np.isfinite(pb).all()
    assert not np.array_equal(pa, p0) and not np.array_equal(pa, pb)
    wf.close()
