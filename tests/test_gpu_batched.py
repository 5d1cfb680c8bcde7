"""The batched workload of many procedural scenes (displaced icospheres on a
ground quad, Lambertian and dielectric objects, 1-4 quad lights, uneven image
sizes and spp) on the GPU:

* a few scenes of the batch, at reduced resolution and 4 spp, against the
  oracle at equal sampler streams (relMSE <= 1e-4, ray counts within 0.2%),
  through both batched renderers (the megakernel batch and the wavefront pool);
* all 64 scenes (seeds 1-64) at 1 spp in one call: each job's image equals
  its own render, the batch's ray count is the sum of the jobs' counts, and the
  wavefront pool gives the megakernel batch's images.
"""
import sys

import numpy as np
import pytest

from rs_illume_amd import scenes as S

pytestmark = pytest.mark.gpu
REL_MSE_BAR = 1e-4
SEEDS = (1, 2, 17)


def _say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _small(cam, side=96):
    c = cam.copy()
    w, h = int(c["width"][0]), int(c["height"][0])
    c["width"] = min(w, side)
    c["height"] = min(h, side)
    return c


def _size(cam):
    return int(cam["width"][0]), int(cam["height"][0])


@pytest.mark.parametrize("renderer", ["megakernel_batch", "wavefront"])
def test_procedural_scenes_match_oracle(gpu, oracle, renderer):
    import torch
    from rs_illume_amd.render import Scene, Wavefront, render_batch
    spp, depth = 4, 8
    descs = [S.procedural_scene(s) for s in SEEDS]
    scenes = [Scene(d) for d, _, _ in descs]
    cams = [_small(c) for _, c, _ in descs]
    jobs = [(sc, cam, spp, depth, seed) for sc, cam, seed in zip(scenes, cams, SEEDS)]
    _say("case procedural vs oracle", renderer, [(s, len(d.idx), _size(c)) for s, (d, _, _), c in
                                                 zip(SEEDS, descs, cams)])
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    if renderer == "wavefront":
        wf = Wavefront(max_paths=1 << 16, max_depth=depth, block=64, graph_iters=4)
        film, offs = wf.render_batch(jobs, ray_counts=counts)
    else:
        film, offs = render_batch(jobs, ray_counts=counts)
    film = film.cpu().numpy()
    total = np.zeros(2, np.int64)
    for (d, _, _), cam, seed, off in zip(descs, cams, SEEDS, offs):
        w, h = _size(cam)
        g = film[off:off + w * h * 3].reshape(h, w, 3)
        o, oc = oracle.OracleRenderScene(d.arrays()).render_sum(cam, spp, depth, seed=seed)
        rm = S.rel_mse(g / spp, o / spp)
        _say("seed", seed, "relMSE", rm, "oracle rays", oc)
        assert np.isfinite(g).all()
        assert rm <= REL_MSE_BAR, (seed, rm)
        total += np.asarray(oc, np.int64)
    gc = counts.cpu().numpy()
    _say("batch rays", gc, "oracle rays", total)
    assert abs(int(gc[0]) - int(total[0])) <= 0.002 * total[0] + 10
    assert abs(int(gc[1]) - int(total[1])) <= 0.002 * total[1] + 10
    for sc in scenes:
        sc.close()


def test_64_scene_batch_properties(gpu):
    import torch
    from rs_illume_amd.render import PathIntegrator, Scene, Wavefront, render_batch
    descs = [S.procedural_scene(s) for s in range(1, 65)]
    # the device builder keeps set-up to seconds; hits do not depend on the tree
    scenes = [Scene(d, builder="LBVH") for d, _, _ in descs]
    jobs = [(sc, cam, 1, 8, seed) for seed, (sc, (_, cam, _)) in enumerate(zip(scenes, descs), 1)]
    n_px = sum(_size(c)[0] * _size(c)[1] for _, c, _ in descs)
    _say("case 64-scene batch, 1 spp,", n_px, "pixels,", sum(len(d.idx) for d, _, _ in descs), "triangles")
    counts = torch.zeros(2, dtype=torch.int64, device="cuda")
    film, offs = render_batch(jobs, ray_counts=counts)
    torch.cuda.synchronize()
    film_np = film.cpu().numpy()
    assert np.isfinite(film_np).all()
    # every job's own render: same paths, so the same sums (float atomics sum a
    # pixel's samples in any order) and the same ray counts
    total = torch.zeros(2, dtype=torch.int64, device="cuda")
    for j, (sc, cam, spp, depth, seed) in enumerate(jobs):
        c = torch.zeros(2, dtype=torch.int64, device="cuda")
        one = PathIntegrator(cam, spp=spp, max_depth=depth, seed=seed).render_sum(sc, ray_counts=c)
        total += c
        if j in (0, 3, 16, 40, 63):
            w, h = _size(cam)
            got = film_np[offs[j]:offs[j] + w * h * 3].reshape(h, w, 3)
            ref = one.cpu().numpy()
            assert np.allclose(got, ref, rtol=1e-5, atol=1e-5), (j, np.abs(got - ref).max())
    torch.cuda.synchronize()
    _say("batch rays", counts.cpu().numpy(), "sum of the jobs' rays", total.cpu().numpy())
    assert torch.equal(counts, total)
    # the wavefront pool over the same batch
    wf = Wavefront(max_paths=1 << 21, max_depth=8, block=64, graph_iters=4)
    wcounts = torch.zeros(2, dtype=torch.int64, device="cuda")
    wfilm, woffs = wf.render_batch(jobs, ray_counts=wcounts)
    torch.cuda.synchronize()
    assert woffs == offs
    wfilm_np = wfilm.cpu().numpy()
    assert np.allclose(wfilm_np, film_np, rtol=1e-4, atol=1e-4), np.abs(wfilm_np - film_np).max()
    wc, mc = wcounts.cpu().numpy(), counts.cpu().numpy()
    _say("wavefront rays", wc)
    assert np.all(np.abs(wc - mc) <= 0.002 * mc + 10)
    wf.close()
    for sc in scenes:
        sc.close()
