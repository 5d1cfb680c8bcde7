"""Equal-quality comparison of the guided and the plain GPU path tracer:
wall time until the running image reaches a relMSE target against a
high-spp plain reference.

Both estimators are progressive: passes of ``spp_per_pass`` samples per pixel
are accumulated into one radiance sum (every pass uses its own sample
indices, so the passes are independent samples of the same estimator), and
after each pass the relMSE of the running mean against the reference is
evaluated on the device. Only the render passes are timed (device
synchronised before and after each one); the error evaluation is not. The
guided tracer trains its proposal online inside its passes, and that
training time is part of its time.
"""
import time

import numpy as np
import torch


def rel_mse_dev(img, ref, eps=0.01):
    """relMSE = mean over pixels and channels of (I - R)^2 / (R^2 + eps), on
    the device in float64."""
    d = img.double() - ref.double()
    return float((d * d / (ref.double() ** 2 + eps)).mean())


def progressive(render_pass, ref, target, spp_per_pass, max_spp, eps=0.01):
    """render_pass(film, sample_offset) adds spp_per_pass samples per pixel to
    film. Returns a dict with the time/spp at which relMSE <= target was first
    reached (None if not within max_spp) and the convergence curve."""
    film = torch.zeros_like(ref)
    spp, elapsed = 0, 0.0
    curve = []
    reached = None
    while spp < max_spp:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        render_pass(film, spp)
        torch.cuda.synchronize()
        elapsed += time.perf_counter() - t0
        spp += spp_per_pass
        e = rel_mse_dev(film / spp, ref, eps)
        curve.append((spp, round(elapsed, 5), e))
        if e <= target:
            reached = {"seconds": elapsed, "spp": spp, "rel_mse": e}
            break
    return {"reached": reached, "curve": curve, "spp": spp, "seconds": elapsed}


def guided_vs_plain(scene, camera, ref, target=0.01, spp_per_pass=8, max_spp=4096, max_depth=8, guide=None,
                    guide_frac=0.5, train_batch=65536, plain="megakernel", seed=1, max_paths=1 << 21, log=None,
                    train_steps=16, warmup_steps=0, product=False, train_budget=None):
    """Time-to-relMSE of the plain and the guided tracer on one scene.
    ``plain`` selects the plain renderer ("megakernel" or "wavefront"). The
    guided tracer samples the BSDF alone (training on those paths) until the
    proposal has taken ``warmup_steps`` Adam steps, then uses one-sample MIS
    with probability ``guide_frac`` (``product``: the proposal times a vMF
    fit of each vertex's BSDF lobe); every pass is an unbiased estimate.
    With ``train_budget`` the proposal trains online until it has taken that
    many Adam steps, then stays frozen (its passes then keep no training
    vertices)."""
    from .guide import Guide
    from .render import GuideCfg, PathIntegrator, Wavefront, WavefrontPathIntegrator
    w, h = int(np.asarray(camera)["width"].reshape(-1)[0]), int(np.asarray(camera)["height"].reshape(-1)[0])
    if plain == "megakernel":
        pi = PathIntegrator(camera, spp=spp_per_pass, max_depth=max_depth, seed=seed)
        plain_pass = lambda film, off: pi.render_sum(scene, rgb_sum=film, spp=spp_per_pass, sample_offset=off)
    else:
        wpi = WavefrontPathIntegrator(scene, camera, spp=spp_per_pass, max_depth=max_depth, seed=seed,
                                      max_paths=max_paths)
        plain_pass = lambda film, off: wpi.render_sum(rgb_sum=film, spp=spp_per_pass, sample_offset=off)
    plain_pass(torch.zeros((h, w, 3), dtype=torch.float32, device="cuda"), 0)  # warm-up
    rp = progressive(plain_pass, ref, target, spp_per_pass, max_spp)
    if log:
        log("plain", rp["reached"], rp["curve"][-1])
    g = guide if guide is not None else Guide(seed=seed)
    wf = Wavefront(max_paths=max_paths, max_depth=max_depth)
    gpi = WavefrontPathIntegrator(scene, camera, spp=spp_per_pass, max_depth=max_depth, seed=seed + 1000, guide=g,
                                  guide_frac=guide_frac, train=train_steps, train_batch=train_batch, wavefront=wf,
                                  product=product)
    # warm-up on a throw-away guide so graph capture is not timed
    wg = WavefrontPathIntegrator(scene, camera, spp=1, max_depth=max_depth, seed=seed + 2000, guide=Guide(seed=seed + 1),
                                 guide_frac=guide_frac, train=True, train_batch=train_batch, wavefront=wf,
                                 product=product)
    wg.render_sum(rgb_sum=torch.zeros((h, w, 3), dtype=torch.float32, device="cuda"))
    state = {"warm": warmup_steps > 0}

    def guided_pass(film, off):
        if state["warm"]:
            gpi.gcfg.guide_frac = 0.0
            gpi.render_sum(rgb_sum=film, spp=spp_per_pass, sample_offset=off)
            if g.stats()[0] >= warmup_steps:
                state["warm"] = False
                gpi.gcfg.guide_frac = guide_frac
        else:
            gpi.render_sum(rgb_sum=film, spp=spp_per_pass, sample_offset=off)
        if train_budget is not None and gpi.gcfg.train and g.stats()[0] >= train_budget:
            gpi.gcfg.train = 0

    rg = progressive(guided_pass, ref, target, spp_per_pass, max_spp)
    if log:
        log("guided", rg["reached"], rg["curve"][-1])
    out = {"plain": rp, "guided": rg, "target": target, "spp_per_pass": spp_per_pass, "plain_renderer": plain,
           "train_steps": train_steps, "warmup_steps": warmup_steps, "guide_frac": guide_frac, "product": product,
           "train_budget": train_budget, "adam_steps": g.stats()[0]}
    if rp["reached"] and rg["reached"]:
        out["time_ratio"] = rg["reached"]["seconds"] / rp["reached"]["seconds"]
    return out


def _timed_pass(fn, reps=5):
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def frozen_ceiling(scene, camera, ref, spp=16, seeds=(11, 12, 13, 14), train_passes=48, train_spp=4,
                   guide_frac=0.25, product=True, max_depth=8, max_paths=1 << 21):
    """What the proposal can buy at best: trained for train_passes guided
    passes (untimed), then frozen; relMSE at equal spp of the guided and the
    plain (megakernel) tracer, averaged over seeds, and the cost of a pass of
    each. ratio_equal_spp x cost_ratio is the equal-quality time ratio the
    guided tracer would reach if training cost nothing."""
    from .guide import Guide
    from .render import PathIntegrator, Wavefront, WavefrontPathIntegrator
    g = Guide(seed=1, lr=1e-3)
    wf = Wavefront(max_paths=max_paths, max_depth=max_depth)
    tr = WavefrontPathIntegrator(scene, camera, spp=train_spp, max_depth=max_depth, seed=50, guide=g,
                                 guide_frac=guide_frac, train=16, wavefront=wf, product=product)
    for k in range(train_passes):
        tr.render_sum(sample_offset=train_spp * k)
    steps, loss = g.stats()
    ep, eg = [], []
    for s in seeds:
        p = PathIntegrator(camera, spp=spp, max_depth=max_depth, seed=s)
        ep.append(rel_mse_dev(p.render_sum(scene) / spp, ref))
        gp = WavefrontPathIntegrator(scene, camera, spp=spp, max_depth=max_depth, seed=s + 100, guide=g,
                                     guide_frac=guide_frac, train=0, wavefront=wf, product=product)
        eg.append(rel_mse_dev(gp.render_sum() / spp, ref))
    film = torch.zeros_like(ref)
    p4 = PathIntegrator(camera, spp=4, max_depth=max_depth, seed=7)
    g4 = WavefrontPathIntegrator(scene, camera, spp=4, max_depth=max_depth, seed=8, guide=g, guide_frac=guide_frac,
                                 train=0, wavefront=wf, product=product)
    tp = _timed_pass(lambda: p4.render_sum(scene, rgb_sum=film))
    tg = _timed_pass(lambda: g4.render_sum(rgb_sum=film))
    wf.close()
    ratio = float(np.mean(eg)) / float(np.mean(ep))
    return {"adam_steps": steps, "last_loss": round(loss, 4), "spp": spp, "seeds": len(seeds),
            "rel_mse_plain": round(float(np.mean(ep)), 6), "rel_mse_guided": round(float(np.mean(eg)), 6),
            "ratio_equal_spp": round(ratio, 4), "plain_4spp_pass_ms": round(tp * 1e3, 3),
            "guided_4spp_pass_ms": round(tg * 1e3, 3), "cost_ratio": round(tg / tp, 3),
            "time_ratio_bound_free_training": round(ratio * tg / tp, 3), "guide_frac": guide_frac,
            "product": product}


def _train_passes(res=512, passes=3, spp=16, train=4, batch=65536):
    """Guided training passes on the glossy small-light Cornell box (the
    workload profiled for the proposal kernels)."""
    from . import scenes as S
    from .guide import Guide
    from .render import Scene, Wavefront, WavefrontPathIntegrator
    sc = Scene(S.cornell_box(0, glossy=True, light_size=0.2, light_L=26.5625))
    cam = S.cornell_camera(res, res)
    g = Guide(seed=1)
    wf = Wavefront(max_paths=1 << 21, max_depth=8)
    gpi = WavefrontPathIntegrator(sc, cam, spp=spp, max_depth=8, seed=0, guide=g, guide_frac=0.5, train=train,
                                  train_batch=batch, wavefront=wf)
    for k in range(passes):
        gpi.render_sum(sample_offset=spp * k)
    torch.cuda.synchronize()
    print("adam steps, last loss:", g.stats(), flush=True)


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="guided training passes (profiling workload)")
    ap.add_argument("--res", type=int, default=512)
    ap.add_argument("--passes", type=int, default=3)
    ap.add_argument("--spp", type=int, default=16)
    ap.add_argument("--train", type=int, default=4)
    ap.add_argument("--batch", type=int, default=65536)
    a = ap.parse_args()
    _train_passes(a.res, a.passes, a.spp, a.train, a.batch)
