"""Rendering with the neural radiance cache.

  * the cache trains online (its loss falls) and the image stays finite and close in mean
    to the plain tracer's (the cache is biased, the bound here is loose: 15%);
  * equal-time comparison on the triangle soup (diffuse, albedo 0.8, constant sky): the
    image MSE against a 2048-spp plain reference is lower with the cache than without it
    for the same wall time, warm-up passes of the cache included.  This is the gain the
    learned part brings; the numbers are printed for the record.
"""
import time

import numpy as np
import pytest

from crender_amd import Context, Nrc, Renderer, scene_from_spec, scenes

pytestmark = pytest.mark.gpu

RES = (256, 256)


def _soup():
    spec = scenes.soup(200_000, seed=2, res=RES)
    spec.mode = 1
    spec.max_bounces = 16
    spec.rr_start = 3
    return spec


def _renderer(ctx, sc, spec, seed):
    r = Renderer(ctx, sc, RES[0], RES[1], spec.max_bounces)
    r.set_resolution(*RES)
    r.set_options(mode=spec.mode, rr_start=spec.rr_start, seed=seed)
    return r


@pytest.fixture(scope="module")
def setup():
    spec = _soup()
    ctx = Context(0)
    sc = scene_from_spec(ctx, spec)
    ref = _renderer(ctx, sc, spec, seed=99)
    ref.render(2048)
    img = ref.raw() / 2048.0
    ref.close()
    yield ctx, sc, spec, img
    sc.close()
    ctx.close()


def test_cache_trains_and_renders(setup):
    ctx, sc, spec, ref = setup
    nrc = Nrc(ctx, max_batch=65536, seed=3)
    r = _renderer(ctx, sc, spec, seed=5)
    try:
        r.set_nrc(nrc, train_paths=16384, train_steps=4)
        losses = []
        for _ in range(40):
            r.render(1)
            losses.append(r.current_stats()["nrc_loss"])
        img = r.raw() / 40.0
    finally:
        r.close()
        nrc.close()
    assert np.all(np.isfinite(img))
    assert abs(img.mean() - ref.mean()) < 0.15 * ref.mean(), (img.mean(), ref.mean())
    assert np.mean(losses[-10:]) < np.mean(losses[:3])


def _mse(a, b):
    return float(np.mean((a.astype(np.float64) - b) ** 2))


def test_equal_time_mse_gain(setup):
    ctx, sc, spec, ref = setup
    # plain tracer: passes in a fixed wall-time budget
    budget = 2.0
    plain = _renderer(ctx, sc, spec, seed=7)
    t0 = time.perf_counter()
    n_plain = 0
    while time.perf_counter() - t0 < budget:
        plain.render(1)
        n_plain += 1
    mse_plain = _mse(plain.raw() / n_plain, ref)
    plain.close()
    # cache: warm-up passes (training only, counted in the budget), then render
    nrc = Nrc(ctx, max_batch=65536, seed=4)
    r = _renderer(ctx, sc, spec, seed=8)
    try:
        r.set_nrc(nrc, train_paths=65536, warmup_passes=8)     # library defaults otherwise
        t0 = time.perf_counter()
        r.render(8)
        t_warm = time.perf_counter() - t0
        n_nrc = 0
        while time.perf_counter() - t0 < budget:
            r.render(1)
            n_nrc += 1
        assert r.current_sample_count() == n_nrc
        mse_nrc = _mse(r.raw() / n_nrc, ref)
    finally:
        r.close()
        nrc.close()
    print(f"equal time {budget:.1f}s: plain {n_plain} spp MSE {mse_plain:.3e}; cache warm-up {t_warm:.2f}s + "
          f"{n_nrc} spp MSE {mse_nrc:.3e}; ratio {mse_plain / max(mse_nrc, 1e-30):.2f}")
    # the cache must at least halve the error of the plain tracer at equal time
    assert mse_nrc < 0.5 * mse_plain


@pytest.mark.parametrize("bad", [np.nan, np.inf])
def test_diverged_cache_falls_back_to_the_plain_tracer(setup, bad):
    """Non-finite weights (set through crt_nrc_set_params) make the cache's predictions
    non-finite: the renderer drops the wavefronts that used them, detaches the cache and
    renders those passes again with the plain tracer, so the image equals, bit for bit, what
    a renderer without the cache accumulates over the same passes."""
    ctx, sc, spec = setup[:3]
    print(f"diverged cache ({bad})", flush=True)
    plain = _renderer(ctx, sc, spec, seed=31)
    plain.render(6)
    ref = plain.raw()
    plain.close()
    nrc = Nrc(ctx, seed=3)
    w = nrc.params()
    w[::7] = bad
    nrc.set_params(w)
    r = _renderer(ctx, sc, spec, seed=31)
    r.set_nrc(nrc)
    r.render(6)
    img = r.raw()
    st = r.current_stats()
    assert r.current_sample_count() == 6
    assert np.isfinite(img).all()
    assert np.array_equal(img, ref)
    assert st["nrc_detached"] == 1 and st["nrc_discarded_passes"] >= 6
    # attaching a healthy cache again clears the fallback
    nrc2 = Nrc(ctx, seed=4)
    r.set_nrc(nrc2)
    r.render(2)
    st = r.current_stats()
    assert st["nrc_detached"] == 0 and np.isfinite(r.raw()).all()
    r.close()
    nrc.close()
    nrc2.close()
