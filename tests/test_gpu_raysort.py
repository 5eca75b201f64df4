"""Ray-queue reordering (CRT_FLAG_SORT_RAYS) changes only the order in which secondary rays
are traced: every path's result, and so the image and the ray counts, must be bit-identical
to the unsorted renderer's."""
import numpy as np
import pytest

from crender_amd import Context, Renderer, scene_from_spec, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["soup", "sunlit"])
def test_sorted_equals_unsorted(which):
    spec = scenes.soup(50_000, seed=3) if which == "soup" else scenes.sunlit(4, mode=1)
    res = (160, 90)
    outs = []
    for sort in (False, True):
        ctx = Context(0, sort_rays=sort)
        sc = scene_from_spec(ctx, spec)
        r = Renderer(ctx, sc, res[0], res[1], 6)
        r.set_resolution(*res)
        r.set_options(mode=spec.mode, rr_start=2, seed=4)
        r.render(3)
        st = r.current_stats()
        outs.append((r.raw(), st["extension_rays"], st["shadow_rays"], st["total_rays"]))
        r.close()
        sc.close()
        ctx.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert outs[0][1:] == outs[1][1:]
