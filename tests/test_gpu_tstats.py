"""The instrumented traversal (CRT_FLAG_TRAVERSAL_STATS) returns the same hits as the
production kernel (bit-identical image and ray counts) and plausible visit counts."""
import numpy as np
import pytest

from crender_amd import Context, Renderer, scene_from_spec, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("width", [4, 8])
def test_instrumented_traversal_matches_and_counts(width):
    spec = scenes.sunlit(4, mode=1)
    res = (80, 60)
    outs = []
    for stats in (False, True):
        ctx = Context(0, bvh_width=width, traversal_stats=stats)
        sc = scene_from_spec(ctx, spec)
        r = Renderer(ctx, sc, res[0], res[1], 5)
        r.set_resolution(*res)
        r.set_options(mode=spec.mode, rr_start=2, seed=3)
        r.render(2)
        outs.append((r.raw(), r.current_stats()))
        r.close()
        sc.close()
        ctx.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    plain, inst = outs[0][1], outs[1][1]
    assert plain["node_visits"] == 0 and plain["tri_tests"] == 0
    rays = inst["extension_rays"] + inst["shadow_rays"]
    assert inst["extension_rays"] == plain["extension_rays"] and inst["shadow_rays"] == plain["shadow_rays"]
    # every ray that enters the tree tests the root; leaves hold 1..16 triangles
    assert rays <= inst["node_visits"] < 200 * rays
    assert inst["leaf_visits"] <= inst["tri_tests"] <= 16 * inst["leaf_visits"]
