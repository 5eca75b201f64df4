"""Material-sorted shading (CRT_FLAG_SORT_MATERIALS) only changes the order in which a
bounce's hits are shaded: every path's result, and so the image and the ray counts, must be
bit-identical to the renderer that shades in queue order."""
import numpy as np
import pytest

from crender_amd import Context, Renderer, scene_from_spec, scenes

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("which", ["terrain", "sunlit", "cornell"])
def test_material_sorted_equals_unsorted(which):
    spec = {"terrain": lambda: scenes.terrain(grid=64, spheres=120, seed=5, res=(96, 54)),
            "sunlit": lambda: scenes.sunlit(4, mode=1),
            "cornell": lambda: scenes.cornell(1)}[which]()
    res = (96, 54)
    outs = []
    for sort in (False, True):
        ctx = Context(0, sort_materials=sort)
        sc = scene_from_spec(ctx, spec)
        r = Renderer(ctx, sc, res[0], res[1], 6)
        r.set_resolution(*res)
        r.set_options(mode=spec.mode, rr_start=2, seed=9)
        r.render(3)
        st = r.current_stats()
        outs.append((r.raw(), r.current_albedos(), st["extension_rays"], st["shadow_rays"], st["total_rays"]))
        r.close()
        sc.close()
        ctx.close()
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert outs[0][2:] == outs[1][2:]
