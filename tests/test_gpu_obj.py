"""The seeded Cornell box written as OBJ + MTL, loaded through crt_obj_load (the reference's
asset_loader::load_model path) and lit by restoring the emitter's emission (the OBJ loader
keeps emission 0, asset_loader.cpp:233, as the UI's material editor would), renders the
same image, bit for bit, as the in-memory scene: same triangles, same primID order."""
import numpy as np
import pytest

from crender_amd import Context, ObjModel, Renderer, Scene, scene_from_spec, scenes

pytestmark = pytest.mark.gpu


def test_cornell_obj_renders_like_the_spec(tmp_path):
    spec = scenes.cornell(1)
    path = tmp_path / "cornell.obj"
    scenes.write_obj(spec, path)
    ctx = Context(0)
    ref_sc = scene_from_spec(ctx, spec)
    obj = ObjModel(path)
    sc = Scene(ctx)
    mid = obj.add_to(sc)
    sc.set_materials(mid, spec.materials)        # emission back on the light
    sc.set_skybox(spec.sky)
    sc.set_sun_enabled(spec.sun_enabled)
    sc.set_camera(spec.camera)
    sc.commit()
    imgs = []
    for s in (ref_sc, sc):
        r = Renderer(ctx, s, 64, 64, spec.max_bounces)
        r.set_resolution(64, 64)
        r.set_options(mode=0, rr_start=spec.rr_start, seed=3)
        r.render(4)
        imgs.append(r.raw())
        r.close()
    assert np.array_equal(imgs[0], imgs[1])
    assert imgs[0].mean() > 0
    obj.close()
    sc.close()
    ref_sc.close()
    ctx.close()
