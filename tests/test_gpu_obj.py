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


def test_textured_obj_renders_like_the_spec(tmp_path):
    """The textured / alpha cut-out scene written as OBJ + MTL + PNG and loaded through
    crt_obj_load: the decoded textures (flipped back, byte / 255), the per-corner texture
    coordinates (negative and > 1 ones included) and the alpha cut-outs render the same image,
    bit for bit, as the in-memory scene with the same 8-bit textures.  The OBJ loader keeps
    only Kd and map_Kd, so the glass / metal types and the clear pane's alpha are restored
    from the spec, as the UI's material editor would."""
    from _objfiles import quantise_textures, write_textured_obj

    spec, imgs = quantise_textures(scenes.cutout())
    path = tmp_path / "cutout.obj"
    write_textured_obj(spec, imgs, path)
    ctx = Context(0)
    ref_sc = scene_from_spec(ctx, spec)
    obj = ObjModel(path)
    assert [m["tex"] for m in obj.materials] == [m.tex for m in spec.materials]
    for got, want in zip(obj.textures, spec.textures):
        np.testing.assert_array_equal(got, want)
    sc = Scene(ctx)
    mid = obj.add_to(sc)
    sc.set_materials(mid, spec.materials)
    sc.set_skybox(spec.sky)
    sc.set_skybox_rotation(*map(float, spec.sky_rotation))
    sc.set_sun_enabled(spec.sun_enabled)
    sc.set_sun(**spec.sun)
    sc.set_camera(spec.camera)
    sc.commit()
    print("textured OBJ render", spec.res, flush=True)
    imgs_out = []
    for s in (ref_sc, sc):
        r = Renderer(ctx, s, spec.res[0], spec.res[1], spec.max_bounces)
        r.set_resolution(*spec.res)
        r.set_options(mode=spec.mode, rr_start=spec.rr_start, seed=4)
        r.render(4)
        imgs_out.append((r.raw(), r.current_albedos()))
        r.close()
    assert np.array_equal(imgs_out[0][0], imgs_out[1][0])
    assert np.array_equal(imgs_out[0][1], imgs_out[1][1])
    alb = imgs_out[0][1].reshape(-1, 4)[:, :3]
    assert len(np.unique(np.round(alb, 4), axis=0)) > 8     # the checker and fence colours show
    obj.close()
    sc.close()
    ref_sc.close()
    ctx.close()
