"""The C-ABI library loads and exports every function include/crt.h declares (CPU only)."""
import ctypes
from pathlib import Path

import pytest

from crender_amd import _abi


def test_header_functions_are_bound():
    names = _abi.header_functions()
    assert len(names) >= 40
    missing = [n for n in names if n not in _abi.SIGNATURES]
    assert not missing, f"declared in crt.h but not bound in _abi.SIGNATURES: {missing}"


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_abi.LIBPATH))
    missing = [n for n in _abi.header_functions() if not hasattr(lib, n)]
    assert not missing, f"libcrt.so lacks {missing}"


def test_cpu_only_entry_points():
    L = _abi.lib()
    assert L.crt_abi_version() == 1
    cam = _abi.Camera()
    L.crt_camera_default(ctypes.byref(cam))
    assert tuple(cam.position) == (5.0, 5.0, 0.0) and cam.fov == 75.0   # camera.h:21
    assert list(cam.matrix) == [1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]
    sun = _abi.Sun()
    L.crt_sun_default(ctypes.byref(sun))
    assert sun.intensity == 100.0 and abs(sun.size - 3.14159265359 / 48) < 1e-7   # components.h:23-29
    m = _abi.Material()
    L.crt_material_default(ctypes.byref(m))
    assert (m.type, m.ior, m.roughness, m.reflectiveness, m.emission, m.tex) == (1, 1.5, 0.5, 1.0, 0.0, -1)


def test_camera_update_matches_glm_order():
    """camera.cpp:54-68: T(p) * Ry(x) * Rx(y) * Rz(z), checked against numpy."""
    import numpy as np
    from crender_amd import scenes

    L = _abi.lib()
    cam = _abi.Camera()
    L.crt_camera_default(ctypes.byref(cam))
    pos, rot = (1.0, 2.0, -3.0), (30.0, -20.0, 10.0)
    for k in range(3):
        cam.position[k] = pos[k]
        cam.rotation[k] = rot[k]
    L.crt_camera_update(ctypes.byref(cam))
    ref = scenes.camera_matrix(pos, rot)
    np.testing.assert_allclose(np.array(cam.matrix), ref, atol=1e-6)
    # a pure yaw of 90 degrees turns +z (the camera's forward) into +x
    m = scenes.camera_matrix((0, 0, 0), (90, 0, 0)).reshape(4, 4).T
    np.testing.assert_allclose(m[:3, :3] @ [0, 0, 1], [1, 0, 0], atol=1e-6)


def test_errors_without_gpu_are_reported_not_fatal():
    """Invalid arguments return codes with a message; the process keeps running."""
    L = _abi.lib()
    rc = L.crt_context_create(None, None)
    assert rc == -1
    assert b"null" in L.crt_last_error()


@pytest.mark.parametrize("field,value,word", [("bvh_width", 3, b"bvh_width"), ("bvh_prim_cost", -1.0, b"prim_cost"),
                                             ("bvh_prim_cost", float("nan"), b"prim_cost"),
                                             ("trace_refill", 65, b"trace_refill"), ("trace_park_min", 100, b"park_min"),
                                             ("trace_stack", 12, b"trace_stack")])
def test_context_options_are_checked_before_the_device(field, value, word):
    """Out-of-range context options are refused with a message before any device is used."""
    L = _abi.lib()
    o = _abi.Options()
    L.crt_options_default(ctypes.byref(o))
    setattr(o, field, value)
    h = ctypes.c_void_p()
    assert L.crt_context_create(ctypes.byref(o), ctypes.byref(h)) == -1
    assert word in L.crt_last_error() and not h.value


def test_struct_layouts_match_header(tmp_path):
    """Every ctypes mirror of a crt.h struct has the C struct's size and field offsets
    (compiled with the host C compiler against the header)."""
    import ctypes as C
    import shutil
    import subprocess

    from crender_amd import _abi as A

    cc = shutil.which("gcc") or shutil.which("cc")
    if cc is None:
        pytest.skip("no host C compiler")
    pairs = {"crt_material": A.Material, "crt_image": A.Image, "crt_model_data": A.ModelData, "crt_sun": A.Sun,
             "crt_camera": A.Camera, "crt_options": A.Options, "crt_render_options": A.RenderOptions,
             "crt_stats": A.Stats, "crt_bvh_info": A.BvhInfo, "crt_rays": A.Rays, "crt_hits": A.Hits,
             "crt_records": A.Records,
             "crt_nrc_options": A.NrcOptions, "crt_nrc_render_options": A.NrcRenderOptions,
             "crt_post_options": A.PostOptions, "crt_denoise_options": A.DenoiseOptions, "crt_job": A.Job,
             "crt_batch_options": A.BatchOptions, "crt_batch_stats": A.BatchStats, "crt_tune_result": A.TuneResult}
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "crt.h"', "int main(void) {"]
    for cname, py in pairs.items():
        lines.append(f'printf("{cname} @sizeof %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["return 0; }"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    inc = Path(__file__).resolve().parent.parent / "include"
    subprocess.run([cc, "-std=c11", f"-I{inc}", str(src), "-o", str(exe)], check=True, capture_output=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in pairs.items():
        assert got[(cname, "@sizeof")] == C.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)
