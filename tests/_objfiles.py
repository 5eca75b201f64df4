"""Writes a SceneSpec with textures as Wavefront OBJ + MTL + PNG files, for tests that load
scenes through crt_obj_load (asset_loader::load_model, asset_loader.cpp:182-303)."""
import struct
import zlib
from pathlib import Path

import numpy as np


def png_bytes(rgba: np.ndarray) -> bytes:
    """An 8-bit RGBA PNG of an (h, w, 4) uint8 array, rows top to bottom."""
    h, w = rgba.shape[:2]
    raw = b"".join(b"\x00" + bytes(rgba[y].reshape(-1)) for y in range(h))

    def chunk(t, d):
        return struct.pack(">I", len(d)) + t + d + struct.pack(">I", zlib.crc32(t + d) & 0xFFFFFFFF)

    return (b"\x89PNG\r\n\x1a\n" + chunk(b"IHDR", struct.pack(">IIBBBBB", w, h, 8, 6, 0, 0, 0)) +
            chunk(b"IDAT", zlib.compress(raw)) + chunk(b"IEND", b""))


def quantise_textures(spec):
    """The spec with every texture rounded to 8 bits, as decoding a PNG gives it back
    (byte / 255 in FP32); returns (spec, the uint8 images)."""
    imgs = [np.clip(np.rint(np.asarray(t, np.float64) * 255.0), 0, 255).astype(np.uint8) for t in spec.textures]
    spec.textures = [im.astype(np.float32) / np.float32(255.0) for im in imgs]
    return spec, imgs


def write_textured_obj(spec, imgs, path) -> None:
    """OBJ with per-corner `vt` and `f v/vt`, MTL with Kd and map_Kd (one PNG per texture).
    The loader flips images vertically (asset_loader.cpp), so each PNG is written flipped."""
    path = Path(path)
    for i, im in enumerate(imgs):
        (path.parent / f"tex{i}.png").write_bytes(png_bytes(np.ascontiguousarray(im[::-1])))
    mtl = path.with_suffix(".mtl")
    with open(mtl, "w") as f:
        for i, m in enumerate(spec.materials):
            c = m.colour
            f.write(f"newmtl m{i}\nKd {c[0]:.9g} {c[1]:.9g} {c[2]:.9g}\n")
            if m.tex >= 0:
                f.write(f"map_Kd tex{m.tex}.png\n")
            f.write("\n")
    tris = np.asarray(spec.vertex_indices, dtype=np.int64).reshape(-1, 3)
    tidx = np.asarray(spec.texture_indices, dtype=np.int64).reshape(-1, 3)
    mi = np.asarray(spec.material_indices, dtype=np.int64)
    with open(path, "w") as f:
        f.write(f"# {spec.name}\nmtllib {mtl.name}\n")
        for v in np.asarray(spec.vertices, dtype=np.float32):
            f.write(f"v {v[0]:.9g} {v[1]:.9g} {v[2]:.9g}\n")
        for t in np.asarray(spec.texture_coords, dtype=np.float32):
            f.write(f"vt {t[0]:.9g} {t[1]:.9g}\n")
        cur = None
        for t, u, m in zip(tris, tidx, mi):
            if m != cur:
                f.write(f"usemtl m{m}\n")
                cur = m
            f.write(f"f {t[0] + 1}/{u[0] + 1} {t[1] + 1}/{u[1] + 1} {t[2] + 1}/{u[2] + 1}\n")
