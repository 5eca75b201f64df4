"""Image export (crt_image_write): host code, runs without a GPU.

Each writer is read back by an independent decoder written from the format's public
specification (PNG: zlib + filter 0 rows; OpenEXR: scanline, uncompressed, half; Radiance
RGBE; PFM) and compared with what asset_loader.cpp:89-178 computes from the same buffer.
"""
import struct
import zlib

import numpy as np
import pytest

from crender_amd import CrtError, write_image


@pytest.fixture
def img():
    rng = np.random.default_rng(3)
    a = rng.random((5, 7, 4), dtype=np.float32) * 1.3
    a[0, 0] = [0.0, 1.0, 2.0, 0.5]
    a[1, 2] = [65504.0, 1e-6, 3.0e-5, 1.0]   # half max, subnormal range
    return a


def _png_rows(data):
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w = 8, b"", 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        t = data[pos + 4:pos + 8]
        d = data[pos + 8:pos + 8 + n]
        assert struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0] == zlib.crc32(t + d) & 0xFFFFFFFF
        if t == b"IHDR":
            w, h, depth, ctype = struct.unpack(">IIBB", d[:10])
            assert depth == 8 and ctype == 6
        elif t == b"IDAT":
            idat += d
        pos += 12 + n
    raw = zlib.decompress(idat)
    rows = np.frombuffer(raw, np.uint8).reshape(h, 1 + 4 * w)
    assert np.all(rows[:, 0] == 0)
    return rows[:, 1:].reshape(h, w, 4)


def test_png(tmp_path, img):
    p = tmp_path / "a.png"
    write_image(p, img)
    got = _png_rows(p.read_bytes())
    want = np.floor(np.minimum(img * np.float32(255.0), 255.0)).astype(np.uint8)
    np.testing.assert_array_equal(got, want)


def test_exr_half_bgr(tmp_path, img):
    p = tmp_path / "a.exr"
    write_image(p, img)
    d = p.read_bytes()
    assert struct.unpack("<ii", d[:8]) == (20000630, 2)
    pos, attrs = 8, {}
    while d[pos] != 0:
        e = d.index(b"\0", pos)
        name = d[pos:e].decode()
        e2 = d.index(b"\0", e + 1)
        typ = d[e + 1:e2].decode()
        size = struct.unpack("<i", d[e2 + 1:e2 + 5])[0]
        attrs[name] = (typ, d[e2 + 5:e2 + 5 + size])
        pos = e2 + 5 + size
    pos += 1
    assert attrs["compression"][1] == b"\0"
    assert attrs["channels"][1].startswith(b"B\0")
    H, W = img.shape[:2]
    assert struct.unpack("<iiii", attrs["dataWindow"][1]) == (0, 0, W - 1, H - 1)
    offs = struct.unpack(f"<{H}Q", d[pos:pos + 8 * H])
    for y in range(H):
        o = offs[y]
        yy, n = struct.unpack("<ii", d[o:o + 8])
        assert yy == y and n == W * 6
        planes = np.frombuffer(d[o + 8:o + 8 + n], np.float16).reshape(3, W)
        for k, c in enumerate((2, 1, 0)):
            np.testing.assert_array_equal(planes[k], img[y, :, c].astype(np.float16))


def test_hdr_rgbe(tmp_path, img):
    p = tmp_path / "a.hdr"
    write_image(p, img)
    d = p.read_bytes()
    head, _, body = d.partition(b"\n\n")
    assert head.startswith(b"#?RADIANCE")
    res, _, px = body.partition(b"\n")
    H, W = img.shape[:2]
    assert res == f"-Y {H} +X {W}".encode()
    q = np.frombuffer(px, np.uint8).reshape(H, W, 4).astype(np.float64)
    dec = np.where(q[..., 3:] > 0, (q[..., :3] + 0.0) * np.ldexp(1.0, (q[..., 3:] - 136).astype(int)), 0.0)
    want = np.power(img[..., :3].astype(np.float64), 2.2)
    m = want.max(axis=2, keepdims=True)
    assert np.all(np.abs(dec - want) <= m / 128 + 1e-30)


def test_pfm_exact(tmp_path, img):
    p = tmp_path / "a.pfm"
    write_image(p, img)
    d = p.read_bytes()
    lines = d.split(b"\n", 3)
    assert lines[0] == b"PF" and lines[2] == b"-1.0"
    W, H = map(int, lines[1].split())
    a = np.frombuffer(lines[3], "<f4").reshape(H, W, 3)[::-1]
    np.testing.assert_array_equal(a, img[..., :3])


def test_jpg_unsupported_and_bad_args(tmp_path, img):
    with pytest.raises(CrtError):
        write_image(tmp_path / "a.jpg", img)
    with pytest.raises(CrtError):
        write_image(tmp_path / "nodir" / "a.png", img)
