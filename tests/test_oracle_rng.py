"""Pins the oracle's Philox4x32-10 (and the scene generator's numpy copy).

* Published known-answer vectors of Philox4x32-10 (Random123 kat_vectors).
* Cross-run against rocRAND's Philox (rocrand_philox4x32_10.h, compiled host-side).
* u32 -> float mapping: (x >> 8) * 2^-24 in [0, 1).
"""
import json
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

from crender_amd import philox as P

GOLDEN = Path(__file__).resolve().parent / "golden"
KATS = json.loads((GOLDEN / "philox_kat.json").read_text())["vectors"]


def _h(x):
    return [int(v, 16) for v in x]


@pytest.mark.parametrize("kat", KATS, ids=lambda k: k["name"])
def test_oracle_philox_kat(oracle, kat):
    out = oracle.philox([_h(kat["ctr"])], _h(kat["key"]))[0]
    assert list(out) == _h(kat["out"])


@pytest.mark.parametrize("kat", KATS, ids=lambda k: k["name"])
def test_numpy_philox_kat(kat):
    c = _h(kat["ctr"])
    k = _h(kat["key"])
    out = P.philox4x32_10(c[0], c[1], c[2], c[3], k[0], k[1])
    assert [int(x) for x in out] == _h(kat["out"])


def test_kat_catches_a_wrong_round_count(oracle):
    """The KAT distinguishes 10 rounds from 9: a plausible slip is caught."""
    import numpy as np

    kat = KATS[2]
    c, k = _h(kat["ctr"]), _h(kat["key"])

    def rounds(n):
        c0, c1, c2, c3 = (np.uint32(x) for x in c)
        k0, k1 = k
        for r in range(n):
            if r:
                k0 = (k0 + 0x9E3779B9) & 0xFFFFFFFF
                k1 = (k1 + 0xBB67AE85) & 0xFFFFFFFF
            p0 = 0xD2511F53 * int(c0)
            p1 = 0xCD9E8D57 * int(c2)
            c0, c1, c2, c3 = ((p1 >> 32) ^ int(c1) ^ k0) & 0xFFFFFFFF, p1 & 0xFFFFFFFF, \
                ((p0 >> 32) ^ int(c3) ^ k1) & 0xFFFFFFFF, p0 & 0xFFFFFFFF
        return [int(c0), int(c1), int(c2), int(c3)]

    assert rounds(10) == _h(kat["out"])
    assert rounds(9) != _h(kat["out"])


@pytest.mark.skipif(shutil.which("hipcc") is None and not Path("/opt/rocm/bin/hipcc").exists(),
                    reason="hipcc not available")
def test_oracle_matches_rocrand(oracle, tmp_path):
    exe = tmp_path / "probe"
    src = Path(__file__).resolve().parent / "rocrand_probe.cpp"
    hipcc = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    subprocess.run([hipcc, "-O1", "-std=c++17", "--offload-arch=gfx950", str(src), "-o", str(exe)], check=True,
                   capture_output=True)
    rng = np.random.default_rng(123)
    ctr = rng.integers(0, 2 ** 32, size=(4096, 4), dtype=np.uint64).astype(np.uint32)
    ctr[:16] = 0xFFFFFFFF
    keys = rng.integers(0, 2 ** 32, size=(4096, 2), dtype=np.uint64).astype(np.uint32)
    lines = "\n".join(" ".join(f"{v:x}" for v in list(c) + list(k)) for c, k in zip(ctr, keys))
    proc = subprocess.run([str(exe)], input=lines, capture_output=True, text=True, check=True)
    ref = np.array([[int(x, 16) for x in l.split()] for l in proc.stdout.strip().splitlines()], dtype=np.uint32)
    ours = np.stack([oracle.philox(c[None], k)[0] for c, k in zip(ctr, keys)])
    assert np.array_equal(ours, ref)
    # and the numpy generator stream
    np_out = np.stack(P.philox4x32_10(ctr[:, 0], ctr[:, 1], ctr[:, 2], ctr[:, 3], 0, 0), axis=1)
    zero_key = np.stack([oracle.philox(c[None], [0, 0])[0] for c in ctr[:64]])
    assert np.array_equal(np_out[:64], zero_key)


def test_u01_range(oracle):
    L = oracle.lib()
    assert L.orc_u01(0) == 0.0
    assert L.orc_u01(0xFFFFFFFF) == np.float32(1.0 - 2.0 ** -24)
    assert L.orc_u01(0x80000000) == 0.5
    x = np.array([0, 255, 256, 0xFFFFFFFF, 12345678], np.uint32)
    assert np.array_equal(P.u01(x), np.array([L.orc_u01(int(v)) for v in x], np.float32))
