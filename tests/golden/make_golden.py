"""Writes the golden fixtures under tests/golden/ with the CPU oracle.

The reference ships no golden vectors, test scenes or images, so these fixtures are
the oracle's own outputs on seeded inputs, committed so that later changes of either
implementation are compared against fixed data:
  * cornell_rays.npz   4096 jittered camera rays of the Cornell box and their FP64
                       brute-force closest hits (prim, t, u, v);
  * soup_rays.npz      2048 random rays against a 20k-triangle soup, FP64 brute force;
  * cornell_compat.npz 64x64, 4 spp compat render (FP64 oracle, Philox seed 5) with AOVs.
Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "oracle"))

import pyoracle as O  # noqa: E402
from crender_amd import scenes  # noqa: E402


def cornell_rays():
    spec = scenes.cornell(1)
    o, d = scenes.rays_from_camera(spec, (64, 64), seed=31)
    p, t, u, v = O.intersect(spec.corners(), o, d, O.BRUTE64)
    return {"org": o, "dir": d, "prim": p, "t": t, "u": u, "v": v}


def soup_rays():
    spec = scenes.soup(20_000, seed=13)
    o, d = scenes.diffuse_rays(2048, seed=14)
    p, t, u, v = O.intersect(spec.corners(), o, d, O.BRUTE64)
    return {"org": o, "dir": d, "prim": p, "t": t, "u": u, "v": v}


def cornell_compat():
    spec = scenes.cornell(1)
    r = O.render(spec, res=(64, 64), spp=4, seed=5, precision=64, trav=O.BRUTE64, nthreads=1)
    return {"raw": r["raw"], "albedo": r["albedo"], "normal": r["normal"], "depth": r["depth"],
            "ext_rays": np.array(r["stats"]["ext_rays"]), "ref_segments": np.array(r["stats"]["ref_segments"])}


FIXTURES = {"cornell_rays": cornell_rays, "soup_rays": soup_rays, "cornell_compat": cornell_compat}


if __name__ == "__main__":
    for name, fn in FIXTURES.items():
        np.savez_compressed(HERE / f"{name}.npz", **fn())
        print("wrote", HERE / f"{name}.npz")
