"""Image parity between the HIP renderer and the CPU oracle (imported by the GPU tests).

Both sides run the same estimator on identical Philox streams with FP contraction off and
the same FP32 rounding of every operation (transcendental functions are the FP64 ones
rounded once), so a pixel agrees bit for bit unless one of its paths took another branch.
A pixel outside the tolerance must be *classified*: it counts as a near-tie when the oracle
itself changes that pixel once its closest-hit decisions are made in FP64 (the same FP32
shading, hit distances and barycentrics from an FP64 test, by brute force or over the
oracle's BVH), i.e. the pixel's value rests on an FP32 hit decision that precision can
flip.  Anything else is unexplained.
"""
from __future__ import annotations

import numpy as np

RTOL = 1e-5
ATOL = 1e-7


def close(a, b, rtol=RTOL, atol=ATOL):
    """Per-pixel agreement of (H, W, C) images: every channel within rtol relative (+atol)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return (np.abs(a - b) <= atol + rtol * np.abs(b)).all(axis=-1)


def classify(gpu, ref, alt):
    """gpu, ref: running sums of the HIP renderer and the FP32 oracle; alt: a callable giving
    the FP32 oracle's image with FP64 closest-hit decisions (only called when some pixel is
    outside the tolerance).  Returns a dict of counts and fractions."""
    ok = close(gpu, ref)
    exact = (np.asarray(gpu) == np.asarray(ref)).all(axis=-1)
    bad = ~ok
    sensitive = ~close(ref, alt()) if bad.any() else np.zeros_like(bad)
    return {
        "pixels": int(ok.size),
        "within_tol": float(ok.mean()),
        "bit_exact": float(exact.mean()),
        "outside": int(bad.sum()),
        "near_tie": int((bad & sensitive).sum()),
        "unexplained": int((bad & ~sensitive).sum()),
    }


def assert_parity(gpu, ref, alt, min_within=0.999):
    c = classify(gpu, ref, alt)
    msg = f"image parity {c}"
    print(msg, flush=True)
    assert c["within_tol"] >= min_within, msg
    assert c["unexplained"] == 0, msg
    return c
