"""Structure of the device-built wide BVH (crt_scene_bvh_export), decoded in numpy.

Every leaf-order position belongs to exactly one leaf, the primID table is a permutation,
and every child box -- FP32 (4-wide) or 8-bit quantised (8-wide) -- contains the bounds of
every triangle below it.  Sizes cover the multi-workgroup binning stage, the per-segment
workgroup stage and its per-wave subtree stage (segments <= 2048, subtrees <= 128), plus
degenerate inputs (all centroids equal, a single triangle)."""
import numpy as np
import pytest

from crender_amd import Context, scene_from_spec, scenes

pytestmark = pytest.mark.gpu
LEAF = 0x80000000


def _decode(nodes, width):
    nodes = nodes.reshape(-1, 128)
    out = []
    for nd in nodes:
        if width == 4:
            f = nd[:96].view(np.float32).reshape(6, 4)
            lo = np.stack([f[0], f[2], f[4]], 1)
            hi = np.stack([f[1], f[3], f[5]], 1)
            ch = nd[96:112].view(np.uint32)
            nc = int(nd[112:116].view(np.uint32)[0])
        else:
            org = nd[:12].view(np.float32)
            meta = int(nd[12:16].view(np.uint32)[0])
            ex = [(meta >> (8 * k)) & 0xFF for k in range(3)]
            nc = meta >> 24
            ch = nd[16:48].view(np.uint32)
            q = nd[48:96].reshape(6, 8).astype(np.float64)
            sc = np.array([2.0 ** (e - 127) for e in ex])
            lo = org[None, :].astype(np.float64) + np.stack([q[0], q[2], q[4]], 1) * sc
            hi = org[None, :].astype(np.float64) + np.stack([q[1], q[3], q[5]], 1) * sc
        out.append((lo[:nc], hi[:nc], ch[:nc].astype(np.int64)))
    return out


def _check(spec, width):
    ctx = Context(0, bvh_width=width)
    sc = scene_from_spec(ctx, spec)
    nodes, prims = sc.bvh_export()
    info = sc.bvh_info()
    sc.close()
    ctx.close()
    _check_tree(spec, width, nodes, prims, info)


def _check_tree(spec, width, nodes, prims, info):
    ntri = spec.ntri
    assert np.array_equal(np.sort(prims), np.arange(ntri))
    tri = spec.corners().reshape(-1, 3, 3).astype(np.float64)[prims]     # leaf order
    tlo, thi = tri.min(1), tri.max(1)
    dec = _decode(nodes, width)
    assert len(dec) == info["node_count"]
    covered = np.zeros(ntri, np.int64)

    def visit(ref):
        """bounds of the subtree of a child reference; checks containment on the way."""
        if ref & LEAF:
            first, cnt = ref & 0x07FFFFFF, ((ref >> 27) & 0xF) + 1
            covered[first:first + cnt] += 1
            return tlo[first:first + cnt].min(0), thi[first:first + cnt].max(0)
        lo, hi, ch = dec[ref]
        blo, bhi = np.full(3, np.inf), np.full(3, -np.inf)
        for j, c in enumerate(ch):
            slo, shi = visit(int(c))
            assert np.all(lo[j] <= slo) and np.all(hi[j] >= shi), (ref, j)
            blo, bhi = np.minimum(blo, slo), np.maximum(bhi, shi)
        return blo, bhi

    import sys
    sys.setrecursionlimit(10000)
    visit(0)
    assert np.all(covered == 1)


@pytest.mark.parametrize("width", [4, 8])
@pytest.mark.parametrize("n", [1, 50, 3000, 100_000])
def test_soup_bvh_structure(width, n):
    print(f"bvh width={width} n={n}", flush=True)
    _check(scenes.soup(n, seed=17), width)


def test_coincident_centroids():
    spec = scenes.soup(600, seed=3)
    v = spec.vertices.reshape(-1, 3, 3).copy()
    v[:] = v[0]                      # every triangle identical
    spec.vertices = v.reshape(-1, 3)
    _check(spec, 8)


def test_concurrent_commits_share_the_build_scratch():
    """Builds on one GPU take turns on the builder's scratch buffer: two threads committing
    scenes of different sizes on two contexts at once both get complete, correct trees (the
    second size also makes the buffer grow while the other thread may be building)."""
    import threading

    specs = [scenes.soup(30_000, seed=21), scenes.soup(120_000, seed=22)]
    errors, exported = [], [None, None]

    def work(k):
        try:
            ctx = Context(0)
            sc = scene_from_spec(ctx, specs[k])
            for _ in range(4):
                sc.commit()
            exported[k] = (sc.bvh_export(), sc.bvh_info())
            sc.close()
            ctx.close()
        except Exception as e:   # reported by the main thread
            errors.append(e)

    print("bvh: two threads committing", flush=True)
    threads = [threading.Thread(target=work, args=(k,)) for k in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join(120)
    assert not errors, errors
    for k in range(2):
        (nodes, prims), info = exported[k]
        _check_tree(specs[k], 8, nodes, prims, info)
