"""Throughput benchmark of the crt path tracer on one MI355X.

Workload (one step = one sample pass of the render hot path): random triangle soup of
1,000,000 triangles (centroids uniform in the unit cube, vertex offsets N(0, 0.005^2)),
1920x1080 pinhole camera, 1 spp, primary rays plus 3 cosine-weighted diffuse bounces
(depth 4), compat estimator, constant sky, sun off.  Geometry and rays live in HBM
before the timed region starts; the BVH build is timed separately.

Output: one JSON line.  ``value`` is Mrays/s = (extension + shadow rays traced) / wall
time of the timed passes.  ``roofline`` covers the dominant kernel (closest-hit
traversal): algorithmic bytes = 48 B per ray (32 B ray in, 16 B hit out) over the
kernel's mean launch time measured with HIP events on the library's stream.
``cpu_baseline`` times the CPU oracle's restatement of the reference loop (FP32
binned-SAH BVH, one task per row, barrier per pass) on a band of rows of the same frame.

Three more measurements ride on the same line (skip with --no-nrc / --no-batch / --no-terrain):
``nrc`` -- the metric's second half: image MSE against a 4096-spp plain render of the same
scene at 1920x1080 (pbr, depth 16), plain tracer vs neural radiance cache, as MSE-vs-time
curves; ``time_ratio`` = the cache's time to reach the plain tracer's MSE at T over T.
``batch`` -- 64 independent scenes rendered by one batched call after the autotuner:
aggregate Mrays/s and per-scene latency.
``terrain`` -- the 14.8M-triangle terrain with objects at 1920x1080, 64 spp, pbr, sun + sky,
depth 8: Mrays/s and ms per spp.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import os
import subprocess
import sys
import tempfile
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Mrays/s at 1080p; equal-time MSE vs 4096-spp render, plain vs neural cache"
HBM_PEAK_GBS = 8000.0
BYTES_PER_RAY = 48


def _host_info() -> dict:
    """nproc, the cores this process may run on, and the CPU model (lscpu)."""
    info = {"nproc": os.cpu_count(), "affinity": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in ("Model name", "Socket(s)", "Thread(s) per core", "Core(s) per socket"):
                info[k.strip()] = v.strip()
    except (OSError, subprocess.SubprocessError):
        pass
    return info


def _baseline_library() -> tuple[str, str]:
    """The oracle's sources built for this host as the reference's Release build is
    (-O3 -march=native, FP contraction allowed); the prebuilt checker build otherwise."""
    out = Path(tempfile.mkdtemp(prefix="crt_baseline_")) / "liboracle_native.so"
    try:
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "baseline", f"BASELINE_OUT={out}"], check=True,
                       capture_output=True, timeout=300)
        return str(out), "-O3 -march=native"
    except (OSError, subprocess.SubprocessError):
        return str(ROOT / "oracle" / "liboracle.so"), "-O2 -march=x86-64-v2 -ffp-contract=off (native build failed)"


def cpu_baseline(spec, res, bounces, seconds_target: float) -> dict:
    """The oracle's restatement of the reference loop (one task per row, a barrier per sample
    pass, renderer.cpp:116-144,240-256) on this host's cores: median of 3 timed runs after one
    warm-up, each on the same band of rows sized for about seconds_target / 3."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle as O  # CPU baseline leg only

    libpath, flags = _baseline_library()
    O.use_library(libpath)
    host = _host_info()
    # every core the process may use; on a GPU box that is its CPU share (OMP_NUM_THREADS)
    threads = host["affinity"]
    if os.environ.get("OMP_NUM_THREADS", "").isdigit():
        threads = min(threads, int(os.environ["OMP_NUM_THREADS"]))
    t0 = time.perf_counter()
    bvh = O.Bvh(spec.corners())
    build_s = time.perf_counter() - t0
    W, H = res
    kw = dict(res=res, max_bounces=bounces, seed=0, trav=O.BVH32, bvh=bvh, nthreads=threads, aovs=False)
    # warm-up and calibration: one band of 64 rows
    r = O.render(spec, spp=1, rows=(H // 2 - 32, H // 2 + 32), **kw)
    per_row = max(r["stats"]["seconds"], 1e-6) / 64
    run_s = seconds_target / 3
    if per_row * H < run_s:
        nrows, r0 = H, 0
        spp = int(min(256, max(1, round(run_s / (per_row * H)))))
    else:
        nrows = int(min(H, max(8, run_s / per_row)))
        r0, spp = max(0, H // 2 - nrows // 2), 1
    runs = []
    for _ in range(3):
        st = O.render(spec, spp=spp, rows=(r0, r0 + nrows), **kw)["stats"]
        runs.append(st)
    rates = sorted((s["ext_rays"] + s["shadow_rays"]) / s["seconds"] / 1e6 for s in runs)
    med = runs[[((s["ext_rays"] + s["shadow_rays"]) / s["seconds"] / 1e6) for s in runs].index(rates[1])]
    rays = med["ext_rays"] + med["shadow_rays"]
    return {
        "value": rates[1],
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"median of 3 runs after a warm-up, each rows {r0}-{r0 + nrows} of the {W}x{H} frame x {spp} spp "
                  f"({nrows * W * spp} paths, {rays} rays, {med['seconds']:.1f} s); oracle restatement of "
                  f"renderer.cpp:240-384 built {flags}, one task per row, barrier per pass, CPU BVH (FP32 "
                  f"binned SAH, built in {build_s:.1f} s, untimed); Embree is not available on this image",
        "runs_mrays_s": rates,
        "host": host,
        "ref_counting_mrays_s": med["ref_segments"] / med["seconds"] / 1e6,
    }


# Sources that decide what the closest-hit traversal reads and writes: HBM traffic measured
# with profiler counters holds only for the build they were measured on.
TRAVERSAL_SOURCES = ("crender_amd/csrc/traverse.hip", "crender_amd/csrc/crt_math.h", "crender_amd/csrc/crt_internal.h",
                     "crender_amd/csrc/bvh_build.hip", "crender_amd/csrc/integrator.hip")
PMC_FILE = ROOT / "profiles" / "trace_closest_soup1m_1080p_pmc.json"
# the closest-hit traversal the headline workload launches (default width, stack, no instances)
TRACE_KERNEL = "k_trace<8, false, false, 16u, false, false>"


def traversal_source_hash() -> str:
    h = hashlib.sha256()
    for rel in TRAVERSAL_SOURCES:
        h.update(rel.encode())
        h.update((ROOT / rel).read_bytes())
    return h.hexdigest()[:16]


def _pmc_traffic(rays_per_launch: float):
    """HBM bytes per closest-hit launch (bytes per traced ray from the committed counter
    summary times the rays of an average timed launch), if the summary was measured on the
    current traversal sources; else (None, why)."""
    if not PMC_FILE.exists():
        return None, "no counter summary"
    pmc = json.loads(PMC_FILE.read_text())
    now = traversal_source_hash()
    if pmc.get("source_hash") != now:
        return None, (f"stale: {PMC_FILE.name} was measured on traversal sources {pmc.get('source_hash')}, "
                      f"the current ones hash to {now}")
    return (pmc["hbm_bytes_per_ray"] * rays_per_launch,
            f"{PMC_FILE.name} (rocprofv3 FETCH_SIZE + WRITE_SIZE per traced ray, copy-calibrated, "
            f"commit {pmc.get('commit')}, source {now})")


def _counter_rows(d: Path) -> list:
    import csv

    files = sorted(d.rglob("*counter_collection.csv"))
    if not files:
        raise SystemExit(f"no counter_collection.csv under {d}")
    return [r for f in files for r in csv.DictReader(open(f))]


def _per_dispatch(rows, counter: str, match=None) -> dict:
    """{dispatch id: (kernel name, summed counter value)} of the dispatches whose kernel
    name contains `match`."""
    out = {}
    for r in rows:
        if r["Counter_Name"] != counter or (match and match not in r["Kernel_Name"]):
            continue
        k = int(r["Dispatch_Id"])
        name, v = out.get(k, (r["Kernel_Name"], 0.0))
        out[k] = (name, v + float(r["Counter_Value"]))
    return dict(sorted(out.items()))


def _scale(rows, counter: str, known_bytes: float, match: str, launches: int) -> float:
    """Bytes moved per byte the counter reports: the `launches` largest dispatches of the
    calibration kernel each move `known_bytes`."""
    vals = sorted((v for _, v in _per_dispatch(rows, counter, match).values()), reverse=True)[:launches]
    if len(vals) < launches:
        raise SystemExit(f"calibration: {len(vals)} dispatches of '{match}' with {counter}")
    return known_bytes / (1024.0 * sum(vals) / len(vals))


def pmc_summary(d: Path) -> int:
    """Turns the counter runs under `d` into the traffic summary bench.py reads
    (profiles/trace_closest_soup1m_1080p_pmc.json) and prints a markdown table.  Expects
    d/fetch, d/write (FETCH_SIZE / WRITE_SIZE of a --warmup 0 headline run whose JSON line is
    d/fetch.json), d/copy_fetch, d/copy_write (--copy-calibration, JSON d/copy.json) and
    d/gather_fetch (--gather-calibration, JSON d/gather.json)."""
    run = json.loads((d / "fetch.json").read_text().strip().splitlines()[-1])
    if run.get("warmup") != 0:
        raise SystemExit("the counter run must have --warmup 0 (all its traversal launches are counted)")
    copy = json.loads((d / "copy.json").read_text().strip().splitlines()[-1])
    gather = json.loads((d / "gather.json").read_text().strip().splitlines()[-1])
    fs = _scale(_counter_rows(d / "copy_fetch"), "FETCH_SIZE", copy["copy_bytes_read"], "", copy["launches"])
    ws = _scale(_counter_rows(d / "copy_write"), "WRITE_SIZE", copy["copy_bytes_written"], "", copy["launches"])
    # scattered reads: the 128-byte lines the gathered 64-byte rows lie in, plus the index
    g_known = gather["unique_lines"] * gather["line_bytes"] + gather["index_bytes"]
    gs = _scale(_counter_rows(d / "gather_fetch"), "FETCH_SIZE", g_known, "gather", gather["launches"])
    fetch = _per_dispatch(_counter_rows(d / "fetch"), "FETCH_SIZE", TRACE_KERNEL)
    write = _per_dispatch(_counter_rows(d / "write"), "WRITE_SIZE", TRACE_KERNEL)
    if len(fetch) != len(write) or not fetch:
        raise SystemExit(f"{len(fetch)} FETCH_SIZE and {len(write)} WRITE_SIZE dispatches of {TRACE_KERNEL}")
    rays = run["rays"]["extension"]
    kib_f = sum(v for _, v in fetch.values())
    kib_w = sum(v for _, v in write.values())
    total = 1024.0 * (kib_f * fs + kib_w * ws)
    total_g = 1024.0 * (kib_f * gs + kib_w * ws)
    try:
        commit = subprocess.run(["git", "-C", str(ROOT), "rev-parse", "--short", "HEAD"], capture_output=True,
                                text=True, timeout=10).stdout.strip() or None
    except (OSError, subprocess.SubprocessError):
        commit = None
    summary = {
        "kernel": TRACE_KERNEL,
        "workload": run["config"]["workload"],
        "source_hash": traversal_source_hash(),
        "commit": commit,
        "dispatches": len(fetch),
        "extension_rays": rays,
        "fetch_scale": fs,
        "write_scale": ws,
        "gather_fetch_scale": gs,
        "hbm_bytes_per_ray": total / rays,
        # the same counts scaled as scattered 128-byte line reads (the gather calibration): node and
        # triangle fetches are scattered, so the copy scale may overstate them
        "hbm_bytes_per_ray_gather_scaled": total_g / rays,
        "algorithmic_bytes_per_ray": BYTES_PER_RAY,
    }
    PMC_FILE.write_text(json.dumps(summary, indent=1) + "\n")
    print(f"Copy calibration: FETCH_SIZE scale {fs:.3f}, WRITE_SIZE scale {ws:.3f} (bytes moved / bytes "
          f"counted).  Gather calibration ({gather['gather_rows']} random {gather['row_bytes']}-byte rows, "
          f"{gather['unique_lines']} distinct 128-byte lines): FETCH_SIZE scale {gs:.3f} (line bytes + index "
          f"bytes / bytes counted).\n")
    print("| dispatch | FETCH_SIZE KiB | WRITE_SIZE KiB | scaled MB |\n|---|---|---|---|")
    for k, (kf, kw) in zip(fetch, zip(fetch.values(), write.values())):
        print(f"| {k} | {kf[1]:.0f} | {kw[1]:.0f} | {1024 * (kf[1] * fs + kw[1] * ws) / 1e6:.1f} |")
    print(f"\n{rays} extension rays over {len(fetch)} dispatches: {total / rays:.1f} bytes per ray copy-scaled, "
          f"{total_g / rays:.1f} gather-scaled (algorithmic {BYTES_PER_RAY}), {total / len(fetch) / 1e6:.1f} MB "
          f"per dispatch on average copy-scaled; "
          f"written to {PMC_FILE.relative_to(ROOT) if PMC_FILE.is_relative_to(ROOT) else PMC_FILE} "
          f"(source {summary['source_hash']}, commit {commit}).")
    return 0


def _mse(img, ref) -> float:
    d = img.astype(np.float64) - ref
    return float(np.mean(d * d))


def _curve(r, checkpoints, ref, t_start=0.0):
    """Renders up to each checkpoint spp; returns [(spp, seconds of rendering, MSE)]."""
    out, done, t = [], 0, t_start
    for c in checkpoints:
        t0 = time.perf_counter()
        r.render(c - done)
        t += time.perf_counter() - t0
        done = c
        out.append((c, t, _mse(r.raw() / c, ref)))
    return out


def _time_to_reach(curve, target):
    """Seconds at which a (spp, seconds, MSE) curve first reaches `target` MSE (log-log
    interpolation between checkpoints); None if it never does."""
    if curve and curve[0][2] <= target:
        return curve[0][1]
    for (_, t0, m0), (_, t1, m1) in zip(curve, curve[1:]):
        if m0 > target >= m1:
            f = (np.log(m0) - np.log(target)) / (np.log(m0) - np.log(m1))
            return float(np.exp(np.log(t0) + f * (np.log(t1) - np.log(t0))))
    return None


def nrc_equal_time(ctx, sc, ref_spp: int) -> dict:
    """Equal-time MSE against a ref_spp plain render at 1920x1080: plain tracer vs the neural
    radiance cache (warm-up passes counted), pbr estimator, depth 16 with Russian roulette
    from bounce 3, the soup scene of the Mrays/s workload."""
    from crender_amd import Nrc, Renderer

    res, bounces, rr = (1920, 1080), 16, 3

    def mk(seed):
        r = Renderer(ctx, sc, res[0], res[1], bounces)
        r.set_resolution(*res)
        r.set_options(mode=1, rr_start=rr, seed=seed)
        return r

    t0 = time.perf_counter()
    ref = mk(1000)
    done = 0
    while done < ref_spp:
        k = min(512, ref_spp - done)
        ref.render(k)
        done += k
        print(f"# nrc reference {done}/{ref_spp} spp, {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    refimg = ref.raw() / ref_spp
    ref.close()
    plain = mk(1)
    pc = _curve(plain, [4, 8, 16, 32, 64], refimg)
    plain.close()
    nrc = Nrc(ctx, seed=5)
    r = mk(2)
    warm = 4
    r.set_nrc(nrc, warmup_passes=warm)
    tw0 = time.perf_counter()
    r.render(warm)
    t_warm = time.perf_counter() - tw0
    nc = _curve(r, [4, 8, 16, 32, 64], refimg, t_start=t_warm)   # the plain tracer's checkpoints
    st = r.current_stats()
    r.close()
    nrc.close()
    T = pc[-1][1]
    target = pc[-1][2]
    t_match = _time_to_reach(nc, target)
    return {
        "workload": "soup 1M tris, 1920x1080, pbr, depth 16, RR from bounce 3; cache: term bounce 1 (footprint c=0.01), "
                    "65536 training paths/pass, 4 bounces past the query, 24 Adam steps/pass, 4 warm-up passes",
        "reference_spp": ref_spp,
        "plain": [{"spp": a, "seconds": b, "mse": c} for a, b, c in pc],
        "cache": [{"spp": a, "seconds": b, "mse": c} for a, b, c in nc],
        "cache_warmup_seconds": t_warm,
        "cache_train_gpu_ms_per_pass": st["ms_nrc_train"] / max(1, warm + 64),   # GPU time of the training steps
        "equal_time_seconds": T,
        "plain_mse_at_T": target,
        "cache_seconds_to_plain_mse": t_match,
        "time_ratio": (t_match / T) if t_match else None,
    }


def traversal_counts(spec, res, bounces, args) -> dict:
    """Node visits, leaf visits and triangle tests per traced ray, from one pass of the same
    workload with the instrumented traversal kernel (outside the timed region)."""
    from crender_amd import Context, Renderer, scene_from_spec

    ctx = Context(0, bvh_width=args.width, max_leaf=args.max_leaf, trace_waves_per_cu=args.waves_per_cu,
                  trace_stack=args.stack, trace_refill=args.refill, trace_park_min=args.park_min,
                  trace_leaf_rounds=args.leaf_rounds, bvh_prim_cost=args.prim_cost,
                  traversal_stats=True)
    sc = scene_from_spec(ctx, spec)
    r = Renderer(ctx, sc, res[0], res[1], bounces)
    r.set_resolution(*res)
    r.set_options(mode=0, rr_start=0, seed=0)
    r.render(1)
    st = r.current_stats()
    r.close()
    sc.close()
    ctx.close()
    rays = max(1, st["extension_rays"] + st["shadow_rays"])
    return {"nodes_per_ray": st["node_visits"] / rays, "leaves_per_ray": st["leaf_visits"] / rays,
            "tris_per_ray": st["tri_tests"] / rays, "rays": rays,
            # lanes busy per wave iteration of the node / leaf loop (SIMD utilisation)
            "node_loop_lanes": st["node_visits"] / max(1, 64 * st["wave_node_steps"]),
            "leaf_loop_lanes": st["leaf_visits"] / max(1, 64 * st["wave_leaf_steps"]),
            # idle lanes of node-loop iterations by cause, as fractions of all lanes
            "node_loop_idle_no_ray": st["node_idle_no_ray"] / max(1, 64 * st["wave_node_steps"]),
            "node_loop_idle_done": st["node_idle_done"] / max(1, 64 * st["wave_node_steps"]),
            "node_loop_idle_leaf_wait": st["node_idle_leaf_wait"] / max(1, 64 * st["wave_node_steps"])}


def terrain_bench(args) -> dict:
    """The terrain with objects (2047^2 x 2 terrain triangles + 20k icospheres, 14.8M
    triangles), 1920x1080, pbr, sun + sky, depth 8, Russian roulette from bounce 3: a 64-spp
    render after an 8-spp warm-up, geometry resident in HBM."""
    from crender_amd import Context, Renderer, scene_from_spec, scenes

    spec = scenes.terrain()
    W, H = spec.res
    ctx = Context(0, bvh_width=args.width, max_leaf=args.max_leaf, trace_waves_per_cu=args.waves_per_cu,
                  trace_stack=args.stack, trace_refill=args.refill, trace_park_min=args.park_min,
                  trace_leaf_rounds=args.leaf_rounds, bvh_prim_cost=args.prim_cost,
                  wavefront_paths=args.wavefront_paths)
    sc = scene_from_spec(ctx, spec)
    first_ms, rebuild_ms, info = rebuilds(sc, 3)
    r = Renderer(ctx, sc, W, H, spec.max_bounces)
    r.set_resolution(W, H)
    r.set_options(mode=spec.mode, rr_start=spec.rr_start, seed=0)
    r.render(8)
    s0 = r.current_stats()
    t0 = time.perf_counter()
    r.render(spec.spp)
    wall = time.perf_counter() - t0
    s1 = r.current_stats()
    r.close()
    sc.close()
    ctx.close()
    ext = s1["extension_rays"] - s0["extension_rays"]
    shadow = s1["shadow_rays"] - s0["shadow_rays"]
    return {
        "workload": f"terrain with objects, {spec.ntri} tris, {W}x{H}, {spec.spp} spp, pbr, sun + sky, "
                    f"depth {spec.max_bounces}, roulette from bounce {spec.rr_start}",
        "mrays_per_s": (ext + shadow) / wall / 1e6,
        "ms_per_spp": 1e3 * wall / spec.spp,
        "extension_rays": ext,
        "shadow_rays": shadow,
        "bvh_build_ms": float(np.median(rebuild_ms)),
        "bvh_mtris_per_s": spec.ntri / (float(np.median(rebuild_ms)) * 1e-3) / 1e6,
        "bvh_first_build_ms": first_ms,
        "bvh_rebuild_ms": rebuild_ms,
    }


def batch_bench(seed: int = 7) -> dict:
    """64 independent scenes (soup / terrain, 1k-500k tris, 128^2-1024^2, 1-64 spp, depth 6)
    rendered by crt_render_batch after crt_autotune."""
    from crender_amd import Context, autotune, render_batch, scene_from_spec, scenes

    suite = scenes.batch_suite(64, seed)
    ctx = Context(0)
    scs = [scene_from_spec(ctx, spec) for spec, _, _, _ in suite]
    jobs = [dict(scene=sc, width=w, height=h, spp=spp, max_bounces=6, mode=spec.mode, rr_start=0, seed=k)
            for k, (sc, (spec, w, h, spp)) in enumerate(zip(scs, suite))]
    tune = autotune(ctx, jobs)
    print(f"# batch autotune {tune}", file=sys.stderr, flush=True)
    _, ms, st = render_batch(ctx, jobs)
    for sc in scs:
        sc.close()
    ctx.close()
    lat = np.sort(np.asarray(ms))
    return {
        "workload": "64 scenes (soup / terrain), 1k-500k tris log-uniform, 128^2-1024^2, spp 1/4/16/64, depth 6",
        "triangles": int(sum(spec.ntri for spec, _, _, _ in suite)),
        "paths": int(sum(w * h * spp for _, w, h, spp in suite)),
        "mrays_per_s": st["rays_per_second"] / 1e6,
        "ms_total": st["ms_total"],
        "extension_rays": st["extension_rays"],
        "shadow_rays": st["shadow_rays"],
        "replays": st["replays"],
        "used_graph": bool(st["used_graph"]),
        "latency_ms": {"p50": float(np.percentile(lat, 50)), "p90": float(np.percentile(lat, 90)),
                       "max": float(lat[-1])},
        "tuned": tune,
    }


def rebuilds(sc, n: int):
    """The scene commit's build time, then n rebuilds of the same scene (the builder keeps
    its scratch memory from the first build on); returns (first, [rebuilds], bvh info)."""
    first = sc.bvh_info()["build_ms"]
    ms = []
    for _ in range(n):
        sc.commit()
        ms.append(round(sc.bvh_info()["build_ms"], 3))
    return first, ms, sc.bvh_info()


def build_only(args) -> int:
    """BVH build throughput of the triangle soup: the commit, then rebuilds of the same scene
    (the first build in a process also maps the builder's scratch pool)."""
    from crender_amd import Context, scene_from_spec, scenes

    spec = scenes.soup(args.tris, seed=2)
    ctx = Context(0, bvh_width=args.width, max_leaf=args.max_leaf, bvh_prim_cost=args.prim_cost)
    sc = scene_from_spec(ctx, spec)
    ms = [sc.bvh_info()["build_ms"]]
    for _ in range(args.build_only - 1):
        sc.commit()
        ms.append(sc.bvh_info()["build_ms"])
    info = sc.bvh_info()
    rebuilds = ms[1:] or ms
    print(json.dumps({"metric": "BVH build", "tris": args.tris, "build_ms": [round(x, 3) for x in ms],
                      "first_build_ms": ms[0], "median_rebuild_ms": float(np.median(rebuilds)),
                      "mtris_per_s": args.tris / (float(np.median(rebuilds)) * 1e-3) / 1e6,
                      "wide_nodes": info["node_count"], "binary_nodes": info["bvh2_node_count"],
                      "sah_cost": info["sah_cost"], "depth": info["depth"]}), flush=True)
    sc.close()
    ctx.close()
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tris", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=0, help="BVH width (0 = library default)")
    ap.add_argument("--waves-per-cu", type=int, default=0, help="persistent traversal occupancy (0 = default)")
    ap.add_argument("--stack", type=int, default=0, help="LDS stack entries per traversal lane (0 = default)")
    ap.add_argument("--max-leaf", type=int, default=0, help="BVH leaf size limit (0 = library default)")
    ap.add_argument("--refill", type=int, default=0, help="idle lanes that make a traversal wave fetch rays (0 = default)")
    ap.add_argument("--prim-cost", type=float, default=0.0,
                    help="triangle-test cost relative to a wide-node test in the BVH collapse (0 = default)")
    ap.add_argument("--leaf-rounds", type=int, default=0,
                    help="leaf-loop rounds per leaf phase (0 = default, above 64 = unlimited)")
    ap.add_argument("--park-min", type=int, default=0, help="searching lanes below which a wave tests its parked leaves "
                                                            "(0 = default)")
    ap.add_argument("--wavefront-paths", type=int, default=0, help="camera paths per wavefront (0 = default)")
    ap.add_argument("--sort-rays", action="store_true", help="reorder secondary rays before tracing")
    ap.add_argument("--sort-materials", action="store_true", help="shade each bounce's hits grouped by material")
    ap.add_argument("--cpu-seconds", type=float, default=15.0, help="CPU work of the 3 timed baseline runs")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-rendering measurement")
    ap.add_argument("--no-terrain", action="store_true", help="skip the 14.8M-triangle terrain measurement")
    ap.add_argument("--no-nrc", action="store_true", help="skip the neural radiance cache equal-time MSE")
    ap.add_argument("--nrc-ref-spp", type=int, default=4096)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (read + write) 5 times, for scaling "
                         "profiler byte counters; prints the bytes moved per launch")
    ap.add_argument("--gather-calibration", type=int, default=0, metavar="MIB",
                    help="only gather 2^22 random 64-byte rows of an MIB-MiB table 5 times (index_select), for "
                         "scaling profiler byte counters on scattered reads; prints the bytes moved per launch")
    ap.add_argument("--pmc-summary", type=Path, default=None, metavar="DIR",
                    help="only summarise the counter runs under DIR into the committed traffic figure")
    ap.add_argument("--build-only", type=int, default=0, metavar="N",
                    help="only build the soup's BVH N times (the scene commit, then N - 1 rebuilds) and print "
                         "each build's time")
    args = ap.parse_args(argv)
    if args.pmc_summary:
        return pmc_summary(args.pmc_summary)
    if args.gather_calibration:
        import torch

        rows = args.gather_calibration * (1 << 20) // 64
        table = torch.zeros((rows, 16), dtype=torch.float32, device="cuda:0")
        idx_h = np.random.default_rng(0).integers(0, rows, 1 << 22)
        idx = torch.from_numpy(idx_h).to("cuda:0")
        for _ in range(5):
            out = table.index_select(0, idx)   # reads 8 B of index and one 64-byte row per output row
        torch.cuda.synchronize()
        del out
        print(json.dumps({"gather_rows": 1 << 22, "unique_rows": int(np.unique(idx_h).size), "row_bytes": 64,
                          "unique_lines": int(np.unique(idx_h // 2).size), "line_bytes": 128,
                          "index_bytes": 8 << 22, "launches": 5}), flush=True)
        return 0
    if args.copy_calibration:
        import torch

        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(5):
            torch.mul(src, 1.0, out=dst)   # one elementwise kernel: reads 4n bytes, writes 4n bytes
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "launches": 5}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (each box has one GPU)",
              file=sys.stderr)
        return 2
    if args.build_only:
        return build_only(args)

    from crender_amd import Context, Renderer, scene_from_spec, scenes

    res, bounces = (1920, 1080), 4
    spec = scenes.soup(args.tris, seed=2, res=res)
    ctx = Context(0, bvh_width=args.width, max_leaf=args.max_leaf, phase_timing=True,
                  trace_waves_per_cu=args.waves_per_cu, trace_stack=args.stack,
                  trace_refill=args.refill, trace_park_min=args.park_min,
                  trace_leaf_rounds=args.leaf_rounds, bvh_prim_cost=args.prim_cost, sort_rays=args.sort_rays,
                  sort_materials=args.sort_materials, wavefront_paths=args.wavefront_paths)
    dev = ctx.device_name()
    sc = scene_from_spec(ctx, spec)
    first_ms, rebuild_ms, info = rebuilds(sc, 5)
    r = Renderer(ctx, sc, res[0], res[1], bounces)
    r.set_resolution(*res)
    r.set_options(mode=0, rr_start=0, seed=0)
    r.render(args.warmup)
    s0 = r.current_stats()
    t0 = time.perf_counter()
    r.render(args.steps)          # synchronous: returns after the last pass completed
    t1 = time.perf_counter()
    s1 = r.current_stats()
    wall = t1 - t0
    ext = s1["extension_rays"] - s0["extension_rays"]
    shadow = s1["shadow_rays"] - s0["shadow_rays"]
    segs = s1["total_rays"] - s0["total_rays"]
    rays = ext + shadow
    ms_ext = s1["ms_extend"] - s0["ms_extend"]
    ms_busy = s1["ms_extend_busy"] - s0["ms_extend_busy"]
    launches = s1["extend_launches"] - s0["extend_launches"]
    gpu_ms = s1["ms_render"] - s0["ms_render"]
    achieved = BYTES_PER_RAY * ext / (ms_ext * 1e-3) / 1e9 if ms_ext > 0 else 0.0
    traffic, traffic_note = _pmc_traffic(ext / max(launches, 1))
    out = {
        "metric": METRIC,
        "value": rays / wall / 1e6,
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * wall / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"triangle soup {args.tris} tris, 1920x1080, 1 spp, primary + 3 diffuse bounces "
                        f"(depth 4), compat estimator, sun off, constant sky",
            "bvh_width": info["width"],
            "waves_per_cu": args.waves_per_cu or 8,
            "trace_stack": args.stack or 16,
            # sample passes per wavefront: floor(wavefront paths / (W*H)) in [1, 8], 2^24 paths by default
            "passes_per_wavefront": max(1, min(8, (args.wavefront_paths or (1 << 24)) // (res[0] * res[1]))),
            "wavefronts_in_flight": 2,
            "sort_rays": args.sort_rays,
            "sort_materials": args.sort_materials,
            "device": dev,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "traffic_source": traffic_note,
            "kernel": "k_trace (closest hit)",
            "bytes_per_launch": BYTES_PER_RAY * ext / max(launches, 1),
            "mean_launch_ms": ms_ext / max(launches, 1),
            # two wavefronts are in flight on two streams, so closest-hit launches overlap and
            # each one's span covers time shared with the other: the same bytes over the time
            # during which any closest-hit launch ran
            "busy_ms": ms_busy,
            "achieved_busy": BYTES_PER_RAY * ext / (ms_busy * 1e-3) / 1e9 if ms_busy > 0 else 0.0,
        },
        "rays": {"extension": ext, "shadow": shadow, "ref_counting_segments": segs},
        "samples_per_s": res[0] * res[1] * args.steps / wall,
        "ref_counting_mrays_s": segs / wall / 1e6,
        "gpu_ms_per_step": gpu_ms / args.steps,
        "bvh": {"build_ms": float(np.median(rebuild_ms)),
                "mtris_per_s": args.tris / (float(np.median(rebuild_ms)) * 1e-3) / 1e6,
                "first_build_ms": first_ms, "rebuild_ms": rebuild_ms,
                "nodes": info["node_count"], "depth": info["depth"], "sah_cost": info["sah_cost"],
                "bytes": info["bytes"]},
    }
    r.close()
    if not args.no_nrc:
        out["nrc"] = nrc_equal_time(ctx, sc, args.nrc_ref_spp)
    sc.close()
    ctx.close()
    out["traversal"] = traversal_counts(spec, res, bounces, args)
    if not args.no_terrain:
        out["terrain"] = terrain_bench(args)
    if not args.no_batch:
        out["batch"] = batch_bench()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(spec, res, bounces, args.cpu_seconds)
        out["gpu_over_cpu"] = out["value"] / out["cpu_baseline"]["value"]
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
