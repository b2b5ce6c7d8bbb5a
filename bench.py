"""Benchmark of the path-tracing hot path on one MI355X.

Default workload: triangle-mesh path tracing of a noise-displaced geodesic
icosphere of 1,003,520 triangles (frequency 224, the nearest geodesic count to
1M) on a ground plane, Plastic (roughness 0.2) over Matte, one point light
plus a Perez sun/sky, 1920x1080 film, maxBounce 8, Russian roulette from
bounce 5.  One step = one render call of --spp samples per pixel over the
whole film (64 by default: one converging frame).

Prints ONE JSON line with the throughput (path samples/s), Mrays/s, the
roofline of the dominant kernel (BVH closest-hit traversal of the wavefront
extend stage, timed with HIP events on the library's stream inside the timed
region) and a CPU baseline (the oracle, one thread, on a bounded pixel
window of the same frame).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
METRIC = "path samples/s + Mrays/s; p50/p99 request latency, batched vs plain kernel"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(sc, width, height, spp, max_bounce, min_rr, budget_s):
    """Oracle (float32 restatement of CPURenderer), one thread, on a centred
    pixel window sized to take about `budget_s` seconds."""
    from tests.oracle_lib import OracleScene
    o = OracleScene(sc)
    cx, cy = width // 2, height // 2
    # probe a small window to estimate the rate
    w0 = 8
    t = time.perf_counter()
    o.render(width, height, spp, seed=99, max_bounce=max_bounce, min_rr=min_rr, threads=1,
             window=(cx - w0 // 2, cy - w0 // 2, cx + w0 // 2, cy + w0 // 2))
    dt = max(time.perf_counter() - t, 1e-3)
    rate = w0 * w0 * spp / dt
    side = int(max(8, min(256, (rate * budget_s / spp) ** 0.5)))
    win = (cx - side // 2, cy - side // 2, cx - side // 2 + side, cy - side // 2 + side)
    t = time.perf_counter()
    o.render(width, height, spp, seed=99, max_bounce=max_bounce, min_rr=min_rr, threads=1, window=win)
    dt = time.perf_counter() - t
    o.close()
    return {"value": side * side * spp / dt, "unit": "path samples/s", "cores": 1, "kind": "port",
            "sample": f"{side}x{side} pixel window at the image centre, {spp} spp, one thread, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--triangles", type=int, default=1_000_000)
    ap.add_argument("--width", type=int, default=1920)
    ap.add_argument("--height", type=int, default=1080)
    ap.add_argument("--max-bounce", type=int, default=8)
    ap.add_argument("--kernel", choices=["wavefront", "plain"], default="wavefront")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-perez", action="store_true")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (multi-GPU sharding is a design "
                                   "note only)", "n_gpus": args.gpus}))
        sys.exit(2)

    from dayflower_amd import scene as S
    from dayflower_amd.renderer import GPURenderer

    t0 = time.perf_counter()
    sc = S.icosphere_scene(args.triangles, seed=0x5EED, width=args.width, height=args.height,
                           with_perez=not args.no_perez)
    log(f"bench: scene {sc.n_triangles} triangles generated in {time.perf_counter() - t0:.1f} s")
    r = GPURenderer()
    r.setScene(sc)
    r.setSeed(0x5EED)
    r.setMaximumBounce(args.max_bounce)
    r.setMinimumBounceRussianRoulette(5)
    r.setKernel(0 if args.kernel == "wavefront" else 1)
    t0 = time.perf_counter()
    r.setup()
    log(f"bench: setup (BVH build + upload) {time.perf_counter() - t0:.1f} s")

    # algorithmic traversal counts of the extend stage (untimed pass)
    r.setCountTraversal(True)
    r.render(max(1, args.spp // 8))
    cst = r.stats()
    r.setCountTraversal(False)
    nodes_per_ray = cst["extend_nodes"] / max(1, cst["extend_rays"])
    tris_per_ray = cst["extend_tris"] / max(1, cst["extend_rays"])
    log(f"bench: extend visits per ray: {nodes_per_ray:.2f} nodes, {tris_per_ray:.2f} triangles")

    for i in range(args.warmup):
        r.render(args.spp)
        log(f"bench: warmup {i + 1}/{args.warmup} {r.stats()['seconds']:.3f} s")
    r.clear()
    r.setProfile(args.kernel == "wavefront")
    tot = {"samples": 0, "closest": 0, "any": 0, "ext_s": 0.0, "ext_launch": 0, "ext_rays": 0, "dev_s": 0.0}
    t0 = time.perf_counter()
    for i in range(args.steps):
        r.render(args.spp)
        st = r.stats()
        tot["samples"] += st["path_samples"]
        tot["closest"] += st["rays_closest"]
        tot["any"] += st["rays_any"]
        tot["ext_s"] += st["extend_seconds"]
        tot["ext_launch"] += st["extend_launches"]
        tot["ext_rays"] += st["extend_rays"]
        tot["dev_s"] += st["seconds"]
        log(f"bench: step {i + 1}/{args.steps} {st['seconds']:.3f} s")
    wall = time.perf_counter() - t0
    r.setProfile(False)
    r.dispose()

    value = tot["samples"] / wall
    roofline = None
    if tot["ext_s"] > 0:
        # algorithmic bytes per closest-hit ray: ray in (32) + hit out (16) +
        # 64 B per two-wide node fetched + 36 B of positions per triangle test
        bytes_per_ray = 32 + 16 + 64 * nodes_per_ray + 36 * tris_per_ray
        traffic = bytes_per_ray * tot["ext_rays"] / max(1, tot["ext_launch"])
        achieved = bytes_per_ray * tot["ext_rays"] / tot["ext_s"] / 1e9
        roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                    "kernel": "k_extend (BVH closest hit)", "algorithmic_bytes_per_launch": round(traffic),
                    "avg_launch_ms": round(tot["ext_s"] / max(1, tot["ext_launch"]) * 1e3, 3),
                    "bytes_per_ray": round(bytes_per_ray, 1)}
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(sc, args.width, args.height, args.spp, args.max_bounce, 5, args.cpu_seconds)
        except Exception as e:  # the baseline is informative; never hide the GPU number
            cpu = {"error": str(e)}
    out = {
        "metric": METRIC,
        "value": round(value, 1),
        "unit": "path samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene generator, no dataset)",
        "config": {"workload": f"displaced icosphere path tracing ({sc.n_triangles} triangles), "
                               f"{args.width}x{args.height}, {args.spp} spp per step, maxBounce {args.max_bounce}, "
                               f"RR from 5, {'Perez sky + ' if not args.no_perez else ''}point light",
                   "kernel": args.kernel, "triangles": sc.n_triangles, "width": args.width, "height": args.height,
                   "spp": args.spp, "max_bounce": args.max_bounce},
        "mrays_per_s": round((tot["closest"] + tot["any"]) / wall / 1e6, 1),
        "mrays_closest_per_s": round(tot["closest"] / wall / 1e6, 1),
        "mrays_any_per_s": round(tot["any"] / wall / 1e6, 1),
        "device_seconds": round(tot["dev_s"], 3),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
