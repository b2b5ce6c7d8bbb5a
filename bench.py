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
region; algorithmic bytes priced per ray as 48 + 32 N_node + 36 N_tri with
N_node and N_tri counted by the oracle's FP64 walk of the same BVH over a
sample of the frame's own rays), the CPU baseline (the oracle on all host
cores and on one core, on a bounded pixel window of the same frame, checked
against the GPU film of the same random streams) and the batched leg (64
mixed jobs: the plain per-job kernel against the packed wavefront launch).
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


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_icosphere1m_1080p.txt")
DEFAULT_FRAME = {"triangles": 1_000_000, "width": 1920, "height": 1080, "spp": 64, "max_bounce": 8, "no_perez": False,
                 "scene": "icosphere", "kernel": "wavefront"}


def hbm_traffic(kernel, args):
    """HBM bytes per launch of `kernel` on the default frame workload, as
    measured by rocprofv3 FETCH_SIZE + WRITE_SIZE and scaled by the same
    counters on a copy of known size (lines `bytes_per_dispatch.<kernel> = N`
    of the committed traffic summary); None for other workloads."""
    if any(getattr(args, k) != v for k, v in DEFAULT_FRAME.items()) or not os.path.exists(TRAFFIC_FILE):
        return None
    with open(TRAFFIC_FILE) as f:
        for line in f:
            key, _, val = line.partition("=")
            if key.strip() == f"bytes_per_dispatch.{kernel}":
                return int(float(val))
    return None


def copy_calibration(mib):
    """A fill and a device-to-device copy of `mib` MiB before the render, so a
    counter run of this process also records kernels of known traffic (the
    fill writes mib MiB; the copy reads and writes mib MiB each)."""
    import torch
    n = int(mib) << 20
    a = torch.empty(n, dtype=torch.uint8, device="cuda")
    b = torch.empty_like(a)
    a.fill_(7)
    torch.cuda.synchronize()
    b.copy_(a)
    torch.cuda.synchronize()
    log(f"bench: copy calibration: fill {mib} MiB, copy {mib} MiB")
    del a, b


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(sc, width, height, spp, max_bounce, min_rr, seed, budget_s, gpu_film=None):
    """The oracle (float32 restatement of CPURenderer) on a centred pixel
    window of the same frame and random streams: on all host cores (the
    reported value, window sized to about `budget_s` seconds) and on one
    core (about a third of that).  When the GPU film of the same streams is
    given, the window is compared with it path by path."""
    import numpy as np
    from tests.oracle_lib import OracleScene
    o = OracleScene(sc)
    cores = max(1, min(16, os.cpu_count() or 1))
    cx, cy = width // 2, height // 2

    def window(side):
        return (cx - side // 2, cy - side // 2, cx - side // 2 + side, cy - side // 2 + side)

    def timed(side, threads):
        t = time.perf_counter()
        out = o.render(width, height, spp, seed=seed, max_bounce=max_bounce, min_rr=min_rr, threads=threads,
                       window=window(side))
        return time.perf_counter() - t, out

    # rate probes on small windows, then windows sized to the budget
    dt1, _ = timed(8, 1)
    rate1 = 64 * spp / max(dt1, 1e-3)
    side1 = int(max(8, min(256, (rate1 * budget_s / 3.0 / spp) ** 0.5)))
    dt1, _ = timed(side1, 1)
    side = int(max(16, min(1024, (rate1 * cores * budget_s / spp) ** 0.5)))
    dt, (s, q, c) = timed(side, cores)
    o.close()
    res = {"value": side * side * spp / dt, "unit": "path samples/s", "cores": cores, "kind": "port",
           "sample": f"{side}x{side} pixel window at the image centre, {spp} spp, {cores} threads, {dt:.1f} s",
           "cpu_model": _cpu_model(),
           "one_core": {"value": round(side1 * side1 * spp / dt1, 1), "cores": 1,
                        "sample": f"{side1}x{side1} window, {spp} spp, {dt1:.1f} s"}}
    if gpu_film is not None:
        x0, y0, x1, y1 = window(side)
        ref = (s / np.maximum(c, 1)[..., None])[y0:y1, x0:x1]
        g = gpu_film[y0:y1, x0:x1]
        px = np.isclose(g, ref, rtol=1e-3, atol=1e-4).all(-1).mean()
        rel = abs(float(g.mean()) - float(ref.mean())) / max(float(ref.mean()), 1e-12)
        res["same_streams_vs_gpu"] = {"pixels_equal": round(float(px), 5), "mean_rel_diff": float(f"{rel:.3g}"),
                                      "window": f"{side}x{side} at the centre, {spp} spp"}
    return res


def traversal_counts(sample, nodes, tris):
    """Node visits and triangle tests per ray of the oracle's FP64 walk of the
    product BVH over the sampled rays, beside the GPU's counts of the same
    rays (which include the speculative traversal's extra visits), and the
    agreement of the closest hits."""
    import numpy as np
    from tests import oracle_lib
    rec = np.concatenate([sample["o"], sample["tmax"][:, None], sample["d"],
                          sample["any"].view(np.float32)[:, None], np.zeros((len(sample["o"]), 4), np.float32)], 1)
    nv, nt, hit, _ = oracle_lib.bvh4_walk(nodes, tris, rec)
    out = {}
    for name, sel in (("extend", sample["any"] == 0), ("shadow_any", sample["any"] == 1)):
        if not sel.any():
            continue
        d = {"rays": int(sel.sum()), "nodes_per_ray": round(float(nv[sel].mean()), 3),
             "tris_per_ray": round(float(nt[sel].mean()), 3),
             "gpu_nodes_per_ray": round(float(sample["nodes"][sel].mean()), 3),
             "gpu_tris_per_ray": round(float(sample["tris"][sel].mean()), 3)}
        if name == "extend":
            d["hit_triangle_agreement"] = round(float((hit[sel] == sample["tri"][sel]).mean()), 6)
        else:
            d["occlusion_agreement"] = round(float(((hit[sel] >= 0) == (sample["prim"][sel] >= 0)).mean()), 6)
        out[name] = d
    return out


def valu_profile():
    """VALU busy / utilisation of the timed traversal and shade kernels from
    the committed counter summary of the default frame."""
    path = os.path.join(ROOT, "profiles", "valu_icosphere1m_1080p.txt")
    if not os.path.exists(path):
        return None
    rows = {}
    with open(path) as f:
        for line in f:
            if line.startswith("#") and "|" in line and "k_" in line:
                cols = [c.strip() for c in line[1:].split("|")]
                try:
                    rows[cols[0]] = {"busy": float(cols[4]), "utilization": float(cols[5])}
                except (ValueError, IndexError):
                    pass
    names = {"k_extend": "k_trace_queue<256, 16, 8, false, false, true>",
             "k_shadow_any": "k_trace_queue<128, 16, 4, false, true, true>", "k_shade": "k_shade<512, false, 2>"}
    out = {k: rows[v] for k, v in names.items() if v in rows}
    return out or None


def batched_leg(n_jobs=64, autotune=True):
    """64 mixed jobs: the plain per-job kernel (one launch per job per
    sample) against the packed persistent wavefront launch."""
    from dayflower_amd import abi, workloads as W
    from dayflower_amd.renderer import GPURenderer
    t0 = time.perf_counter()
    scenes = W.batched_scenes()
    r = GPURenderer()
    ids = [r.addScene(sc) for sc in scenes]
    log(f"batched: 4 scenes resident in {time.perf_counter() - t0:.1f} s")
    jobs = W.batched_jobs(ids, n_jobs=n_jobs)
    total = sum(j.width * j.height * j.spp for j in jobs)
    t0 = time.perf_counter()
    tuned = r.autotune(jobs) if autotune else tuple(r.getWavefront().values())
    log(f"batched: autotune -> block {tuned[0]}, pool {tuned[1]}, sort keys {tuned[2]} ({time.perf_counter() - t0:.1f} s)")
    res = {}
    for name, mode in (("plain", abi.BATCH_PLAIN), ("packed", abi.BATCH_PACKED)):
        r.batchRender(jobs[:2], mode=mode, readback=False)  # warm
        t0 = time.perf_counter()
        r.batchRender(jobs, mode=mode, readback=False)
        wall = time.perf_counter() - t0
        st = r.stats()
        res[name] = {"seconds": round(wall, 3), "samples_per_s": round(total / wall, 1),
                     "mrays_per_s": round((st["rays_closest"] + st["rays_any"]) / wall / 1e6, 1)}
        log(f"batched: {name} {wall:.2f} s, {total / wall / 1e6:.1f} M samples/s")
    r.dispose()
    return {"workload": f"{len(jobs)} mixed render jobs in one batch (res 128^2..1024^2, spp 1..256, "
                        "icospheres 10k..2M triangles, Matte/Glass/Metal/Disney, maxBounce 8)",
            "path_samples": total, "autotuned": {"block": tuned[0], "pool_paths": tuned[1], "sort_keys": tuned[2]},
            "plain": res["plain"], "packed": res["packed"],
            "speedup_packed_over_plain": round(res["plain"]["seconds"] / res["packed"]["seconds"], 3)}


def run_batched(args):
    out = batched_leg(args.jobs)
    print(json.dumps({"metric": METRIC, "config": {k: out[k] for k in ("workload", "path_samples", "autotuned")},
                      "plain": out["plain"], "packed": out["packed"],
                      "speedup_packed_over_plain": out["speedup_packed_over_plain"],
                      "unit": "path samples/s", "n_gpus": 1}), flush=True)


def run_serving(args):
    """Poisson request stream against a resident 2M-triangle scene: the
    continuous-batching server on the whole stream, and a FIFO of plain
    per-request launches on the stream's first --plain-prefix seconds."""
    import threading
    import numpy as np
    from dayflower_amd import abi, workloads as W
    from dayflower_amd.renderer import GPURenderer, Server
    sc = W.mixed_material_icosphere(2_000_000, 0x5EED)
    r = GPURenderer()
    sid = r.addScene(sc)
    stream = W.request_stream(sid, rate=args.rate, duration=args.duration)
    total = sum(j.width * j.height * j.spp for _, j in stream)
    log(f"serving: {len(stream)} requests over {args.duration} s, {total / 1e9:.2f} G samples")
    out = {}
    # continuous batching server
    pool = args.pool or (1 << 24)
    srv = Server(r, hbm_budget_bytes=64 << 30, max_active=256, pool_paths=pool)
    t0 = time.perf_counter()
    ids = []
    last_log = t0
    for at, job in stream:
        now = time.perf_counter() - t0
        if at > now:
            time.sleep(at - now)
        ids.append(srv.submit(job))
        if time.perf_counter() - last_log > 20:
            last_log = time.perf_counter()
            log(f"serving: {len(ids)}/{len(stream)} submitted, {srv.stats()['completed']} done at {last_log - t0:.0f} s")
    for k, i in enumerate(ids):
        srv.wait(i)
        if time.perf_counter() - last_log > 20:
            last_log = time.perf_counter()
            log(f"serving: waited for {k + 1}/{len(ids)}, {srv.stats()['completed']} done at {last_log - t0:.0f} s")
    wall = time.perf_counter() - t0
    lat = np.array([srv.latency(i) for i in ids])
    st = srv.stats()
    srv.close()
    out["served"] = {"p50_s": round(float(np.percentile(lat, 50)), 4), "p99_s": round(float(np.percentile(lat, 99)), 4),
                     "samples_per_s": round(total / wall, 1), "wall_s": round(wall, 2), "peak_active": st["peak_active"],
                     "peak_bytes": st["peak_bytes"], "busy_s": round(st["busy_seconds"], 2)}
    log(f"serving: server p50 {out['served']['p50_s']} s p99 {out['served']['p99_s']} s, {total / wall / 1e6:.1f} M samples/s")
    # FIFO of plain per-request launches (one job at a time) on a prefix
    prefix = [(at, j) for at, j in stream if at < args.plain_prefix]
    ptotal = sum(j.width * j.height * j.spp for _, j in prefix)
    # the server's figures on the same prefix, for a like-for-like ratio
    lat_prefix = lat[:len(prefix)]
    done = {}
    t0 = time.perf_counter()
    submit_t = {}

    def feeder():
        for k, (at, _) in enumerate(prefix):
            now = time.perf_counter() - t0
            if at > now:
                time.sleep(at - now)
            submit_t[k] = time.perf_counter()

    th = threading.Thread(target=feeder)
    th.start()
    for k in range(len(prefix)):
        while k not in submit_t:
            time.sleep(0.0005)
        r.batchRender([prefix[k][1]], mode=abi.BATCH_PLAIN, readback=True)
        done[k] = time.perf_counter() - submit_t[k]
        if time.perf_counter() - last_log > 20:
            last_log = time.perf_counter()
            log(f"serving: plain {k + 1}/{len(prefix)} requests done")
    th.join()
    pwall = time.perf_counter() - t0
    plat = np.array([done[k] for k in range(len(prefix))])
    out["plain_fifo"] = {"p50_s": round(float(np.percentile(plat, 50)), 4), "p99_s": round(float(np.percentile(plat, 99)), 4),
                         "samples_per_s": round(ptotal / pwall, 1), "wall_s": round(pwall, 2),
                         "requests": len(prefix), "stream_prefix_s": args.plain_prefix}
    log(f"serving: plain p50 {out['plain_fifo']['p50_s']} s p99 {out['plain_fifo']['p99_s']} s, {ptotal / pwall / 1e6:.1f} M samples/s")
    r.dispose()
    print(json.dumps({"metric": METRIC, "config": {"workload": f"Poisson {args.rate} req/s for {args.duration} s, "
                      "2M-triangle displaced icosphere, res 256^2..1024^2, spp 4..64, 64 GB admission budget",
                      "requests": len(stream), "path_samples": total, "pool_paths": pool},
                      "served": out["served"], "plain_fifo": out["plain_fifo"],
                      "served_prefix_latency": {"p50_s": round(float(np.percentile(lat_prefix, 50)), 4),
                                                "p99_s": round(float(np.percentile(lat_prefix, 99)), 4)},
                      "throughput_ratio": round(out["served"]["samples_per_s"] / out["plain_fifo"]["samples_per_s"], 3),
                      "unit": "path samples/s", "n_gpus": 1}), flush=True)


def run_soup_ao(args):
    """BVH stress: 5M-triangle soup, 2048 x 2048 primary rays plus 16
    ambient-occlusion rays of radius 0.05 per hit, per step."""
    from dayflower_amd import abi, scene as S
    from dayflower_amd.renderer import GPURenderer
    res = 2048
    t0 = time.perf_counter()
    sc = S.soup_scene(5_000_000, seed=7, width=res, height=res)
    log(f"soup_ao: scene generated in {time.perf_counter() - t0:.1f} s")
    r = GPURenderer()
    r.setScene(sc)
    r.setSeed(8)
    r.setRenderingAlgorithm(abi.ALGO_AMBIENT_OCCLUSION)
    r.setSamples(16)
    r.setMaximumDistance(0.05)
    r.setup()
    for _ in range(args.warmup):
        r.render(1)
    closest = anyhit = 0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        r.render(1)
        st = r.stats()
        closest += st["rays_closest"]
        anyhit += st["rays_any"]
    wall = time.perf_counter() - t0
    r.dispose()
    print(json.dumps({"metric": METRIC, "config": {"workload": "5M-triangle soup, 2048x2048 primary rays + 16 AO rays "
                      "(radius 0.05) per hit per step", "triangles": sc.n_triangles, "steps": args.steps},
                      "mrays_per_s": round((closest + anyhit) / wall / 1e6, 1),
                      "mrays_closest_per_s": round(closest / wall / 1e6, 1), "mrays_any_per_s": round(anyhit / wall / 1e6, 1),
                      "ms_per_step": round(wall / args.steps * 1e3, 2), "unit": "Mrays/s", "n_gpus": 1}), flush=True)


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
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-job batched leg of the default line")
    ap.add_argument("--no-perez", action="store_true")
    ap.add_argument("--scene", choices=["icosphere", "mixed"], default="icosphere",
                    help="icosphere: Plastic mesh (default frame); mixed: the Matte/Glass/Metal/Disney icosphere "
                         "of the batched and serving workloads")
    ap.add_argument("--config", choices=["frame", "batched", "serving", "soup_ao"], default="frame")
    ap.add_argument("--jobs", type=int, default=64)
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--duration", type=float, default=2.0)
    ap.add_argument("--plain-prefix", type=float, default=2.0,
                    help="serving: the plain FIFO leg serves the requests arriving in the first this many seconds")
    ap.add_argument("--pool", type=int, default=0, help="path pool capacity (0: library default)")
    ap.add_argument("--sort-keys", type=int, default=-1, help="0 (off) or 1 (on) (default: library)")
    ap.add_argument("--sample-stride", type=int, default=256,
                    help="every n-th traced ray of the counting pass is recounted by the oracle")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="fill and copy MIB MiB on the device first (reference kernels for HBM counters)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus must be 1 (multi-GPU sharding is a design "
                                   "note only)", "n_gpus": args.gpus}))
        sys.exit(2)

    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
    if args.config == "batched":
        return run_batched(args)
    if args.config == "serving":
        return run_serving(args)
    if args.config == "soup_ao":
        return run_soup_ao(args)

    from dayflower_amd import scene as S
    from dayflower_amd.renderer import GPURenderer

    t0 = time.perf_counter()
    if args.scene == "mixed":
        from dayflower_amd import workloads as W
        sc = W.mixed_material_icosphere(args.triangles, 0x5EED, with_perez=not args.no_perez, width=args.width,
                                        height=args.height)
    else:
        sc = S.icosphere_scene(args.triangles, seed=0x5EED, width=args.width, height=args.height,
                               with_perez=not args.no_perez)
    log(f"bench: scene {sc.n_triangles} triangles generated in {time.perf_counter() - t0:.1f} s")
    seed = 0x5EED
    r = GPURenderer()
    r.setScene(sc)
    r.setSeed(seed)
    r.setMaximumBounce(args.max_bounce)
    r.setMinimumBounceRussianRoulette(5)
    r.setKernel(0 if args.kernel == "wavefront" else 1)
    r.setWavefront(pool_paths=args.pool or None, sort_keys=None if args.sort_keys < 0 else args.sort_keys)
    wf = r.getWavefront()
    bvh_width = r.getBvhWidth()
    t0 = time.perf_counter()
    r.setup()
    log(f"bench: setup (BVH build + upload) {time.perf_counter() - t0:.1f} s")

    # untimed counting pass: visit counts of the GPU traversal, and every
    # n-th traced ray recorded for the oracle's recount
    sample = bvh = None
    cst = None
    if args.kernel == "wavefront":
        r.setCountTraversal(True, sample_stride=args.sample_stride)
        r.render(max(1, args.spp // 8))
        cst = r.stats()
        if not args.no_cpu:
            sample = r.ray_sample()
            bvh = r.bvh()
        r.setCountTraversal(False)
        log(f"bench: GPU visits per ray: extend {cst['extend_nodes'] / max(1, cst['extend_rays']):.2f} nodes, "
            f"{cst['extend_tris'] / max(1, cst['extend_rays']):.2f} triangles; shadow "
            f"{cst['shadow_nodes'] / max(1, cst['shadow_rays']):.2f} nodes, "
            f"{cst['shadow_tris'] / max(1, cst['shadow_rays']):.2f} triangles"
            + (f"; {len(sample['o'])} rays recorded" if sample is not None else ""))

    for i in range(args.warmup):
        r.render(args.spp)
        log(f"bench: warmup {i + 1}/{args.warmup} {r.stats()['seconds']:.3f} s")
    r.clear()
    r.setProfile(args.kernel == "wavefront")
    tot = {"samples": 0, "closest": 0, "any": 0, "ext_s": 0.0, "ext_launch": 0, "ext_rays": 0, "dev_s": 0.0,
           "sh_s": 0.0, "sh_rays": 0, "shade_s": 0.0}
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
        tot["sh_s"] += st["shadow_seconds"]
        tot["sh_rays"] += st["shadow_rays"]
        tot["shade_s"] += st["shade_seconds"]
        tot["dev_s"] += st["seconds"]
        log(f"bench: step {i + 1}/{args.steps} {st['seconds']:.3f} s")
    wall = time.perf_counter() - t0
    r.setProfile(False)
    # untimed: one frame of the first spp samples, for the CPU comparison of the same streams
    gpu_film = None
    if not args.no_cpu:
        r.clear()
        r.render(args.spp)
        gpu_film, _ = r.film(args.width, args.height)
    r.dispose()
    value = tot["samples"] / wall

    cpu = counts = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(sc, args.width, args.height, args.spp, args.max_bounce, 5, seed, args.cpu_seconds,
                               gpu_film)
            if sample is not None and len(sample["o"]):
                counts = traversal_counts(sample, *bvh)
        except Exception as e:  # the baseline is informative; never hide the GPU number
            cpu = {"error": str(e)}
        log(f"bench: cpu baseline {json.dumps(cpu)}")
        log(f"bench: traversal counts {json.dumps(counts)}")

    roofline = stages = None
    if tot["ext_s"] > 0:
        launches = max(1, tot["ext_launch"])
        gpu_cnt = {"k_extend": (cst["extend_nodes"] / max(1, cst["extend_rays"]), cst["extend_tris"] / max(1, cst["extend_rays"])),
                   "k_shadow": (cst["shadow_nodes"] / max(1, cst["shadow_rays"]), cst["shadow_tris"] / max(1, cst["shadow_rays"]))}
        orc = {}
        if counts:
            if "extend" in counts:
                orc["k_extend"] = (counts["extend"]["nodes_per_ray"], counts["extend"]["tris_per_ray"])
            if "shadow_any" in counts:
                orc["k_shadow"] = (counts["shadow_any"]["nodes_per_ray"], counts["shadow_any"]["tris_per_ray"])

        # Algorithmic bytes per traversal query: ray in (32) + hit out (16) +
        # 32 per node visit + 36 of vertex positions per triangle test, with
        # the visit counts of the oracle's FP64 walk of the same BVH over the
        # sampled rays (the GPU's own counts when no CPU leg ran).
        def bytes_per(nodes, tris):
            return 32 + 16 + 32 * nodes + 36 * tris

        kern = {}
        for name, secs, rays, what in (("k_extend", tot["ext_s"], tot["ext_rays"], "k_trace_queue on the extension rays: BVH closest hit"),
                                       ("k_shadow", tot["sh_s"], tot["sh_rays"], "k_trace_queue on the any-hit shadow queue + k_shadow_closest")):
            n_node, n_tri = orc.get(name, gpu_cnt[name])
            bpr = bytes_per(n_node, n_tri)
            kern[name] = {"seconds": secs, "rays": rays, "nodes_per_ray": n_node, "tris_per_ray": n_tri,
                          "bytes_per_ray": bpr, "achieved": bpr * rays / max(secs, 1e-12) / 1e9, "what": what,
                          "gpu_nodes_per_ray": gpu_cnt[name][0], "gpu_tris_per_ray": gpu_cnt[name][1],
                          # the four-wide node record read per visit (six planes x four children + four references)
                          "node_record_GBs": (48 + 112 * gpu_cnt[name][0] + 36 * gpu_cnt[name][1]) * rays / max(secs, 1e-12) / 1e9}
        stages = {name: {"ms_per_launch": round(k["seconds"] / launches * 1e3, 3), "rays": k["rays"],
                         "nodes_per_ray": round(k["nodes_per_ray"], 2), "tris_per_ray": round(k["tris_per_ray"], 2),
                         "gpu_nodes_per_ray": round(k["gpu_nodes_per_ray"], 2), "gpu_tris_per_ray": round(k["gpu_tris_per_ray"], 2),
                         "bytes_per_ray": round(k["bytes_per_ray"], 1), "achieved_GBs": round(k["achieved"], 1),
                         "node_record_GBs": round(k["node_record_GBs"], 1)}
                  for name, k in kern.items()}
        stages["k_shade"] = {"ms_per_launch": round(tot["shade_s"] / launches * 1e3, 3)}
        dom = max(kern, key=lambda n: kern[n]["seconds"])
        k = kern[dom]
        traffic = hbm_traffic({"k_extend": "k_extend", "k_shadow": "k_shadow_any"}[dom], args)
        avg_s = k["seconds"] / launches
        valu = valu_profile()
        roofline = {"bound": "hbm", "achieved": round(k["achieved"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                    "frac": round(k["achieved"] / HBM_PEAK_GBS, 4), "traffic": traffic,
                    "kernel": f"{dom} ({k['what']})",
                    "algorithmic_bytes_per_launch": round(k["bytes_per_ray"] * k["rays"] / launches),
                    "avg_launch_ms": round(avg_s * 1e3, 3), "bytes_per_ray": round(k["bytes_per_ray"], 1),
                    "counts_from": ("oracle FP64 walk of the same BVH over "
                                    f"{counts['extend' if dom == 'k_extend' else 'shadow_any']['rays']} sampled rays"
                                    if dom in orc else "GPU counting pass"),
                    "dram_frac": round(traffic / avg_s / 1e9 / HBM_PEAK_GBS, 4) if traffic else None,
                    "valu": (valu or {}).get({"k_extend": "k_extend", "k_shadow": "k_shadow_any"}[dom]),
                    "limiter": "VALU issue: divergent traversal loops keep under half of each wave's lanes active, "
                               "with dependent node fetches; the DRAM rate stays below a third of peak"}
    batched = None
    if not args.no_batched and args.kernel == "wavefront":
        try:
            batched = batched_leg(args.jobs)
        except Exception as e:  # informative; never hide the frame number
            batched = {"error": str(e)}
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
        "config": {"workload": f"{'mixed-material ' if args.scene == 'mixed' else ''}displaced icosphere path tracing ({sc.n_triangles} triangles), "
                               f"{args.width}x{args.height}, {args.spp} spp per step, maxBounce {args.max_bounce}, "
                               f"RR from 5, {'Perez sky + ' if not args.no_perez else ''}point light",
                   "kernel": args.kernel, "triangles": sc.n_triangles, "width": args.width, "height": args.height,
                   "spp": args.spp, "max_bounce": args.max_bounce,
                   "wavefront": wf if args.kernel == "wavefront" else None, "bvh_width": bvh_width},
        "mrays_per_s": round((tot["closest"] + tot["any"]) / wall / 1e6, 1),
        "mrays_closest_per_s": round(tot["closest"] / wall / 1e6, 1),
        "mrays_any_per_s": round(tot["any"] / wall / 1e6, 1),
        "device_seconds": round(tot["dev_s"], 3),
        "roofline": roofline,
        "stages": stages,
        "traversal_counts": counts,
        "cpu_baseline": cpu,
        "batched": batched,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
