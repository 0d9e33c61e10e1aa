"""Benchmark: Mrays/s of the reference path tracer on the 4M-triangle instanced scene, one MI355X.

Workload (the ray-throughput configuration): 4,096 instances of 16 distinct 1,024-triangle meshes
(4,194,304 triangles instanced, two-level BVH: one structure per mesh and one over the instances)
under random rigid transforms, GGX roughness U[0.05, 1], one sun plus 64 radius-limited point lights,
1920 x 1080 jittered primaries, 2 spp, 8 bounces, next-event estimation with one any-hit shadow ray
per light that can reach the point. A step is one frame: one ge_rt_path_trace dispatch over the whole
image, on device-resident scene, structures, accumulator and coverage.

Rays = closest-hit queries + any-hit queries issued, counted by the kernel's own counters in a
separate counted replay of the same frames (counters are compiled out of the timed kernel).
value = rays / wall seconds of the K timed frames. The BVH build is timed separately.

roofline: the dominant kernel of the frame (the traversal walk with the larger share of device time:
k_closest, the closest-hit walk of one bounce, or k_any, the any-hit shadow walk of one bounce). Its
algorithmic bytes per launch are 48 B a ray (32 in, 16 out) + 64 B per node popped + 48 B per
triangle tested + 64 B per instance entered, from the counted replay's per-kind counters divided by
its launches per frame; its mean launch time is the library's own HIP events around each launch on
its stream, over the timed frames. Peak: MI355X's 8 TB/s. `traffic` is that kernel's measured HBM
bytes per launch from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary, when one is there.

cpu_baseline: the oracle (oracle/, a scalar C++ port of the same tracer over the same LBVH) on a
band of rows of the same frame, all host threads, same seeds.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WIDTH, HEIGHT = 1920, 1080
PEAK_HBM_GBPS = 8000.0
METRIC = "Mrays/s (closest+any-hit); 1080p time-to-equal-RMSE, cached vs plain tracer"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 PMC summary, if one is there."""
    path = os.path.join(ROOT, "profiles", "path_trace_c2_pmc.md")
    if not os.path.exists(path):
        return None
    key = f"traffic_bytes_per_launch {kernel}:"
    for line in open(path):
        if line.startswith(key):
            return float(line.split(":", 1)[1].strip())
    return None


def cpu_baseline(s, pt, target_seconds, sample_base):
    """The oracle on whole frames of the same workload (same scene, seeds and sample numbering) until
    `target_seconds` of CPU time have passed: a bounded sample, at least one frame."""
    from oracle import binding as ob
    sc = ob.Scene(s.desc())
    threads = ob.lib().ge_oracle_max_threads()
    p = type(pt).from_buffer_copy(pt)
    ob.reset_counters()
    frames = 0
    t = time.perf_counter()
    while True:
        p.sample_base = sample_base + frames * p.spp
        sc.path_trace(p)
        frames += 1
        dt = time.perf_counter() - t
        log(f"[bench] cpu frame {frames}: {dt:.1f} s")
        if dt >= target_seconds or frames >= 64:
            break
    c = ob.counters()
    rays = c["closest"] + c["any"]
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{frames} whole {WIDTH}x{HEIGHT} frames at {p.spp} spp from sample_base {sample_base}: "
                      f"{rays} rays in {dt:.2f} s on {threads} threads (oracle/ scalar C++ port, OpenMP 8x8 tiles)"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--megakernel", action="store_true", help="one path per lane instead of the wavefront")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device to device three times: the known-size kernel a rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE pass is scaled by")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the path tracer runs on one MI355X and has no multi-GPU path "
              f"(--gpus {args.gpus} asked)", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from game_engine_amd import rt, synth

    if args.copy_calibration:
        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    log("[bench] synthesizing the 4M-triangle instanced scene")
    s = synth.Synth.config(2, 0, args.instances)
    pt, _ = s.frame(WIDTH, HEIGHT)
    assert pt.spp == 2 and pt.max_bounces == 8 and pt.light_count == 64
    ctx = rt.Context(0)
    scene = ctx.upload(s.desc())
    builds = []
    for _ in range(3):
        scene.build()
        builds.append(ctx.stats()["build_ms"])
    log(f"[bench] BVH build {builds} ms (GPU events)")
    n = WIDTH * HEIGHT
    accum = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    coverage = torch.zeros(n, dtype=torch.int32, device="cuda")
    A, Cv = accum.data_ptr(), coverage.data_ptr()

    def frame(k, count=False):
        pt.sample_base = k * pt.spp
        scene.path_trace(pt, A, Cv, count=count, device=True, megakernel=args.megakernel)

    for w in range(args.warmup):
        frame(w)
        log(f"[bench] warmup {w} {ctx.stats()['kernel_ms']:.2f} ms")
    torch.cuda.synchronize()
    per_step = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        frame(args.warmup + k)
        per_step.append(ctx.stats())
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [x["kernel_ms"] for x in per_step]
    log(f"[bench] timed {args.steps} frames: {wall * 1e3:.1f} ms wall, frame ms {['%.2f' % x for x in kernel_ms]}, "
        f"closest-hit walks ms {['%.2f' % x['closest_ms'] for x in per_step]}, "
        f"any-hit walks ms {['%.2f' % x['any_ms'] for x in per_step]}")

    ctx.reset_stats()
    for k in range(args.steps):
        frame(args.warmup + k, count=True)
    st = ctx.stats()
    rays = st["closest"] + st["any"]
    log(f"[bench] counted: {st}")
    value = rays / wall / 1e6
    mean_ms = sum(kernel_ms) / len(kernel_ms)
    # Algorithmic bytes of each walk kind over the timed frames (counts of the counted replay of the
    # same frames), and their HIP-event times.
    kinds = {
        "k_closest": (48 * st["closest"] + 64 * (st["nodes"] - st["any_nodes"]) +
                      48 * (st["triangles"] - st["any_triangles"]) + 64 * (st["instances"] - st["any_instances"]),
                      sum(x["closest_ms"] for x in per_step), sum(x["closest_launches"] for x in per_step)),
        "k_any": (48 * st["any"] + 64 * st["any_nodes"] + 48 * st["any_triangles"] + 64 * st["any_instances"],
                  sum(x["any_ms"] for x in per_step), sum(x["any_launches"] for x in per_step)),
    }
    if args.megakernel:
        kinds = {"k_path_trace": (48 * rays + 64 * st["nodes"] + 48 * st["triangles"] + 64 * st["instances"],
                                  sum(kernel_ms), args.steps)}
    for name, (b, ms, launches) in kinds.items():
        log(f"[bench] {name}: {b / args.steps / 1e6:.1f} MB a frame over {launches / args.steps:.1f} launches, "
            f"{ms / args.steps:.3f} ms a frame, {b / (ms * 1e-3) / 1e9 if ms else 0:.1f} GB/s")
    dominant = max(kinds, key=lambda k: kinds[k][1])
    total_bytes, total_ms, launches = kinds[dominant]
    per_launch = total_bytes / launches
    launch_ms = total_ms / launches
    achieved = per_launch / (launch_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"4M-triangle instanced scene: {args.instances} instances of 16 x 1,024-triangle meshes, "
                               f"two-level LBVH, {WIDTH}x{HEIGHT}, 2 spp, 8 bounces, 1 sun + 64 point lights "
                               f"(path_trace_main, one frame a step)",
                   "triangles_instanced": args.instances * 1024, "width": WIDTH, "height": HEIGHT, "spp": 2,
                   "max_bounces": 8, "lights": 64, "parallelism": "single GPU"},
        "rays_per_step": rays // args.steps,
        "closest_per_step": st["closest"] // args.steps, "any_per_step": st["any"] // args.steps,
        "bvh_build_ms": round(min(builds), 3),
        "frame_kernel_ms": round(mean_ms, 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": read_traffic(dominant),
                     "kernel": dominant, "launch_ms": round(launch_ms, 4), "launches_per_step": launches / args.steps,
                     "bytes_per_launch": int(per_launch),
                     "walk_ms_per_step": {k: round(v[1] / args.steps, 3) for k, v in kinds.items()}},
    }
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(s, pt, args.cpu_seconds, sample_base=args.warmup * pt.spp)
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu"] = round(value / cpu["value"], 1) if cpu["value"] > 0 else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
