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

roofline: the path-trace kernel's algorithmic bytes per launch, 48 B a ray (32 in, 16 out) + 64 B per
node popped + 48 B per triangle tested + 64 B per instance entered, over its mean launch time
(HIP events on the library's stream), against MI355X's 8 TB/s.

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


def read_traffic():
    """HBM bytes per launch from the committed rocprofv3 PMC summary, if one is there."""
    path = os.path.join(ROOT, "profiles", "path_trace_c2_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1].strip())
    return None


def cpu_baseline(s, pt, target_seconds, sample_base):
    from oracle import binding as ob
    sc = ob.Scene(s.desc())
    threads = ob.lib().ge_oracle_max_threads()
    p = type(pt).from_buffer_copy(pt)
    p.sample_base = sample_base
    mid = HEIGHT // 2
    # Calibrate on 8 rows, then take a band sized to the target time.
    ob.reset_counters()
    t = time.perf_counter()
    sc.path_trace(p, rows=(mid - 4, mid + 4))
    dt = time.perf_counter() - t
    rows = max(8, min(HEIGHT, int(8 * target_seconds / max(dt, 1e-3))))
    r0 = max(0, mid - rows // 2)
    r1 = min(HEIGHT, r0 + rows)
    ob.reset_counters()
    t = time.perf_counter()
    sc.path_trace(p, rows=(r0, r1))
    dt = time.perf_counter() - t
    c = ob.counters()
    rays = c["closest"] + c["any"]
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"rows {r0}..{r1} of the {WIDTH}x{HEIGHT} frame, {p.spp} spp, sample_base {sample_base}: "
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
    kernel_ms = []
    t0 = time.perf_counter()
    for k in range(args.steps):
        frame(args.warmup + k)
        kernel_ms.append(ctx.stats()["kernel_ms"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    log(f"[bench] timed {args.steps} frames: {wall * 1e3:.1f} ms wall, kernel ms {['%.2f' % x for x in kernel_ms]}")

    ctx.reset_stats()
    for k in range(args.steps):
        frame(args.warmup + k, count=True)
    st = ctx.stats()
    rays = st["closest"] + st["any"]
    log(f"[bench] counted: {st}")
    value = rays / wall / 1e6
    per_launch = (48 * rays + 64 * st["nodes"] + 48 * st["triangles"] + 64 * st["instances"]) / args.steps
    mean_ms = sum(kernel_ms) / len(kernel_ms)
    achieved = per_launch / (mean_ms * 1e-3) / 1e9
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
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": read_traffic(),
                     "kernel": "k_path_trace", "kernel_ms": round(mean_ms, 3),
                     "bytes_per_launch": int(per_launch)},
    }
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(s, pt, args.cpu_seconds, sample_base=args.warmup * pt.spp)
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu"] = round(value / cpu["value"], 1) if cpu["value"] > 0 else None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
