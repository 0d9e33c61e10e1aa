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

roofline: the dominant kernel of the frame, k_walks (one launch a bounce: that bounce's closest-hit
rays, then the previous bounce's any-hit shadow rays). Its algorithmic bytes per launch are 48 B a ray (32 in, 16 out) + 64 B per 4-wide node visited (the
whole quantized node: four 16-byte loads) + 48 B per triangle tested + 64 B per instance
entered, from the counted replay's per-kind counters divided by its launches per frame; its mean
launch time is the library's own HIP events around each launch on its stream, over the timed
frames. Peak: MI355X's 8 TB/s. `traffic` is that kernel's measured HBM bytes per launch from the
committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary, and `valu_busy`, `valu_thread_util` and `l2_hit`
its vector-issue share, lane utilisation and L2 hit rate from the same summary: the walk is limited
by vector issue, its node fetches served by the L2.

cpu_baseline: the oracle (oracle/, a scalar C++ port of the same tracer over the same LBVH) on whole
frames of the same workload (same scene, seeds and sample numbering), all host threads, for about
--cpu-seconds.

parity: one whole frame of the GPU (samples from warmup x spp) against the oracle's frame of the
same samples: coverage bit for bit and the largest relative radiance error (bar 1e-4); and the error
against the FP64 oracle on a band of rows.

time_to_equal_rmse: the other half of the metric on the same scene (see cache_measure): the
configuration as stated (jittered camera rays, the sun and 64 point lights, RMSE against a 4096-spp
plain render), the same lights with pixel-centre camera rays (the scene's coherent primaries), and
pixel centres under the sun and sky.
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
NODE_BYTES = 64  # a 4-wide node visit loads the whole quantized node: 4 x 16 B
METRIC = "Mrays/s (closest+any-hit); 1080p time-to-equal-RMSE, cached vs plain tracer"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def read_profile(kernel, figure, summary="path_trace_c2_pmc.md"):
    """A figure of `kernel` from a committed rocprofv3 PMC summary (lines `figure kernel: value`):
    traffic_bytes_per_launch (calibrated FETCH_SIZE + WRITE_SIZE), valu_busy (issued VALU cycles over
    the SIMDs' cycles), valu_thread_util (SQ_THREAD_CYCLES_VALU over 64 x SQ_ACTIVE_INST_VALU),
    l2_hit (TCC hits over requests); None when the summary or the line is not there."""
    path = os.path.join(ROOT, "profiles", summary)
    if not os.path.exists(path):
        return None
    key = f"{figure} {kernel}:"
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


def frame_parity(s, pt, sample_base, gpu_acc, gpu_cov, fp64_rows=(520, 556)):
    """The GPU's frame of `sample_base` against the oracle's frame of the same samples: coverage
    bit for bit, radiance relative error per pixel (the FP32 oracle runs the shader's operation
    order), and the error against the FP64 oracle on a band of rows."""
    import numpy as np

    from oracle import binding as ob
    sc = ob.Scene(s.desc())
    p = type(pt).from_buffer_copy(pt)
    p.sample_base = sample_base
    t = time.perf_counter()
    oa, oc = sc.path_trace(p)
    t32 = time.perf_counter() - t
    rel = np.abs(gpu_acc[:, :3] - oa[:, :3]) / np.maximum(np.abs(oa[:, :3]), 1e-6)
    out = {"sample_base": sample_base, "pixels": int(len(oc)), "coverage_equal": bool(np.array_equal(gpu_cov, oc)),
           "max_rel_err": float(rel.max()), "bit_exact_pixels": float(np.mean(np.all(gpu_acc == oa, axis=1))),
           "bar": 1e-4, "oracle_seconds": round(t32, 2)}
    r0, r1 = fp64_rows
    t = time.perf_counter()
    da, _ = sc.path_trace64(p, rows=(r0, r1))
    band = slice(r0 * WIDTH, r1 * WIDTH)
    g, d = gpu_acc[band, :3].astype(np.float64), da[band, :3]
    e = np.abs(g - d) / np.maximum(np.abs(d), 1e-6)
    out["fp64"] = {"rows": [r0, r1], "rel_l2": float(np.linalg.norm(g - d) / np.linalg.norm(d)),
                   "median_rel": float(np.median(e)), "p99_rel": float(np.quantile(e, 0.99)),
                   "within_1e-4": float(np.mean(e <= 1e-4)), "oracle_seconds": round(time.perf_counter() - t, 2)}
    sc.close()
    return out


SHADOW_METRIC = "Mrays/s (any-hit sun shadow rays, whole frames with 10% of instances rebuilt)"


SHADOW_PMC = "shadow_c3_pmc.md"


def shadow_main(args):
    """The shadow pass at scale: the 50M-triangle scene (4096 x 4096 fBm heightfield, 2 x 4095^2
    triangles, plus 16,384 unique 1,024-triangle cluster groups), 3840 x 2160, one jittered sun ray
    a covered pixel. A step is one frame: move a tenth of the groups (a rotating window), rebuild
    their meshes' structures and the instance structure (GE_RT_BUILD_SUBSET), trace the visibility
    words (the camera rays' closest hits, standing in for the rasterizer's), then the shadow resolve.
    value = shadow rays / wall seconds of the whole frames; the stages are timed with HIP events."""
    import torch

    from game_engine_amd import abi, rt, synth

    W, H = 3840, 2160
    t = time.perf_counter()
    s = synth.Synth.config(3, 0, args.groups, args.terrain)
    n_inst = len(s.arrays()["instances"])
    log(f"[bench] synthesized {s.triangle_count} triangles, {n_inst} instances in {time.perf_counter() - t:.1f} s")
    ctx = rt.Context(0)
    scene = ctx.upload(s.desc())
    scene.build()
    full_build = ctx.stats()["build_ms"]
    log(f"[bench] full build {full_build:.2f} ms")
    groups = n_inst - 1
    moving = max(1, groups // 10)
    pt, rv = s.frame(W, H)
    words = torch.zeros(W * H, dtype=torch.int64, device="cuda")
    out = torch.zeros(W * H, dtype=torch.int32, device="cuda")

    host = {"move": 0.0, "update": 0.0, "build": 0.0, "visibility": 0.0, "shadow": 0.0}

    def frame(k, count=False, refit=False):
        t = time.perf_counter()
        first = 1 + (k * moving) % (groups - moving + 1)
        recs = s.move_instances(first, moving, 1000 + k, 0.5, 0.2)
        meshes = [recs[i].mesh for i in range(moving)]
        t1 = time.perf_counter()
        scene.update_instances(first, recs)
        t2 = time.perf_counter()
        if refit:
            scene.build(abi.BUILD_REFIT)  # the groups are rigid: only the instance structure changes
        else:
            scene.build(abi.BUILD_SUBSET, meshes=meshes)
        build_ms = ctx.stats()["build_ms"]
        t3 = time.perf_counter()
        scene.ray_visibility(rv, words=words.data_ptr(), device=True)
        vis_ms = ctx.stats()["kernel_ms"]
        t4 = time.perf_counter()
        p = synth.Synth.shadow_params(pt, rv, abi.SHADOW_JITTER, k)
        scene.shadow_resolve(p, words.data_ptr(), count=count, out=out.data_ptr(), device=True)
        st = ctx.stats()
        t5 = time.perf_counter()
        for key, dt in zip(host, (t1 - t, t2 - t1, t3 - t2, t4 - t3, t5 - t4)):
            host[key] += dt * 1e3
        return build_ms, vis_ms, st

    for w in range(args.warmup):
        b, v, st = frame(w)
        taken = int((out != -1).sum().item())  # the same count as the timed frames, its kernels loaded
        log(f"[bench] warmup {w}: rebuild {b:.2f} ms, visibility {v:.2f} ms, shadow {st['kernel_ms']:.2f} ms, "
            f"{taken} rays")
    torch.cuda.synchronize()
    for key in host:
        host[key] = 0.0
    stages, rays = [], 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        stages.append(frame(args.warmup + k))
        rays += int((out != -1).sum().item())  # one any-hit ray an entry that asked one
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    build_ms = [x[0] for x in stages]
    vis_ms = [x[1] for x in stages]
    shadow_ms = [x[2]["kernel_ms"] for x in stages]
    walk_ms = [x[2]["any_ms"] for x in stages]
    log(f"[bench] timed {args.steps} frames: {wall * 1e3:.1f} ms wall; rebuild {['%.2f' % x for x in build_ms]}, "
        f"visibility {['%.2f' % x for x in vis_ms]}, shadow {['%.2f' % x for x in shadow_ms]} (walk "
        f"{['%.2f' % x for x in walk_ms]})")
    log(f"[bench] wall ms a frame by call (host + device): " +
        ", ".join(f"{k} {v / args.steps:.2f}" for k, v in host.items()))
    # The same frames with the refit option instead of the rebuild (reported beside, not `value`).
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    refit_stages, refit_rays = [], 0
    for k in range(args.steps):
        refit_stages.append(frame(args.warmup + args.steps + k, refit=True))
        refit_rays += int((out != -1).sum().item())
    torch.cuda.synchronize()
    refit_wall = time.perf_counter() - t0
    log(f"[bench] refit frames: {refit_wall * 1e3:.1f} ms wall; refit {['%.3f' % x[0] for x in refit_stages]}")
    # One counted frame more for the walk's algorithmic bytes a ray.
    ctx.reset_stats()
    frame(args.warmup + args.steps, count=True)
    st = ctx.stats()
    per_ray = (48 * st["any"] + NODE_BYTES * st["any_nodes"] + 48 * st["any_triangles"] + 64 * st["any_instances"]) / st["any"]
    rays_per_frame = rays / args.steps
    launch_ms = sum(walk_ms) / len(walk_ms)
    achieved = per_ray * rays_per_frame / (launch_ms * 1e-3) / 1e9
    log(f"[bench] counted frame: {st}")
    result = {
        "metric": SHADOW_METRIC, "value": round(rays / wall / 1e6, 2), "unit": "Mrays/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"shadow pass at scale: {s.triangle_count} triangles ({args.terrain}^2 fBm heightfield "
                               f"+ {groups} unique 1,024-triangle cluster groups), {W}x{H}, one jittered sun ray a "
                               f"covered pixel, {moving} groups moved and rebuilt a frame",
                   "triangles": s.triangle_count, "width": W, "height": H, "moved_per_frame": moving,
                   "parallelism": "single GPU"},
        "shadow_rays_per_step": int(rays_per_frame),
        "shadow_pass_mrays_per_s": round(rays_per_frame / (sum(shadow_ms) / len(shadow_ms) * 1e-3) / 1e6, 1),
        "refit_option": {"build_ms": round(sum(x[0] for x in refit_stages) / len(refit_stages), 3),
                         "ms_per_step": round(refit_wall / args.steps * 1e3, 3),
                         "mrays_per_s": round(refit_rays / refit_wall / 1e6, 2),
                         "note": "GE_RT_BUILD_REFIT: the moved groups are rigid, so only the instance structure is refitted"},
        "stage_ms": {"rebuild": round(sum(build_ms) / len(build_ms), 3),
                     "visibility": round(sum(vis_ms) / len(vis_ms), 3),
                     "shadow": round(sum(shadow_ms) / len(shadow_ms), 3), "full_build": round(full_build, 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4),
                     "traffic": read_profile("k_sun_walk", "traffic_bytes_per_launch", SHADOW_PMC),
                     "kernel": "k_sun_walk", "launch_ms": round(launch_ms, 4), "bytes_per_ray": round(per_ray, 1),
                     "limiter": "valu issue", "valu_busy": read_profile("k_sun_walk", "valu_busy", SHADOW_PMC),
                     "valu_thread_util": read_profile("k_sun_walk", "valu_thread_util", SHADOW_PMC),
                     "l2_hit": read_profile("k_sun_walk", "l2_hit", SHADOW_PMC)},
    }
    print(json.dumps(result), flush=True)


MASKED_METRIC = "Mrays/s (any-hit sun shadow rays through alpha-tested casters, level-0 masks)"
# Algorithmic bytes a mask evaluation reads: the cluster's material index (4), the material's flags,
# albedo texture, sampler, UV transform and offset (44), the cluster's offsets (8), the packed
# triangle (4), the three corners' half2 UVs (12), the texture and sampler records (24 + 16) and four
# RGBA8 texels (16).
MASK_TEST_BYTES = 128


def masked_main(args):
    """Alpha-tested shadows: the forest of `--trees` trees (8 meshes of an 8-sided trunk and 320 to
    480 alpha-tested leaf cards, on rolling ground; every texture slot, format, filter and address
    mode at level 0), 11520 x 2160 (the reference's forest frame), one sun ray a covered pixel. A step
    is one frame: the visibility words (the camera rays decide masked candidates by their masks,
    standing in for the rasterizer) and the sun shadow pass with GE_SHADOW_ALPHA_MASKS. value =
    shadow rays / wall seconds of the whole frames. The same pass with every caster solid is timed
    beside it on the same words (the reference's 20.6 ms masked against 3.2 ms solid)."""
    import torch

    from game_engine_amd import abi, rt, synth

    W, H = args.width or 11520, args.height or 2160
    t = time.perf_counter()
    s = synth.Synth.config(5, 0, args.trees)
    d = s.desc()
    import numpy as np
    arr = s.arrays()
    cl_tris = arr["clusters"]["triangle_count"].astype(np.int64)
    mesh_tris = np.array([cl_tris[m["first_cluster"]:m["first_cluster"] + m["cluster_count"]].sum()
                          for m in arr["meshes"]])
    instanced = int(mesh_tris[arr["instances"]["mesh"]].sum())
    log(f"[bench] forest of {args.trees} trees: {instanced} triangles instanced ({d.triangle_count} unique), "
        f"{d.instance_count} instances, {d.texture_count} textures in {time.perf_counter() - t:.1f} s")
    ctx = rt.Context(0)
    scene = ctx.upload(d)
    scene.build()
    log(f"[bench] full build {ctx.stats()['build_ms']:.2f} ms")
    pt, rv = s.frame(W, H)
    rv.masks[0] = 1  # a frame with alpha-tested materials
    words = torch.zeros(W * H, dtype=torch.int64, device="cuda")
    out = torch.zeros(W * H, dtype=torch.int32, device="cuda")
    masked_p = synth.Synth.shadow_params(pt, rv, abi.SHADOW_ALPHA_MASKS, 0)
    solid_p = synth.Synth.shadow_params(pt, rv, 0, 0)

    def frame(count=False):
        scene.ray_visibility(rv, words=words.data_ptr(), device=True, count=count)
        vis = ctx.stats()
        scene.shadow_resolve(masked_p, words.data_ptr(), count=count, out=out.data_ptr(), device=True)
        return vis, ctx.stats()

    for w in range(args.warmup):
        v, st = frame()
        log(f"[bench] warmup {w}: visibility {v['kernel_ms']:.2f} ms, masked shadow {st['kernel_ms']:.2f} ms")
    rays = int((out != -1).sum().item())  # one any-hit ray an entry that asked one
    torch.cuda.synchronize()
    vis_ms, shadow_ms, walk_ms = [], [], []
    t0 = time.perf_counter()
    for k in range(args.steps):
        v, st = frame()
        vis_ms.append(v["kernel_ms"])
        shadow_ms.append(st["kernel_ms"])
        walk_ms.append(st["any_ms"])
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    solid_ms = []
    for k in range(args.steps):
        scene.shadow_resolve(solid_p, words.data_ptr(), out=out.data_ptr(), device=True)
        solid_ms.append(ctx.stats()["kernel_ms"])
    blocked_solid = int(((out != -1) & ((out & 1) == 1)).sum().item())
    scene.shadow_resolve(masked_p, words.data_ptr(), out=out.data_ptr(), device=True)
    ctx.synchronize()  # device-pointer calls return with their kernels queued on the library's stream
    blocked = int(((out != -1) & ((out & 1) == 1)).sum().item())
    log(f"[bench] timed {args.steps} frames: {wall * 1e3:.1f} ms wall; visibility {['%.2f' % x for x in vis_ms]}, "
        f"masked shadow {['%.2f' % x for x in shadow_ms]} (walk {['%.2f' % x for x in walk_ms]}), solid shadow "
        f"{['%.2f' % x for x in solid_ms]}; {rays} rays, {blocked} blocked masked, {blocked_solid} solid")
    ctx.reset_stats()
    frame(count=True)
    st = ctx.stats()
    per_ray = (48 * st["any"] + NODE_BYTES * st["any_nodes"] + 48 * st["any_triangles"] + 72 * st["any_instances"] +
               MASK_TEST_BYTES * st["any_mask_tests"]) / st["any"]
    launch_ms = sum(walk_ms) / len(walk_ms)
    achieved = per_ray * rays / (launch_ms * 1e-3) / 1e9
    log(f"[bench] counted frame: {st}")
    mean = lambda x: sum(x) / len(x)
    cpu = None
    if not args.no_cpu:
        cpu = masked_cpu_baseline(s, min(args.cpu_seconds, 30.0))
    result = {
        "metric": MASKED_METRIC, "value": round(rays / (wall / args.steps) / 1e6, 2), "unit": "Mrays/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"alpha-tested forest shadows: {args.trees} trees ({instanced} triangles "
                               f"instanced, {d.instance_count} instances, leaf cards masked by a level-0 RGBA8 sRGB "
                               f"atlas), {W}x{H}, masked visibility words then one sun ray a covered pixel",
                   "triangles": instanced, "width": W, "height": H, "parallelism": "single GPU"},
        "shadow_rays_per_step": rays,
        "masked_shadow_pass_mrays_per_s": round(rays / (mean(shadow_ms) * 1e-3) / 1e6, 1),
        "solid_shadow_pass_mrays_per_s": round(rays / (mean(solid_ms) * 1e-3) / 1e6, 1),
        "masked_over_solid_time": round(mean(shadow_ms) / mean(solid_ms), 3),
        "blocked_fraction": {"masked": round(blocked / max(rays, 1), 4), "solid": round(blocked_solid / max(rays, 1), 4)},
        "stage_ms": {"visibility": round(mean(vis_ms), 3), "masked_shadow": round(mean(shadow_ms), 3),
                     "solid_shadow": round(mean(solid_ms), 3)},
        "mask_tests_per_ray": round(st["any_mask_tests"] / st["any"], 3),
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None, "kernel": "k_sun_walk (masks)",
                     "launch_ms": round(launch_ms, 4), "bytes_per_ray": round(per_ray, 1)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(result), flush=True)


def masked_cpu_baseline(s, target_seconds):
    """The oracle's masked visibility and shadow pass on frames of 1152 x 216 (a hundredth of the
    pixels of the timed frame, same view and sun), all host threads, until `target_seconds` pass."""
    from game_engine_amd import abi, synth
    from oracle import binding as ob
    sc = ob.Scene(s.desc())
    threads = ob.lib().ge_oracle_max_threads()
    pt, rv = s.frame(1152, 216)
    rv.masks[0] = 1
    words = sc.ray_visibility(rv)
    p = synth.Synth.shadow_params(pt, rv, abi.SHADOW_ALPHA_MASKS, 0)
    rays, frames = 0, 0
    t = time.perf_counter()
    while True:
        out, _ = sc.shadow_resolve(p, words)
        rays += int((out != abi.SHADOW_NOT_TAKEN).sum())
        frames += 1
        dt = time.perf_counter() - t
        if dt >= target_seconds or frames >= 4096:
            break
    return {"value": round(rays / dt / 1e6, 3), "unit": "Mrays/s", "cores": threads, "kind": "port",
            "sample": f"{frames} masked sun shadow passes over the oracle's 1152x216 masked visibility words of the "
                      f"same view: {rays} rays in {dt:.2f} s on {threads} threads (oracle/ scalar C++ port)"}


CACHE_METRIC = "1080p time-to-equal-RMSE, cached vs plain tracer"
PEAK_BF16_TFLOPS = 2500.0
CACHE_FLOP_PER_QUERY = 2 * (172 * 64 + 4 * 64 * 64 + 64 * 3)  # the network's algorithmic products


def cache_main(args):
    """The radiance cache workload as its own JSON line (`--workload cache`)."""
    print(json.dumps(cache_measure(args, args.jitter, args.point_lights)), flush=True)


def cache_measure(args, jitter, point_lights, scene_parts=None):
    """The radiance cache against the plain tracer on the same 4M-triangle scene at 1920 x 1080.

    By default the camera rays go through pixel centres and the scene is lit by the sun and sky. With
    jitter, near-mirror GGX surfaces (roughness down to 0.05) catch the delta sun across a pixel's
    footprint, and point lights next to surfaces give 1/d^2 spikes: camera-vertex integrands whose
    variance is so heavy-tailed that even a 4096-spp reference does not settle (--jitter and
    --point-lights turn them on; both tracers share the camera vertex exactly either way).

    A reference of args.ref_spp samples (the plain tracer, seed 0). The plain tracer (seed 1) then
    accumulates 1, 1, 2, 4, ... spp dispatches, its RMSE against the reference taken at each power
    of two; the cached tracer (seed 2) trains online (warm-up frames, then one training update
    and one 1-spp cached frame a step) with its RMSE taken after every frame. Times are wall
    clock of the dispatches alone (the RMSE reductions run between them, untimed). For each plain
    budget, the cached tracer's time to reach that RMSE is interpolated on log-log between frames;
    value = that time over the plain tracer's at args.budget spp."""
    import math

    import numpy as np
    import torch

    from game_engine_amd import abi, rt, synth

    if scene_parts is None:
        log("[bench] synthesizing the 4M-triangle instanced scene")
        s = synth.Synth.config(2, 0, args.instances)
        ctx = rt.Context(0)
        scene = ctx.upload(s.desc())
        scene.build()
    else:
        s, ctx, scene = scene_parts
    pt0, _ = s.frame(WIDTH, HEIGHT)
    lo, hi = s.bounds()
    n = WIDTH * HEIGHT
    flags = 0 if jitter else abi.PT_PIXEL_CENTER
    lights = pt0.light_count if point_lights else 0

    def params(seed, spp, base, width=WIDTH, height=HEIGHT):
        p, _ = s.frame(width, height)
        p.seed, p.spp, p.sample_base, p.flags, p.light_count = seed, spp, base, flags, lights
        return p

    def sync():
        ctx.synchronize()
        torch.cuda.synchronize()

    # ---- the reference ----
    ref_acc = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    cov = torch.zeros(n, dtype=torch.int32, device="cuda")
    batch = 16
    t0 = time.perf_counter()
    last = t0
    for k in range(args.ref_spp // batch):
        scene.path_trace(params(0, batch, k * batch), ref_acc.data_ptr(), cov.data_ptr(), device=True)
        if time.perf_counter() - last > 10:
            sync()
            last = time.perf_counter()
            log(f"[bench] reference {(k + 1) * batch} of {args.ref_spp} spp, {last - t0:.1f} s")
    sync()
    ref_s = time.perf_counter() - t0
    ref = (ref_acc[:, :3] / ref_acc[:, 3:4]).contiguous()
    log(f"[bench] reference: {args.ref_spp} spp in {ref_s:.1f} s, mean radiance {ref.mean().item():.4f}")

    def rmse(acc):
        return math.sqrt(torch.mean((acc[:, :3] / acc[:, 3:4] - ref) ** 2).item())

    # ---- the plain tracer ----
    acc = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    plain = []  # (spp, seconds, rmse)
    done, spent = 0, 0.0
    warm = torch.zeros((n, 4), dtype=torch.float32, device="cuda")
    scene.path_trace(params(9, 1, 0), warm.data_ptr(), cov.data_ptr(), device=True)  # first-dispatch costs
    sync()
    while done < args.budget:
        step = max(1, done)
        sync()
        t = time.perf_counter()
        scene.path_trace(params(1, step, done), acc.data_ptr(), cov.data_ptr(), device=True)
        sync()
        spent += time.perf_counter() - t
        done += step
        plain.append((done, spent, rmse(acc)))
        log(f"[bench] plain {done} spp: {spent * 1e3:.1f} ms, RMSE {plain[-1][2]:.5f}")

    # ---- the cached tracer ----
    # Warm-up: cached frames at the training resolution whose images are dropped, every path
    # training; then full frames, 1 spp each, every args.train_every-th one training in-frame.
    cache = rt.RadianceCache(ctx, lo, hi, seed=0)
    warm_p = params(3, 1, 0, args.train_width, args.train_height)
    warm_acc = torch.zeros((args.train_width * args.train_height, 4), dtype=torch.float32, device="cuda")
    warm_cov = torch.zeros(args.train_width * args.train_height, dtype=torch.int32, device="cuda")
    tp_warm = abi.CacheTrainParams(samples=args.samples, steps=args.train_steps, learning_rate=args.lr, stride=1)
    tp = abi.CacheTrainParams(samples=args.samples, steps=args.train_steps, learning_rate=args.lr, stride=0)
    # First-dispatch costs (the cached tracer's buffers at full size) outside the timed region.
    cache.path_trace(scene, params(9, 1, 0), depth=1, accum=warm.data_ptr(), coverage=cov.data_ptr(), device=True,
                     train=tp)
    cache.set_weights(cache.weights())  # the same initial weights, the optimizer restarted
    acc.zero_()
    cached = []  # (frame, seconds, rmse)
    losses, frame_ms = [], []
    sync()
    t = time.perf_counter()
    for w in range(args.cache_warmup):
        warm_p.sample_base = 100000 + w
        _, _, loss = cache.path_trace(scene, warm_p, depth=1, accum=warm_acc.data_ptr(),
                                      coverage=warm_cov.data_ptr(), device=True, train=tp_warm, want_loss=True)
        losses.append(loss)
    sync()
    spent = time.perf_counter() - t
    warm_s = spent
    log(f"[bench] cache warm-up: {args.cache_warmup} frames at {args.train_width}x{args.train_height} in "
        f"{spent * 1e3:.1f} ms, loss {losses[0]:.3f} -> {losses[-1]:.3f}" if losses else "[bench] no warm-up")
    target_rmse = min(e for _, _, e in plain)
    for f in range(args.cache_frames):
        trains = args.train_every > 0 and f % args.train_every == 0
        sync()
        t = time.perf_counter()
        out = cache.path_trace(scene, params(2, 1, f), depth=1, accum=acc.data_ptr(), coverage=cov.data_ptr(),
                               device=True, train=tp if trains else None, want_loss=trains)
        sync()
        dt = time.perf_counter() - t
        frame_ms.append((dt * 1e3, trains))
        spent += dt
        if trains:
            losses.append(out[2])
        cached.append((f + 1, spent, rmse(acc)))
        if f < 4 or (f + 1) % 16 == 0:
            log(f"[bench] cached frame {f + 1}: {spent * 1e3:.1f} ms, RMSE {cached[-1][2]:.5f}, last frame "
                f"{dt * 1e3:.2f} ms{' (trained)' if trains else ''}, loss {losses[-1]:.3f}")
        if cached[-1][2] <= target_rmse and f >= 8:
            break
    train_frames = [m for m, tr in frame_ms if tr]
    plain_frames = [m for m, tr in frame_ms if not tr]

    def time_to(target):
        prev = None
        for fr, t_s, e in cached:
            if e <= target:
                if prev is None or prev[2] <= e:
                    return t_s
                lt0, lt1 = math.log(prev[1]), math.log(t_s)
                le0, le1 = math.log(prev[2]), math.log(e)
                return math.exp(lt0 + (math.log(target) - le0) * (lt1 - lt0) / (le1 - le0))
            prev = (fr, t_s, e)
        return None

    table = []
    for spp, t_s, e in plain:
        tc = time_to(e)
        table.append({"plain_spp": spp, "plain_ms": round(t_s * 1e3, 2), "rmse": round(e, 6),
                      "cached_ms": None if tc is None else round(tc * 1e3, 2),
                      "ratio": None if tc is None else round(tc / t_s, 4)})
        log(f"[bench] plain {spp} spp RMSE {e:.5f} in {t_s * 1e3:.1f} ms; cached "
            + ("never" if tc is None else f"{tc * 1e3:.1f} ms ({tc / t_s:.3f}x)"))
    head = next((row for row in table if row["plain_spp"] == args.budget), table[-1])
    frames = len(cached)

    # ---- the network's inference kernel alone, at one frame's query count ----
    nq = int(0.5 * n)
    queries = torch.rand((nq, 12), dtype=torch.float32, device="cuda")
    out = torch.zeros((nq, 3), dtype=torch.float32, device="cuda")
    infer_ms = []
    for _ in range(5):
        cache.infer(queries.data_ptr(), out.data_ptr(), device=True, n=nq)
        infer_ms.append(ctx.stats()["kernel_ms"])
    launch_ms = sorted(infer_ms)[len(infer_ms) // 2]
    achieved = nq * CACHE_FLOP_PER_QUERY / (launch_ms * 1e-3) / 1e12
    log(f"[bench] k_cache_infer: {nq} queries in {['%.3f' % x for x in infer_ms]} ms, {achieved:.1f} TFLOP/s")
    value = head["ratio"]
    result = {
        "metric": CACHE_METRIC, "value": value, "unit": "x plain tracer's time", "n_gpus": 1,
        "steps": frames, "warmup": args.cache_warmup,
        "ms_per_step": round(sum(m for m, _ in frame_ms) / max(frames, 1), 3), "higher_is_better": False,
        "scaling": "weak", "vs_baseline": None, "dtype": "bf16 network, f32 tracer", "data": "synthetic",
        "config": {"workload": f"radiance cache on the 4M-triangle instanced scene ({args.instances} instances of "
                               f"16 x 1,024-triangle meshes), {WIDTH}x{HEIGHT}, 1 spp a frame ending into the cache "
                               f"after bounce 1, trained in-frame (every {args.train_every} frames, at most "
                               f"{args.samples} self-training samples, {args.train_steps} steps) after "
                               f"{args.cache_warmup} warm-up frames at {args.train_width}x{args.train_height}, "
                               f"RMSE against a {args.ref_spp}-spp plain render "
                               f"({'jittered' if jitter else 'pixel centres'}, "
                               f"{'sun + 64 point lights' if point_lights else 'sun and sky'}, 8 bounces)",
                   "width": WIDTH, "height": HEIGHT, "reference_spp": args.ref_spp, "plain_budget_spp": args.budget,
                   "warmup_paths": args.train_width * args.train_height, "samples": args.samples,
                   "train_every": args.train_every,
                   "train_steps": args.train_steps, "learning_rate": args.lr, "pixel_center": not jitter,
                   "point_lights": lights, "parallelism": "single GPU"},
        "plain_vs_cached": table,
        "cached_frames": frames, "cached_rmse_final": round(cached[-1][2], 6),
        "cached_warmup_ms": round(warm_s * 1e3, 2),
        "cached_ms_per_frame": {"training": round(sum(train_frames) / max(len(train_frames), 1), 3),
                                "not_training": round(sum(plain_frames) / max(len(plain_frames), 1), 3)},
        "reference_seconds": round(ref_s, 1),
        "roofline": {"bound": "mfma", "achieved": round(achieved, 1), "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_BF16_TFLOPS, 4), "traffic": None, "kernel": "k_cache_infer",
                     "launch_ms": round(launch_ms, 4), "queries_per_launch": nq,
                     "flop_per_query": CACHE_FLOP_PER_QUERY},
    }
    cache.close()
    return result


INVERSE_METRIC = "inverse-rendering iterations/s (path-replay adjoint)"


def inverse_main(args):
    """Inverse rendering on the 500k-triangle scene: 512 x 512, 16 spp, 6 bounces, 1,024 materials
    (albedo and roughness) and 9 light intensities recovered from a target rendered at hidden
    values. A step is one optimisation iteration: the forward render and the path-replay adjoint
    (through the torch.autograd.Function), the L2 loss and an Adam step. value = iterations / wall
    second over the timed steps."""
    import torch

    from game_engine_amd import abi, inverse, rt, synth

    s = synth.Synth.config(4, 0)
    ctx = rt.Context(0)
    scene = ctx.upload(s.desc()).build()
    pt, _ = s.frame(512, 512)
    assert pt.spp == 16 and pt.max_bounces == 6
    r = inverse.InverseRenderer(scene, s.desc(), pt)
    albedo0, rough0, lights0, sun0 = (t.cuda() for t in r.initial())
    gen = torch.Generator(device="cpu").manual_seed(11)
    m = albedo0.shape[0]
    hidden_a = (albedo0 * (0.6 + 0.8 * torch.rand((m, 3), generator=gen)).cuda()).clamp(0.05, 0.95)
    hidden_r = (rough0 * (0.6 + 0.8 * torch.rand(m, generator=gen)).cuda()).clamp(0.05, 1.0)
    with torch.no_grad():
        target = inverse.render(hidden_a, hidden_r, lights0, sun0, r).double()
    albedo = albedo0.clone().requires_grad_(True)
    rough = rough0.clone().requires_grad_(True)
    lights = lights0.clone().requires_grad_(True)
    opt = torch.optim.Adam([albedo, rough, lights], lr=0.01)
    fwd_ms, adj_ms, losses, launches = [], [], [], []

    def step():
        opt.zero_grad()
        img = inverse.render(albedo, rough, lights, sun0, r)
        st_f = ctx.stats()
        fwd_ms.append(st_f["kernel_ms"])
        loss = ((img.double() - target) ** 2).mean()
        loss.backward()
        st_a = ctx.stats()
        adj_ms.append(st_a["kernel_ms"])
        launches.append((st_f["closest_launches"], st_a["closest_launches"]))
        opt.step()
        with torch.no_grad():
            albedo.clamp_(0.0, 1.0)
            rough.clamp_(0.05, 1.0)
            lights.clamp_(min=0.0)
        losses.append(float(loss.detach()))

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    fwd_ms.clear()
    adj_ms.clear()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # Rays of one frame: the forward estimator's closest and any-hit queries (a counted plain frame
    # of the same samples); the adjoint replays exactly these.
    p = abi.PathTraceParams.from_buffer_copy(r.params)
    ctx.reset_stats()
    scene.path_trace(p, count=True)
    st = ctx.stats()
    rays = st["closest"] + st["any"]
    f_ms = sum(fwd_ms) / len(fwd_ms)
    a_ms = sum(adj_ms) / len(adj_ms)
    bytes_per_frame = (48 * rays + NODE_BYTES * st["nodes"] + 48 * st["triangles"] + 64 * st["instances"])
    # The wavefront adjoint replays every ray in a primal pass (each path's radiance) and a gradient
    # pass; when the forward render traced the same parameters it reuses that as the primal. Passes
    # actually run = the adjoint's closest-hit launches over the forward's.
    passes = max(1, round(launches[-1][1] / max(launches[-1][0], 1)))
    achieved = passes * bytes_per_frame / (a_ms * 1e-3) / 1e9
    log(f"[bench] inverse: {args.steps} iterations in {wall * 1e3:.1f} ms; forward {f_ms:.2f} ms, adjoint "
        f"{a_ms:.2f} ms a frame ({passes} pass(es)); {rays} rays a frame; loss {losses[0]:.4g} -> {losses[-1]:.4g}")
    result = {
        "metric": INVERSE_METRIC, "value": round(args.steps / wall, 3), "unit": "iterations/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32 paths, f64 gradient sums",
        "data": "synthetic",
        "config": {"workload": f"inverse rendering: {s.triangle_count} triangles, {m} materials (albedo, roughness) and "
                               f"{lights0.numel()} light intensities, 512x512, 16 spp, 6 bounces, L2 loss against a "
                               f"target at hidden parameters, Adam (forward render + path-replay adjoint a step)",
                   "triangles": s.triangle_count, "materials": m, "width": 512, "height": 512, "spp": 16,
                   "max_bounces": 6, "parallelism": "single GPU"},
        "forward_ms": round(f_ms, 3), "adjoint_ms": round(a_ms, 3), "rays_per_frame": rays,
        "adjoint_mrays_per_s": round(passes * rays / (a_ms * 1e-3) / 1e6, 1), "adjoint_passes": passes,
        "loss_first": losses[0], "loss_last": losses[-1],
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBPS, "unit": "GB/s",
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": None,
                     "kernel": "the adjoint's wavefront passes (k_walks, k_shade<3>, k_resolve<true>)",
                     "launch_ms": round(a_ms, 4), "bytes_per_launch": int(passes * bytes_per_frame)},
    }
    print(json.dumps(result), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-parity", action="store_true", help="skip the whole-frame check against the oracle")
    ap.add_argument("--no-c5", action="store_true", help="skip the time-to-equal-RMSE measurement")
    ap.add_argument("--instances", type=int, default=4096)
    ap.add_argument("--megakernel", action="store_true", help="one path per lane instead of the wavefront")
    ap.add_argument("--workload", choices=["path", "shadow", "cache", "inverse", "masked"], default="path",
                    help="path: the path tracer on the 4M-triangle scene (the headline); shadow: the sun shadow "
                         "pass on the 50M-triangle scene with per-frame rebuilds; cache: time to equal RMSE, "
                         "radiance-cache tracer against the plain tracer, on the 4M-triangle scene; inverse: "
                         "inverse-rendering iterations through the path-replay adjoint; masked: the sun shadow "
                         "pass through alpha-tested casters on the textured forest")
    ap.add_argument("--trees", type=int, default=4096, help="masked workload: trees in the forest")
    ap.add_argument("--width", type=int, default=0, help="masked workload: frame width (0: 11520)")
    ap.add_argument("--height", type=int, default=0, help="masked workload: frame height (0: 2160)")
    ap.add_argument("--ref-spp", type=int, default=4096, help="cache workload: the reference's samples")
    ap.add_argument("--budget", type=int, default=64, help="cache workload: the plain tracer's samples")
    ap.add_argument("--cache-warmup", type=int, default=8,
                    help="cache workload: warm-up training frames (their time counts in the cached tracer's)")
    ap.add_argument("--cache-frames", type=int, default=512, help="cache workload: at most this many frames")
    ap.add_argument("--samples", type=int, default=1 << 16, help="cache workload: training samples a frame")
    ap.add_argument("--train-steps", type=int, default=4, help="cache workload: optimizer steps a frame")
    ap.add_argument("--train-width", type=int, default=256, help="cache workload: warm-up frame width")
    ap.add_argument("--train-height", type=int, default=144, help="cache workload: warm-up frame height")
    ap.add_argument("--lr", type=float, default=5e-3, help="cache workload: Adam's learning rate")
    ap.add_argument("--train-every", type=int, default=1, help="cache workload: frames a training frame")
    ap.add_argument("--jitter", action="store_true",
                    help="cache workload: jittered camera rays (default: pixel centres, where the reference converges)")
    ap.add_argument("--point-lights", action="store_true",
                    help="cache workload: the 64 point lights on (default: sun and sky, a finite-variance estimator)")
    ap.add_argument("--groups", type=int, default=16384, help="shadow workload: cluster groups")
    ap.add_argument("--terrain", type=int, default=4096, help="shadow workload: heightfield samples a side")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first copy 1 GiB device to device three times: the known-size kernel a rocprofv3 "
                         "FETCH_SIZE/WRITE_SIZE pass is scaled by")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU: the path tracer runs on one MI355X and has no multi-GPU path "
              f"(--gpus {args.gpus} asked)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        import torch
        src = torch.empty(1 << 30, dtype=torch.uint8, device="cuda").fill_(1)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
    if args.workload == "shadow":
        shadow_main(args)
        return
    if args.workload == "cache":
        cache_main(args)
        return
    if args.workload == "inverse":
        inverse_main(args)
        return
    if args.workload == "masked":
        masked_main(args)
        return

    import numpy as np
    import torch

    from game_engine_amd import rt, synth

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
        f"walks ms {['%.2f' % x['closest_ms'] for x in per_step]}")

    ctx.reset_stats()
    for k in range(args.steps):
        frame(args.warmup + k, count=True)
    st = ctx.stats()
    # One more frame on its own, samples warmup * spp onwards: the frame the oracle renders first.
    accum.zero_()
    coverage.zero_()
    frame(args.warmup)
    torch.cuda.synchronize()
    check_acc, check_cov = accum.cpu().numpy(), coverage.cpu().numpy().view(np.uint32)
    rays = st["closest"] + st["any"]
    log(f"[bench] counted: {st}")
    value = rays / wall / 1e6
    mean_ms = sum(kernel_ms) / len(kernel_ms)
    # Algorithmic bytes of the walks over the timed frames (counts of the counted replay of the same
    # frames: a launch walks one bounce's closest-hit rays and the previous bounce's shadow rays), and
    # their HIP-event times.
    kinds = {
        "k_walks": (48 * rays + NODE_BYTES * st["nodes"] + 48 * st["triangles"] + 64 * st["instances"],
                    sum(x["closest_ms"] for x in per_step), sum(x["closest_launches"] for x in per_step)),
    }
    if args.megakernel:
        kinds = {"k_path_trace": (48 * rays + NODE_BYTES * st["nodes"] + 48 * st["triangles"] + 64 * st["instances"],
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
                     "frac": round(achieved / PEAK_HBM_GBPS, 4), "traffic": read_profile(dominant, "traffic_bytes_per_launch"),
                     "kernel": dominant, "launch_ms": round(launch_ms, 4), "launches_per_step": launches / args.steps,
                     "bytes_per_launch": int(per_launch),
                     # What limits the walk: vector issue (VALU cycles over SIMD cycles) and divergence
                     # (lanes doing work in an issued instruction), from the committed counter passes;
                     # the L2 serves the node fetches, so HBM traffic is far below the bytes above.
                     "limiter": "valu issue", "valu_busy": read_profile(dominant, "valu_busy"),
                     "valu_thread_util": read_profile(dominant, "valu_thread_util"),
                     "l2_hit": read_profile(dominant, "l2_hit"),
                     "walk_ms_per_step": {k: round(v[1] / args.steps, 3) for k, v in kinds.items()}},
    }
    if not args.no_cpu:
        log("[bench] cpu baseline")
        cpu = cpu_baseline(s, pt, args.cpu_seconds, sample_base=args.warmup * pt.spp)
        out["cpu_baseline"] = cpu
        out["gpu_over_cpu"] = round(value / cpu["value"], 1) if cpu["value"] > 0 else None
    if not args.no_parity:
        log("[bench] parity of one whole frame against the oracle")
        out["parity"] = frame_parity(s, pt, args.warmup * pt.spp, check_acc, check_cov)
        log(f"[bench] parity: {out['parity']}")
    if not args.no_c5:
        # The time-to-equal-RMSE half of the metric, on the same scene: as configured (jittered
        # camera rays, the sun and 64 point lights), the same lights at pixel centres, and pixel
        # centres under the sun and sky.
        c5 = {}
        for name, jit, pl in (("stated", True, True), ("pixel_centres", False, True),
                              ("pixel_centres_sun_sky", False, False)):
            log(f"[bench] time to equal RMSE: {name}")
            r = cache_measure(args, jit, pl, scene_parts=(s, ctx, scene))
            head = next((row for row in r["plain_vs_cached"] if row["plain_spp"] == args.budget),
                        r["plain_vs_cached"][-1])
            c5[name] = {"ratio": r["value"], "plain_spp": head["plain_spp"], "plain_ms": head["plain_ms"],
                        "cached_ms": head["cached_ms"], "rmse": head["rmse"], "workload": r["config"]["workload"],
                        "cached_frames": r["cached_frames"], "cached_ms_per_frame": r["cached_ms_per_frame"],
                        "table": r["plain_vs_cached"], "k_cache_infer": r["roofline"]}
        out["time_to_equal_rmse"] = c5
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
