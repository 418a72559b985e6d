#!/usr/bin/env python3
"""Benchmark of the path-tracing hot path on one MI355X.

Workload (the headline configuration): one 1920x1080 view of the icosphere
field — 64 seeded icospheres at subdivision level 5 (1,310,720 triangles)
displaced by 3-octave value noise (5 %), a 200x200 ground plane, 2,000
alpha-tested leaf quads, GGX metallic-roughness materials, a Hosek-Wilkie sky
(turbidity 3) with its sun at 30 degrees elevation plus 4 point lights;
8 bounces, 1 sample per pixel per frame. A step is one progressive frame
(2,073,600 paths). Frames are traced frames_in_flight at a time (default 256)
by one generate/extend/shade/shadow wavefront sequence replayed from a
captured hipGraph; samples are still accumulated in frame order, so the
image is bit-identical to tracing one frame per pass. Inputs are resident in
HBM before timing.

Prints one JSON line (rank 0). Extra fields:
  gsample_bounces_per_s  path segments traced (primary + extension rays) / s
  roofline               extend (closest-hit traversal) kernel: algorithmic
                         bytes per launch / measured launch time vs HBM peak
  cpu_baseline           the FP32 CPU transcription on this host's cores,
                         on a bounded number of whole frames
  time_to_relmse_0.01    wall time for plain progressive accumulation and for
                         the denoised image to reach relMSE 0.01 against the
                         plain 4096-frame image
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "Msamples/s at 1080p; time to relMSE 0.01: denoised vs plain path (1 GPU)"
HBM_PEAK_GBPS = 8000.0
NODE_BYTES = {4: 128, 8: 256}
TRI_BYTES = 48           # 3 x float4 leaf-ordered triangle
INST_BYTES = 64          # InstTrav record (world->object 3x4 + BLAS root)
RAY_QUEUE_BYTES = 32 + 4 + 16 + 4   # ray in (o, d), slot, hit record out
BUILDERS = {0: "LBVH (GPU)", 1: "binned SAH (host)", 2: "PLOC clustering (GPU)", 3: "binned SAH (GPU)"}


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=512)
    p.add_argument("--warmup", type=int, default=256)
    p.add_argument("--width", type=int, default=1920)
    p.add_argument("--height", type=int, default=1080)
    p.add_argument("--bounces", type=int, default=8)
    p.add_argument("--bvh-width", type=int, default=4)
    p.add_argument("--bvh-compressed", type=int, default=0)
    p.add_argument("--block", type=int, default=64)
    p.add_argument("--blocks-per-cu", type=int, default=32)
    p.add_argument("--frames-in-flight", type=int, default=256)
    p.add_argument("--refill-lanes", type=int, default=24)
    p.add_argument("--traversal", type=int, default=0, help="0 while-while, 1 if-if")
    p.add_argument("--bvh-builder", type=int, default=3,
                   help="0 LBVH on the GPU, 1 binned SAH on the host, 2 PLOC clustering on the GPU, "
                        "3 binned SAH on the GPU (the host builder's tree)")
    p.add_argument("--continuation-depth", type=int, default=-1,
                   help="bounces traced as wavefront launches before one continuation launch finishes every path "
                        "(0 = every bounce as wavefront launches; -1 = 4 when a pass holds fewer than 64 frames, "
                        "else 0)")
    p.add_argument("--megakernel", action="store_true")
    p.add_argument("--no-roofline", action="store_true")
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-seconds", type=float, default=15.0)
    p.add_argument("--no-denoise", action="store_true")
    p.add_argument("--ref-frames", type=int, default=4096,
                   help="frames of the plain accumulation that scores both methods")
    p.add_argument("--weights", default=str(ROOT / "helios_amd" / "data" / "denoiser_unet.bin"))
    p.add_argument("--workload", default="single", choices=["single", "batch64", "combined4k"],
                   help="single: the 1080p scene (the headline metric); batch64: 64 uneven scenes/views in shared "
                        "wavefront passes, autotuned (prints its own line)")
    p.add_argument("--denoise-probe", action="store_true",
                   help="counter-collection driver: 3 frames with feature planes, then 10 denoiser passes")
    p.add_argument("--probe", action="store_true",
                   help="counter-collection driver: 2 frames of the workload, then a 1 GiB copy kernel "
                        "(the calibration point for FETCH_SIZE / WRITE_SIZE)")
    return p.parse_args()


def cpu_baseline(desc, budget_s):
    """FP32 transcription (oracle) timed on the host's cores on a band of rows."""
    import numpy as np
    import oracle
    threads = os.cpu_count() or 1
    s = oracle.OracleScene(desc)
    w, h = desc.width, desc.height
    img = np.zeros((h, w, 4), np.float32)
    # calibrate on a band of rows through the centre, then render whole
    # progressive frames until the budget is used (at least one frame)
    t0 = time.perf_counter()
    s.render(frames=1, crop=(0, h // 2 - 32, w, h // 2 + 32), threads=threads, image=img)
    t_band = time.perf_counter() - t0
    frames = int(max(1, min(64, budget_s / max(t_band * h / 64, 1e-3))))
    t0 = time.perf_counter()
    s.render(frames=frames, threads=threads, image=img)
    t = time.perf_counter() - t0
    samples = frames * w * h
    ext = int(s.last_stats[0])
    return dict(value=samples / t / 1e6, unit="Msamples/s", cores=threads, kind="port",
                sample=f"{frames} full {w}x{h} progressive frame(s) (1 spp, 8 bounces) of the same scene and camera, "
                       f"{t:.1f} s; FP32 transcription with a CPU BVH2",
                gsample_bounces_per_s=ext / t / 1e9)


def continuation_depth(requested, chunk):
    """-1 picks the measured best for the pass size: with fewer than 64 frames
    in flight the deep-bounce launches are small and tail-bound, and finishing
    every path after 4 wavefront bounces in one continuation launch is faster
    (20 frames: 1.241 -> 1.196 ms per frame); from 64 frames the wavefront
    launches are full and continuing is not faster (256: 1.073 vs 1.081)."""
    if requested >= 0:
        return requested
    return 4 if chunk < 64 else 0


def single_frame(it, scene, desc, surf, pixels, calls=20):
    """The reference's protocol: one progressive frame per render call
    (path_integrator.cpp:48-84), on the single-frame integrator form (camera
    rays as wavefront launches, the rest of every path in one launch).
    Host clock around `calls` synchronized calls after 3 warm-up calls."""
    from helios_amd.abi import HH_FLAG_MEGAKERNEL
    from helios_amd.runtime import camera_struct
    it.set_flags(HH_FLAG_MEGAKERNEL)
    cam = camera_struct(desc.camera)   # the C struct once, as an application keeps its camera
    for _ in range(3):
        it.render(scene, cam, surf, frames=1)
    t0 = time.perf_counter()
    for _ in range(calls):
        it.render(scene, cam, surf, frames=1)   # each call ends with a stream sync
    dt = time.perf_counter() - t0
    ext = it.ray_stats()["extension_rays"]
    it.set_flags(0)
    return {"ms_per_frame": round(dt / calls * 1e3, 4), "msamples_per_s": round(pixels * calls / dt / 1e6, 1),
            "gsample_bounces_per_s": round(ext * calls / dt / 1e9, 4), "calls": calls,
            "integrator": "camera-ray wavefront + continuation megakernel, 1 frame per call"}


def roofline(ctx, it, scene, desc, surf, chunk, node_bytes, a):
    """Roofline of the closest-hit traversal kernels (k_extend_packet at depth
    0, k_extend after), on one pass of `chunk` frames: the launch shape of the
    timed run.

    Pass 1 (HIP events on the launch stream) gives the extension launches'
    time and the rays they traced. Pass 2 (the counting traversal variant,
    every bounce as per-lane wavefront launches) gives node, triangle and
    instance fetches per ray.

    achieved = algorithmic HBM bytes of one launch / its average time, where
    the algorithmic bytes are the launch's queue traffic (56 B per ray: ray
    in 32, slot 4, hit record out 16 + 4) plus one read of the whole
    acceleration structure (every BLAS/TLAS node, leaf triangle and instance
    record once): what the launch must move from HBM with ideal caching.
    traffic = HBM bytes per launch measured by FETCH_SIZE / WRITE_SIZE at the
    same shape (profiles/extend_traffic*.json), scaled by the copy-kernel
    calibration. The per-lane fetch bytes (every lane's node, triangle and
    instance fetches) are served mostly by L1/L2 and are reported beside it,
    as is the issue bound the kernel actually runs into
    (profiles/extend_issue*.json)."""
    ctx.set_profiling(timing=True)
    ctx.reset_profile()
    it.render(scene, desc.camera, surf, frames=chunk)
    prof = ctx.read_profile()
    ctx.set_profiling(timing=True, counting=True)
    ctx.reset_profile()
    it.render(scene, desc.camera, surf, frames=chunk)
    cprof = ctx.read_profile()
    ctx.set_profiling()
    cnt = cprof["counts"]
    crays = max(cprof["extend"]["extension_rays"], 1)
    csrays = max(cprof["shadow"]["shadow_rays"], 1)
    ext = prof["extend"]
    launches = max(ext["launches"], 1)
    avg_ms = ext["ms"] / launches
    rays_per_launch = ext["extension_rays"] / launches
    bs = scene.bvh_stats()
    scene_bytes = ((bs["blas_nodes"] + bs["tlas_nodes"]) * node_bytes + bs["triangles"] * TRI_BYTES +
                   len(desc.instances) * INST_BYTES)
    alg = rays_per_launch * RAY_QUEUE_BYTES + scene_bytes
    achieved = alg / (avg_ms * 1e-3) / 1e9
    per_ray = {"nodes": cnt["ext_nodes"] / crays, "tris": cnt["ext_tris"] / crays, "insts": cnt["ext_insts"] / crays}
    lane_bytes = rays_per_launch * (per_ray["nodes"] * node_bytes + per_ray["tris"] * TRI_BYTES +
                                    per_ray["insts"] * INST_BYTES + RAY_QUEUE_BYTES)
    shape = {"frames_in_flight": chunk, "continuation_depth": a.continuation_depth, "bvh_builder": a.bvh_builder}
    traffic, l2, issue = None, None, None
    # counter profiles exist per launch shape (extend_traffic*.json,
    # extend_issue*.json); the one taken at this run's shape is used
    for tfile in sorted((ROOT / "profiles").glob("extend_traffic*.json")):
        tj = json.loads(tfile.read_text())
        if tj.get("shape") == shape:
            traffic, l2 = tj.get("hbm_bytes_per_launch"), tj.get("l2")
    for ifile in sorted((ROOT / "profiles").glob("extend_issue*.json")):
        ij = json.loads(ifile.read_text())
        if ij.get("shape") == shape:
            issue = {k: ij["kernels"][k] for k in ij.get("kernels", {})}
    r = {
        "bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBPS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBPS, 5), "traffic": traffic,
        "kernel": "k_extend_packet (depth 0) + k_extend (depths 1-%d): closest-hit traversal" % (
            (a.continuation_depth or a.bounces) - 1),
        "launches_per_pass": ext["launches"], "avg_launch_ms": round(avg_ms, 4),
        "rays_per_launch": round(rays_per_launch),
        "alg_bytes_per_launch": round(alg),
        "alg_bytes": "56 B queue traffic per ray + one read of the acceleration structure (%d B)" % scene_bytes,
        "hbm_gbps": round(traffic / (avg_ms * 1e-3) / 1e9, 1) if traffic else None,
        "lane_fetch_bytes_per_launch": round(lane_bytes),
        "lane_fetch_gbps": round(lane_bytes / (avg_ms * 1e-3) / 1e9, 1),
        "l2_read_gbps": l2.get("l2_read_gbps") if l2 else None,
        "l1_hit_rate": l2.get("l1_hit_rate") if l2 else None,
        "l2_hit_rate": l2.get("l2_hit_rate") if l2 else None,
        "issue": issue,
        "per_ray": {k: round(v, 2) for k, v in per_ray.items()},
        "shadow_per_ray": {"nodes": round(cnt["shadow_nodes"] / csrays, 2),
                           "tris": round(cnt["shadow_tris"] / csrays, 2),
                           "insts": round(cnt["shadow_insts"] / csrays, 2)},
        "shape": shape,
    }
    kms = {k: round(v["ms"] / chunk, 4) for k, v in prof.items() if k != "counts" and v["launches"]}
    return r, kms


def relmse(x, r, eps=0.01):
    import numpy as np
    x = np.asarray(x[..., :3], np.float64)
    r = np.asarray(r[..., :3], np.float64)
    return float(np.mean((x - r) ** 2 / (r * r + eps)))


def time_to_relmse(ctx, scene, desc, frame_ms, weights, ref_frames, target=0.01):
    """Wall time for plain progressive accumulation and for the denoised
    image to reach relMSE <= target against the plain ref_frames image.

    One progressive run is checkpointed at ~10 % frame-count steps; at each
    checkpoint the plain image and its denoised version are scored. The
    frame counts found are then timed from scratch on the wall clock (render
    of k frames from a cleared surface, plus one denoiser pass for the
    denoised method), median of 3."""
    import numpy as np
    from helios_amd import runtime
    from helios_amd.denoise import Denoiser
    w, h = desc.width, desc.height
    it = runtime.PathIntegrator(ctx)
    it.set_max_ray_bounces(desc.max_ray_bounces)
    it.set_max_samples(1 << 30)
    ref_surf = runtime.Surface(ctx, w, h)
    t0 = time.perf_counter()
    it.render(scene, desc.camera, ref_surf, frames=ref_frames)
    ref = ref_surf.read()
    ref_s = time.perf_counter() - t0
    del ref_surf

    dn = Denoiser(ctx, seed=0)
    trained = Path(weights).exists()
    if trained:
        dn.set_params(np.fromfile(weights, np.float32))
    surf = runtime.Surface(ctx, w, h)
    surf.enable_features()
    out = runtime.Surface(ctx, w, h)
    it.restart_bake()
    # denoiser time: median of timed calls on a settled input
    it.render(scene, desc.camera, surf, frames=2)
    times = []
    for _ in range(7):
        ctx.sync()
        t0 = time.perf_counter()
        dn.denoise(surf, 2, out)
        ctx.sync()
        times.append(time.perf_counter() - t0)
    dn_ms = float(np.median(times[2:]) * 1e3)

    it.restart_bake()
    done = 0
    k = 1                      # after 1 or 2 frames the running mean holds one sample
    curve = []
    k_plain = k_dn = None
    while k <= ref_frames // 2:
        it.render(scene, desc.camera, surf, frames=k - done)
        done = k
        e_plain = relmse(surf.read(), ref)
        dn.denoise(surf, k, out)
        e_dn = relmse(out.read(), ref)
        curve.append((k, round(e_plain, 5), round(e_dn, 5)))
        if k_dn is None and e_dn <= target:
            k_dn = k
        if k_plain is None and e_plain <= target:
            k_plain = k
        if k_plain is not None and k_dn is not None:
            break
        k = max(k + 1, int(k * 1.1))
    from helios_amd.denoise import forward_bytes, forward_flops
    flops = forward_flops(h, w)
    dn_bytes = forward_bytes(h, w)
    from helios_amd.abi import HH_FLAG_MEGAKERNEL

    def wall(k, denoise, mega):
        """(median wall seconds, relMSE of the image produced) for k frames
        from a cleared surface with the wavefront or the megakernel integrator."""
        it.set_flags(HH_FLAG_MEGAKERNEL if mega else 0)
        ts = []
        for _ in range(3):
            it.restart_bake()
            ctx.sync()
            t0 = time.perf_counter()
            it.render(scene, desc.camera, surf, frames=k)
            if denoise:
                dn.denoise(surf, k, out)
            ctx.sync()
            ts.append(time.perf_counter() - t0)
        e = relmse(out.read() if denoise else surf.read(), ref)
        it.set_flags(0)
        return float(np.median(ts)), e

    def best(k, denoise):
        """The faster integrator form whose image still meets the target."""
        if not k:
            return None, None
        opts = [(wall(k, denoise, m), m) for m in (False, True)]
        ok = [(t, m) for (t, e), m in opts if e <= target] or [(opts[0][0][0], False)]
        t, m = min(ok)
        return t, ("megakernel" if m else "wavefront")

    plain_s, plain_mode = best(k_plain, False)
    dn_s, dn_mode = best(k_dn, True)
    res = {
        "target_relmse": target, "reference_frames": ref_frames, "reference_s": round(ref_s, 2),
        "denoiser_weights": "trained" if trained else "untrained (no weights file)",
        "denoise_ms": round(dn_ms, 3), "throughput_frame_ms": round(frame_ms, 4),
        "denoiser_forward": {"gflop": round(flops / 1e9, 2), "tflops": round(flops / (dn_ms * 1e-3) / 1e12, 1),
                             "frac_of_bf16_mfma_peak": round(flops / (dn_ms * 1e-3) / 2.5e15, 4),
                             "alg_bytes": dn_bytes, "flop_per_byte": round(flops / dn_bytes, 1),
                             "hbm_roofline": {"bound": "hbm (below the 312 FLOP/B ridge of 2.5 PFLOP/s over 8 TB/s)",
                                              "achieved": round(dn_bytes / (dn_ms * 1e-3) / 1e9, 1), "peak": 8000.0,
                                              "unit": "GB/s",
                                              "frac": round(dn_bytes / (dn_ms * 1e-3) / 8e12, 4)},
                             "timing": "host clock around one synchronized hh_denoise call (median of 5)"},
        "plain_frames": k_plain, "denoised_frames": k_dn,
        "plain_s": round(plain_s, 4) if plain_s else None,
        "denoised_s": round(dn_s, 4) if dn_s else None,
        "plain_integrator": plain_mode, "denoised_integrator": dn_mode,
        "timing": "wall clock from a cleared surface to the scored image (render + denoise), median of 3; "
                  "each method uses the faster integrator form (wavefront with the frames in flight of the "
                  "throughput run, or megakernel)",
        "curve_frames_plain_denoised": curve,
    }
    if k_plain and k_dn:
        res["speedup"] = round(res["plain_s"] / res["denoised_s"], 2)
    dn.close()
    return res


def bench_batch(a):
    """64 independent scenes and cameras (128^2..1024^2, 10 k..2 M triangles,
    8 bounces, 16 spp each) rendered together: autotune once, then time the
    16 samples of every view."""
    import numpy as np
    from helios_amd import runtime, scenegen
    t0 = time.perf_counter()
    descs = scenegen.batched_workload(n=64, spp=16)
    gen_s = time.perf_counter() - t0
    ctx = runtime.Context(0)
    t0 = time.perf_counter()
    scenes = [runtime.Scene(ctx, d) for d in descs]
    build_s = time.perf_counter() - t0
    surfs = [runtime.Surface(ctx, d.width, d.height) for d in descs]
    views = [dict(scene=s, surface=f, camera=d.camera, max_ray_bounces=8) for s, f, d in zip(scenes, surfs, descs)]
    # tuned on the frames the timed batch renders (16 per view): with fewer,
    # frames-in-flight settings above that count would all look alike
    tuning, tune_ms = runtime.autotune(ctx, views, frames=16, bvh_width=True)
    for v in views:
        v["spp"] = 16
    runtime.render_batch(ctx, views)          # warm-up (graphs)
    ctx.sync()
    t0 = time.perf_counter()
    runtime.render_batch(ctx, views)
    ctx.sync()
    dt = time.perf_counter() - t0
    px = sum(d.width * d.height for d in descs)
    out = {
        "metric": "Msamples/s, batched 64 views (1 GPU)", "value": round(px * 16 / dt / 1e6, 3), "unit": "Msamples/s",
        "n_gpus": 1, "steps": 1, "warmup": 1, "ms_per_step": round(dt * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic (seeded procedural scenes)",
        "config": {"workload": "64 icosphere-field scenes, seeds 0..63, sides log-uniform 128..1024, triangles "
                               "log-uniform 10k..2M, 8 bounces, 16 spp each, one batch", "pixels": px,
                   "triangles": int(sum(d.num_triangles for d in descs)), "tuning": tuning},
        "autotune_ms_per_frame": round(tune_ms, 3), "scene_gen_s": round(gen_s, 1), "scene_build_s": round(build_s, 2),
    }
    print(json.dumps(out))
    return 0


def bench_combined(a, ref_frames=256, target=0.01):
    """The combined 4K workload: 4 views of 3840x2160 over ~10 M instanced
    triangles (1,000 instances of 40 meshes), 1,024 emissive triangles, sky +
    sun, 16 bounces, rendered together; 2 progressive frames per view plain
    and denoised, and the time for each method to reach relMSE 0.01 against
    the views' plain ref_frames images (mean relMSE over the views)."""
    import numpy as np
    from helios_amd import runtime, scenegen
    from helios_amd.denoise import Denoiser
    from helios_amd.abi import HH_FLAG_LIGHT_POWER
    desc, cams = scenegen.combined_workload(equirect_width=2048)
    ctx = runtime.Context(0)
    ctx.set_tuning(frames_in_flight=4, extend_block=64, persistent_blocks_per_cu=32)
    t0 = time.perf_counter()
    scene = runtime.Scene(ctx, desc)
    build_s = time.perf_counter() - t0
    W, H = desc.width, desc.height
    surfs = [runtime.Surface(ctx, W, H) for _ in cams]
    for s in surfs:
        s.enable_features()
    outs = [runtime.Surface(ctx, W, H) for _ in cams]
    dn = Denoiser(ctx, seed=0)
    if Path(a.weights).exists():
        dn.set_params(np.fromfile(a.weights, np.float32))

    def views(spp, first):
        return [dict(scene=scene, surface=s, camera=c, spp=spp, max_ray_bounces=16, first_frame=first,
                     flags=HH_FLAG_LIGHT_POWER) for s, c in zip(surfs, cams)]

    def clear():
        for s in surfs:
            s.clear()

    # reference images
    clear()
    t0 = time.perf_counter()
    runtime.render_batch(ctx, views(ref_frames, 0))
    refs = [s.read() for s in surfs]
    ref_s = time.perf_counter() - t0

    def err(imgs):
        return float(np.mean([relmse(x, r) for x, r in zip(imgs, refs)]))

    # progressive scan: plain and denoised relMSE by frame count
    clear()
    done, k, curve, k_plain, k_dn = 0, 1, [], None, None
    while k <= ref_frames // 2:
        runtime.render_batch(ctx, views(k - done, done))
        done = k
        e_plain = err([s.read() for s in surfs])
        for s, o in zip(surfs, outs):
            dn.denoise(s, k, o)
        e_dn = err([o.read() for o in outs])
        curve.append((k, round(e_plain, 5), round(e_dn, 5)))
        if k_dn is None and e_dn <= target:
            k_dn = k
        if k_plain is None and e_plain <= target:
            k_plain = k
        if k_plain and k_dn:
            break
        k = max(k + 1, int(k * 1.15))

    def wall(k, denoise):
        ts = []
        for _ in range(3):
            clear()
            ctx.sync()
            t0 = time.perf_counter()
            runtime.render_batch(ctx, views(k, 0))
            if denoise:
                for s, o in zip(surfs, outs):
                    dn.denoise(s, k, o)
            ctx.sync()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts))

    two_plain = curve[1][1] if len(curve) > 1 else None
    two_dn = curve[1][2] if len(curve) > 1 else None
    t2 = wall(2, False)
    res = {
        "metric": "time to relMSE 0.01, denoised vs plain, combined 4K workload (1 GPU)",
        "value": None, "unit": "x", "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "dtype": "fp32",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "4 views of 3840x2160, 1,000 instances of 40 displaced meshes "
                               f"({desc.num_triangles:,} unique / "
                               f"{sum(desc.meshes[m].indices.size // 3 for (_, _, m, _) in desc.instances):,} "
                               "instanced triangles), 1,024 emissive triangles sampled in proportion to power, "
                               "Hosek-Wilkie sky as a 2048x1024 RGBA16F equirectangular map + sun, 16 bounces, "
                               "one batch",
                   "reference_frames": ref_frames, "tuning": ctx.tuning()},
        "scene_build_s": round(build_s, 2), "reference_s": round(ref_s, 1),
        "two_frames_ms": round(t2 * 1e3, 2),
        "two_frames_msamples_per_s": round(4 * W * H * 2 / t2 / 1e6, 1),
        "relmse_after_2_frames": {"plain": two_plain, "denoised": two_dn},
        "plain_frames": k_plain, "denoised_frames": k_dn, "curve_frames_plain_denoised": curve,
    }
    if k_plain and k_dn:
        res["plain_s"] = round(wall(k_plain, False), 4)
        res["denoised_s"] = round(wall(k_dn, True), 4)
        res["value"] = round(res["plain_s"] / res["denoised_s"], 2)
    print(json.dumps(res))
    return 0


def main():
    a = parse()
    if a.workload == "batch64" and a.gpus == 1:
        return bench_batch(a)
    if a.workload == "combined4k" and a.gpus == 1:
        return bench_combined(a)
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; run with --gpus 1 (multi-GPU sharding is "
                                   "design-only: frames split by index, RGBA32F sums reduced over xGMI)"}))
        return 2
    from helios_amd import runtime, scenegen

    desc = scenegen.icosphere_field(seed=0, width=a.width, height=a.height, max_ray_bounces=a.bounces)
    ctx = runtime.Context(0)
    # The timed call traces `steps` frames in passes of `chunk` frames (plus a
    # remainder pass); --probe traces one pass of frames_in_flight frames.
    chunk = a.frames_in_flight if a.probe else max(1, min(a.steps, a.frames_in_flight))
    a.continuation_depth = continuation_depth(a.continuation_depth, chunk)
    ctx.set_tuning(bvh_width=a.bvh_width, extend_block=a.block, persistent_blocks_per_cu=a.blocks_per_cu,
                   frames_in_flight=a.frames_in_flight, bvh_compressed=a.bvh_compressed,
                   refill_lanes=a.refill_lanes, traversal=a.traversal, bvh_builder=a.bvh_builder,
                   continuation_depth=a.continuation_depth)
    node_bytes = 64 if (a.bvh_compressed and a.bvh_width == 4) else NODE_BYTES[a.bvh_width]
    t0 = time.perf_counter()
    scene = runtime.Scene(ctx, desc)
    build_s = time.perf_counter() - t0
    surf = runtime.Surface(ctx, a.width, a.height)
    it = runtime.PathIntegrator(ctx)
    it.set_max_ray_bounces(a.bounces)
    it.set_max_samples(1 << 30)
    if a.megakernel:
        from helios_amd.abi import HH_FLAG_MEGAKERNEL
        it.set_flags(HH_FLAG_MEGAKERNEL)

    if a.denoise_probe:
        from helios_amd.denoise import Denoiser, forward_flops
        import numpy as np
        surf.enable_features()
        it.render(scene, desc.camera, surf, frames=3)
        dn = Denoiser(ctx, seed=0)
        if Path(a.weights).exists():
            dn.set_params(np.fromfile(a.weights, np.float32))
        out_s = runtime.Surface(ctx, a.width, a.height)
        for _ in range(10):
            dn.denoise(surf, 3, out_s)
        ctx.sync()
        print(json.dumps({"probe": "10 denoiser passes at %dx%d" % (a.width, a.height),
                          "gflop_per_pass": forward_flops(a.height, a.width) / 1e9}))
        return 0

    if a.probe:
        it.render(scene, desc.camera, surf, frames=a.frames_in_flight)
        ctx.sync()
        gbps = ctx.bandwidth_probe(1 << 30, 1)
        print(json.dumps({"probe": f"{a.frames_in_flight} frames (one chunk) + 1 GiB copy", "copy_gbps": gbps,
                          "rays": it.ray_stats(),
                          "shape": {"frames_in_flight": chunk, "continuation_depth": a.continuation_depth,
                                    "bvh_builder": a.bvh_builder}}))
        return 0

    # Warm-up first renders the W requested frames, then the timed call's own
    # pass sequence once, so the path state is already sized for the largest
    # pass and every pass's hipGraph is captured and instantiated before the
    # clock starts.
    rem = a.steps % chunk
    it.render(scene, desc.camera, surf, frames=a.warmup)
    it.render(scene, desc.camera, surf, frames=chunk)
    if rem:
        it.render(scene, desc.camera, surf, frames=rem)
    ctx.sync()
    t0 = time.perf_counter()
    it.render(scene, desc.camera, surf, frames=a.steps)   # returns after a stream sync
    ctx.sync()
    dt = time.perf_counter() - t0
    stats = it.ray_stats()
    pixels = a.width * a.height
    msamples = pixels * a.steps / dt / 1e6
    gsb = stats["extension_rays"] / dt / 1e9

    out = {
        "metric": METRIC,
        "value": round(msamples, 3),
        "unit": "Msamples/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(dt / a.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded procedural scene)",
        "config": {
            "workload": "1920x1080 icosphere field: 64 level-5 displaced icospheres (1,310,720 tris) + ground + "
                        "2,000 alpha leaf quads, Hosek-Wilkie sky + sun + 4 point lights, 8 bounces, 1 spp/frame",
            "width": a.width, "height": a.height, "triangles": desc.num_triangles, "bounces": a.bounces,
            "spp_per_frame": 1, "frames_in_flight": chunk, "frames_in_flight_setting": a.frames_in_flight,
            "bvh_width": a.bvh_width,
            "bvh_compressed": bool(a.bvh_compressed), "block": a.block, "blocks_per_cu": a.blocks_per_cu,
            "refill_lanes": a.refill_lanes, "traversal": "if-if" if a.traversal else "while-while",
            "bvh_builder": BUILDERS[a.bvh_builder],
            "integrator": "megakernel" if a.megakernel else "wavefront+hipGraph",
            "continuation_depth": a.continuation_depth,
        },
        "gsample_bounces_per_s": round(gsb, 4),
        "shadow_rays_per_frame": stats["shadow_rays"] / max(stats["frames"], 1),
        "extension_rays_per_frame": stats["extension_rays"] / max(stats["frames"], 1),
        "scene_build_s": round(build_s, 3),
        "time_to_relmse_0.01": None,
    }

    if not a.no_roofline and not a.megakernel:
        out["roofline"], out["kernel_ms_per_frame"] = roofline(ctx, it, scene, desc, surf, chunk, node_bytes, a)

    if not a.megakernel:
        out["single_frame"] = single_frame(it, scene, desc, surf, pixels)

    out["hbm_copy_gbps"] = round(ctx.bandwidth_probe(1 << 30, 4), 1)

    if not a.no_denoise:
        out["time_to_relmse_0.01"] = time_to_relmse(ctx, scene, desc, dt / a.steps * 1e3, a.weights, a.ref_frames)

    if not a.no_cpu_baseline:
        try:
            out["cpu_baseline"] = cpu_baseline(desc, a.cpu_seconds)
        except Exception as e:  # the baseline is reported, never required
            out["cpu_baseline"] = {"error": str(e)}

    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
