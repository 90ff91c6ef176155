#!/usr/bin/env python3
"""Throughput of the photon march on one MI355X: Mrays/s at 1080p hole+disc.

Workload (one "step"): the 1920x1080 hole + thin-disc scene (disc 3..12 r_s,
camera at 20 r_s, inclination 80 deg, fov 60 deg) rendered along its
120-frame circular camera path (3 deg per frame), replayed as one HIP graph:
248,832,000 rays per step, FP32 march with FP64 set-up, then the FP64 pass
over the photon-sphere band and the rays whose FP32 texel lies within the
FP32 error of a texel edge.  Inputs are resident in HBM before the timed
region; outputs stay in HBM.

Prints one JSON line.  `roofline` covers the dominant kernel (the FP32
march, every frame of the path in one launch): algorithmic FLOPs of the
path (blackhole_8_amd/flops.py, from each frame's own per-pixel step counts)
over that launch's HIP-event duration on the context stream, against the
FP32 vector peak; beside it the FMA throughput a plain (not packed) FP32
instruction stream reaches on this device, measured here (bh_fma_issue_rate),
and the march's VALU busy fraction from rocprof counters (profiles/).
`accuracy` compares frames 0, 59 and 119 of the timed graph's own output
with the FP64 oracle.  `spheres` gives the same roofline for the 720p
16-sphere scene's march kernel (march_kernel<true,true,20>) and its whole-frame
time.  `cpu_baseline` is the FP64 oracle (a C restatement
of the reference march) on this host's cores.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

FP32_PEAK_TFLOPS = 157.3        # MI355X FP32 vector peak (published)
TRAFFIC_FILE = ROOT / "profiles" / "march32_disc_orbit_traffic.json"
VALU_FILE = ROOT / "profiles" / "march32_disc_orbit_valu.json"


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def host_threads() -> int:
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def batch_timing(reps: int = 3) -> dict:
    """The 512-scene batch of uneven renders in one persistent launch: best of
    reps runs with the default knobs and with the autotuned ones."""
    import numpy as np

    from blackhole_8_amd import Renderer, flops, scenes, tune

    r = Renderer(0)
    try:
        items, tex = scenes.uneven_batch()
        r.set_textures(tex)
        rays = sum(s.camera.width * s.camera.height for s, _ in items)

        def best_of(tuning):
            runs = [r.render_batch(items, timing=True, tuning=tuning) for _ in range(reps)]
            _, ms, st = min(runs, key=lambda x: x[2].total_ms)
            return ms, st

        ms0, st0 = best_of(None)
        best, rec = tune.autotune(r, items, reps=2)
        ms1, st1 = best_of(best)
        # algorithmic work of the batch, from every scene's own steps and hits
        outs = r.render_batch(items, want=("hit", "steps"))
        flop = sum(flops.frame_flops(o["steps"], o["hit"], bool(s.has_disc))
                   for o, (s, _) in zip(outs, items))
    finally:
        r.close()

    def line(ms, st):
        tf = flop / (st.march_ms * 1e-3) / 1e12
        return {"total_ms": round(st.total_ms, 4), "march_ms": round(st.march_ms, 4),
                "band_ms": round(st.band_ms, 4), "band_rays": int(st.band_rays),
                "mrays_per_s": round(rays / st.total_ms / 1e3, 1),
                "march_tflops": round(tf, 2), "frac_fp32_peak": round(tf / FP32_PEAK_TFLOPS, 4),
                "scene_ms_median": round(float(np.median(ms)), 4),
                "scene_ms_p90": round(float(np.quantile(ms, 0.9)), 4),
                "scene_ms_max": round(float(ms.max()), 4)}

    return {"workload": "uneven_batch: 512 scenes, side log-uniform 32..1024, r_s U[0.5,2], camera "
                        "U[6,50] r_s, inclination U[0,90] deg, disc with p 0.5, FP32, one "
                        "batch call (grid, persistent-tile or lane-refill mode, autotuned)",
            "scenes": len(items), "rays": int(rays), "algorithmic_gflop": round(flop / 1e9, 2),
            "default": line(ms0, st0),
            "tuned": dict(line(ms1, st1), tuning={f: getattr(best, f)
                                                   for f, _ in best._fields_})}


def cpu_side_by_side(iters: int = 20) -> dict:
    """The other workloads' GPU time beside the FP64 oracle on the host's cores
    for the same frame, and the oracle on one thread (the reference's serial
    loop) for one frame of the 1080p disc orbit."""
    import oracle
    from blackhole_8_amd import Renderer, diff, scenes

    threads = host_threads()

    def timed(fn):
        t = time.perf_counter()
        fn()
        return time.perf_counter() - t

    out = {"cpu": cpu_model(), "cpu_threads": threads}
    r = Renderer(0)
    try:
        for name, p, what in (
                ("plain_hole", scenes.plain_hole(),
                 "hole alone, 320x180, FP64 reference arithmetic, escape at 100 r_s"),
                ("spheres", scenes.spheres(), "hole + 16 textured spheres, 1280x720, FP32")):
            r.set_textures(p.textures)
            _, gpu_ms, _ = r.time_render(p.scene, p.options, iters)
            rays = p.scene.camera.width * p.scene.camera.height
            cpu_s = timed(lambda: oracle.render_preset(p, threads))
            out[name] = {"workload": what, "rays": rays, "gpu_ms": round(gpu_ms, 4),
                         "cpu_ms": round(cpu_s * 1e3, 2),
                         "gpu_mrays_per_s": round(rays / gpu_ms / 1e3, 1),
                         "cpu_mrays_per_s": round(rays / cpu_s / 1e6, 3)}
        tex = diff.load_target_textures(r)
        dr = diff.DiffRenderer(r, diff.DiffConfig())
        fwd_ms, _ = dr.time(diff.FIT_TRUE, iters)
        rays = dr.cfg.width * dr.cfg.height
        cpu_s = timed(lambda: oracle.render_diff(dr.cfg.params(diff.FIT_TRUE), tex, threads))
        out["fit_forward"] = {"workload": "differentiable forward of the 512x512 inverse-fit "
                                          "scene, FP32 (oracle in FP64)",
                              "rays": rays, "gpu_ms": round(fwd_ms, 4),
                              "cpu_ms": round(cpu_s * 1e3, 2),
                              "gpu_mrays_per_s": round(rays / fwd_ms / 1e3, 1),
                              "cpu_mrays_per_s": round(rays / cpu_s / 1e6, 3)}
    finally:
        r.close()
    p = scenes.disc_orbit(frames=1)
    rays = p.scene.camera.width * p.scene.camera.height
    cpu_s = timed(lambda: oracle.render_preset(p, 1))
    out["disc_orbit_one_thread"] = {"workload": "frame 0 of the 1080p disc orbit, FP64 oracle "
                                                "on one thread",
                                    "rays": rays, "cpu_ms": round(cpu_s * 1e3, 1),
                                    "cpu_mrays_per_s": round(rays / cpu_s / 1e6, 3)}
    return out


def spheres_timing(iters: int = 20) -> dict:
    """The 16-sphere scene (1280x720, FP32, escape to r = 100 with a
    background): the FP32 march kernel alone (march_kernel<true,true,20>, iters
    back-to-back launches between HIP events on the context stream) against
    the scene's algorithmic work, and the whole frame with its FP64 pass."""
    from blackhole_8_amd import Renderer, flops, scenes

    p = scenes.spheres()
    r = Renderer(0)
    try:
        r.set_textures(p.textures)
        out = r.render(p.scene, p.options, ("hit", "steps"))
        main_ms, total_ms, queued = r.time_render(p.scene, p.options, iters)
    finally:
        r.close()
    s = p.scene
    in_plane = flops.spheres_in_plane(s)
    work = flops.frame_flops(out["steps"], out["hit"], bool(s.has_disc), in_plane,
                             s.n_spheres, background=s.background >= 0)
    achieved = work / (main_ms * 1e-3) / 1e12
    rays = s.camera.width * s.camera.height
    return {"workload": "spheres: hole + 16 textured spheres (radius 0.3..1.5, 5..30 r_s), "
                        "1280x720, FP32, escape at 100 r_s to a background",
            "rays": rays, "steps_per_ray": round(float(out["steps"].mean()), 2),
            "spheres_in_plane_per_ray": round(float(in_plane.mean()), 3),
            "algorithmic_gflop": round(work / 1e9, 4),
            "roofline": {"bound": "valu",
                         "kernel": "march_kernel<true,true,20> (FP32 event march, spheres)",
                         "achieved": round(achieved, 3), "peak": FP32_PEAK_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
                         "launch_ms": round(main_ms, 5)},
            "frame_ms": round(total_ms, 5), "fp64_pass_rays": int(queued),
            "mrays_per_s": round(rays / total_ms / 1e3, 1), "iters": iters}


def adjoint_timing(iters: int = 50) -> dict:
    """Forward+adjoint against the forward on the 512x512 inverse-fit scene
    (FP32): the differentiable forward, the fused forward+adjoint launch, and
    the plain march (reference rule) of the same scene."""
    from blackhole_8_amd import Renderer, abi, diff, scenes

    r = Renderer(0)
    try:
        diff.load_target_textures(r)
        dr = diff.DiffRenderer(r, diff.DiffConfig())
        fwd, fwd_adj = dr.time(diff.FIT_TRUE, iters)
        pl = scenes.fit_plain(diff.FIT_TRUE, dr.cfg.width, dr.cfg.height, dr.cfg.fov,
                                  dr.cfg.r_outer)
        plain, plain_frame, _ = r.time_render(pl.scene, pl.options, iters)
    finally:
        r.close()
    return {"workload": "fit: 512x512 hole + disc, r_s 1.2, inclination 72 deg, camera at "
                        "18 r_s, r_inner 3.5 r_s, FP32; d/d(r_s, incl, dist, r_inner)",
            "fwd_ms": round(fwd, 5), "fwd_adj_ms": round(fwd_adj, 5),
            "ratio_vs_diff_fwd": round(fwd_adj / fwd, 3),
            "plain_fwd_ms": round(plain, 5), "ratio_vs_plain_fwd": round(fwd_adj / plain, 3),
            "plain_frame_ms": round(plain_frame, 5),
            "ratio_vs_plain_frame": round(fwd_adj / plain_frame, 3),
            "note": "plain_fwd_ms is the plain march kernel alone; plain_frame_ms adds its "
                    "FP64 photon-sphere band pass (eager, serial)",
            "iters": iters}


def flat_timing(iters: int = 50) -> dict:
    """The flat-space scene of the ray_tracer_test program (FP64, the
    reference's operation order): kernel time per frame at the reference's
    800x450 and at 1920x1080."""
    from blackhole_8_amd import Renderer, scenes

    r = Renderer(0)
    out = {"workload": "ray_tracer_test scene: 3 textured rectangles + chess plane, "
                       "RayTracer::Prograde 10 chords per pixel, FP64"}
    try:
        for w, h in ((800, 450), (1920, 1080)):
            s, tex = scenes.flat_test(w, h)
            r.set_textures(tex)
            ms = r.time_flat(s, iters)
            out[f"{w}x{h}"] = {"kernel_ms": round(ms, 5),
                               "mrays_per_s": round(w * h / ms / 1e3, 1)}
    finally:
        r.close()
    return out


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=8,
                    help="untimed launches first: the clock ramps up over the first few")
    ap.add_argument("--frames", type=int, default=120)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-frames", type=int, default=12,
                    help="frames of the path in the CPU baseline sample (about 25 "
                         "core-seconds at 12)")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--adjoint-only", action="store_true",
                    help="only time the 512x512 forward / forward+adjoint and exit")
    ap.add_argument("--side-only", action="store_true",
                    help="print only the GPU-beside-CPU timings of the other workloads")
    ap.add_argument("--batch-only", action="store_true",
                    help="only time the 512-scene batch and exit")
    ap.add_argument("--no-flat", action="store_true")
    ap.add_argument("--no-spheres", action="store_true")
    ap.add_argument("--spheres-only", action="store_true",
                    help="only time the 16-sphere 720p scene and exit")
    ap.add_argument("--flat-only", action="store_true",
                    help="only time the flat-space (ray_tracer_test) scene and exit")
    ap.add_argument("--copy-calibration-mib", type=int, default=0,
                    help="only run a plain device copy of this many MiB (3 times), for "
                         "scaling the FETCH_SIZE / WRITE_SIZE counters, and exit")
    args = ap.parse_args()

    if args.copy_calibration_mib:
        import torch

        n = args.copy_calibration_mib << 20
        src = torch.ones(n, dtype=torch.uint8, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": n, "copies": 3}))
        return 0

    if args.adjoint_only:
        print(json.dumps({"adjoint": adjoint_timing()}))
        return 0
    if args.batch_only:
        print(json.dumps({"batch": batch_timing()}))
        return 0
    if args.flat_only:
        print(json.dumps({"flat": flat_timing()}))
        return 0
    if args.spheres_only:
        print(json.dumps({"spheres": spheres_timing()}))
        return 0
    if args.side_only:
        print(json.dumps({"cpu_side_by_side": cpu_side_by_side()}))
        return 0

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU only: the renderer runs on a single MI355X "
              "(rays would shard by frame or tile across GPUs with no exchange). "
              f"Requested --gpus {args.gpus}, WORLD_SIZE {world}.", file=sys.stderr)
        return 2

    import numpy as np

    from blackhole_8_amd import Renderer, abi, flops, scenes

    p = scenes.disc_orbit(frames=args.frames)
    W, H, F = p.width, p.height, len(p.cameras)
    rays_per_step = W * H * F

    r = Renderer(0)
    dev = r.device_info()
    r.set_textures(p.textures)
    path = r.path(p.scene, p.options, p.cameras)

    # ---- timed region: K graph launches of the 120-frame path -------------
    for _ in range(args.warmup):
        path.launch()
    r.lib.bh_synchronize(r.ctx)
    t0 = time.perf_counter()
    gpu_ms = march_sum = band_sum = 0.0
    for _ in range(args.steps):
        gpu_ms += path.launch()
        m_ms, b_ms = path.last_ms()  # the graph's own timing events around its kernels
        march_sum += m_ms
        band_sum += b_ms
    r.lib.bh_synchronize(r.ctx)
    wall = time.perf_counter() - t0
    ms_per_step = wall * 1e3 / args.steps
    value = rays_per_step * args.steps / wall / 1e6
    # frames of the last timed replay, for the accuracy check below
    acc_frames = [0, F // 2 - 1, F - 1] if F >= 3 else list(range(F))
    path_frames = {k: path.read(k) for k in acc_frames}

    # ---- dominant kernel: the fused march launch and algorithmic work -------
    # All frames of the path march in one launch (march_frames_kernel); its
    # duration is the mean over the timed replays of the HIP events the graph
    # records on the context stream around that kernel (bh_path_last_ms), the
    # work comes from each frame's own per-pixel step counts and hit classes.
    march_ms, band_ms = march_sum / args.steps, band_sum / args.steps
    work, band = 0.0, 0
    alg_bytes = 0  # framebuffer (3 B rgb + 1 B class) and one 4-byte texel per disc hit
    for k in range(F):
        sc = p.with_camera(p.cameras[k]).scene
        out = r.render(sc, p.options, ("hit", "steps"))
        work += flops.frame_flops(out["steps"], out["hit"], True)
        alg_bytes += W * H * 4 + 4 * int((out["hit"] == abi.HIT_DISC).sum())
    single = [r.time_render(p.with_camera(p.cameras[k]).scene, p.options, iters=4)
              for k in (0, F // 2)]
    achieved = work / (march_ms * 1e-3) / 1e12
    import ctypes as C

    fma32, fma64 = C.c_double(), C.c_double()
    r._check(r.lib.bh_fma_issue_rate(r.ctx, C.byref(fma32), C.byref(fma64)))
    traffic = None
    if TRAFFIC_FILE.exists():
        try:
            traffic = json.loads(TRAFFIC_FILE.read_text()).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    valu = None
    if VALU_FILE.exists():
        try:
            valu = json.loads(VALU_FILE.read_text()).get("valu_busy")
        except (OSError, ValueError):
            valu = None
    roofline = {
        "bound": "valu",
        "kernel": "march_frames_kernel<false,false,20> (FP32 event march compiled for the reference's 20 steps, all frames in one launch)",
        "achieved": round(achieved, 3),
        "peak": FP32_PEAK_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / FP32_PEAK_TFLOPS, 4),
        "fp32_fma_issue_tflops": round(fma32.value, 2),
        "fp64_fma_issue_tflops": round(fma64.value, 2),
        "frac_of_fma_issue": round(achieved / fma32.value, 4),
        "valu_busy": valu,
        "traffic": traffic,
        "traffic_per": "launch (all frames of the path)",
        "traffic_gbps": round(traffic / (march_ms * 1e-3) / 1e9, 1) if traffic else None,
        "algorithmic_bytes": alg_bytes,
        "algorithmic_gbps": round(alg_bytes / (march_ms * 1e-3) / 1e9, 1),
        "hbm_peak_gbps": 8000,
        "launch_ms": round(march_ms, 4),
        "avg_kernel_ms": round(march_ms / F, 5),
        "band_launch_ms": round(band_ms, 4),
        "single_frame_kernel_ms": round(sum(m for m, _, _ in single) / len(single), 5),
        "flop_per_frame": round(work / F),
        "band_rays_per_frame": round(sum(n for _, _, n in single) / len(single), 1),
    }

    # ---- accuracy of the timed graph's frames and the CPU baseline ----------
    cpu = None
    accuracy = None

    if not args.no_cpu_baseline:
        import oracle
        from tests.parity import compare

        threads = host_threads()
        want = ("rgb", "hit", "phi", "hitpt", "steps")
        frames = [int(round(i * F / args.cpu_frames)) for i in range(args.cpu_frames)]
        cpu_s = 0.0
        ref0 = None
        for k in frames:
            pk = p.with_camera(p.cameras[k])
            t = time.perf_counter()
            ref = oracle.render_preset(pk, threads, want if k == 0 else ("rgb", "hit"))
            cpu_s += time.perf_counter() - t
            if k == 0:
                ref0 = ref
        cpu = {"value": round(W * H * len(frames) / cpu_s / 1e6, 3), "unit": "Mrays/s",
               "cores": threads, "kind": "port",
               "sample": f"frames {frames} of the 120-frame 1080p path, FP64 oracle "
                         f"(C restatement of the reference march), {threads} threads, "
                         f"{cpu_s:.2f} s wall, {cpu_model()}"}
        # frames of the graph's output (the timed launch's planes) against the
        # oracle; frame 0's angle and hit-point errors from a render of the
        # same march kernel with every plane
        accuracy = {"path_frames": {}}
        for k in acc_frames:
            rk = ref0 if k == 0 else oracle.render_preset(p.with_camera(p.cameras[k]),
                                                          threads, ("rgb", "hit"))
            rgb, hit = path_frames[k]
            st = compare({"rgb": rgb, "hit": hit}, rk)
            accuracy["path_frames"][str(k)] = {
                "class_mismatch_frac": st["class_mismatch_frac"],
                "rgb_mismatch_frac": st["rgb_mismatch_frac"]}
        g0 = r.render(p.with_camera(p.cameras[0]).scene, p.options, want)
        st = compare(g0, ref0)
        accuracy["frame0"] = {k: st[k] for k in ("class_mismatch_frac", "rgb_mismatch_frac",
                                                 "phi_max_err", "hit_point_max_err")}
    path.close()
    r.close()

    side = None if args.no_cpu_baseline else cpu_side_by_side()
    adjoint = None if args.no_adjoint else adjoint_timing()
    batch = None if args.no_batch else batch_timing()
    flat = None if args.no_flat else flat_timing()
    spheres = None if args.no_spheres else spheres_timing()

    line = {
        "metric": "Mrays/s at 1080p BH+disc; fwd+adjoint ms/frame vs plain fwd; % FP32 peak",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "disc_orbit: hole + thin disc 3..12 r_s, 1920x1080, camera at "
                               "20 r_s, inclination 80 deg, 120-frame orbit path as one HIP "
                               "graph, FP32 march + FP64 set-up, FP64 pass over the "
                               "photon-sphere band and texel-edge rays",
                   "width": W, "height": H, "frames": F, "rays_per_step": rays_per_step,
                   "gpu_event_ms_per_step": round(gpu_ms / args.steps, 4),
                   "device": dev["name"], "cu": dev["cu"]},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_side_by_side": side,
        "accuracy": accuracy,
        "adjoint": adjoint,
        "batch": batch,
        "flat": flat,
        "spheres": spheres,
    }
    print(json.dumps(line))
    return 0


if __name__ == "__main__":
    sys.exit(main())
