"""DDGI probe-update benchmark on one MI355X.

Workload (the largest single-GPU configuration): plain DDGI, large — 262,144-triangle clustered
scene, 32x16x32 = 16,384 probes x 256 rays, 8x8 irradiance / 16x16 depth FP32 atlases, all probes
updated every step.  A step is one full update: trace (primary rays + shading + shadow rays) +
irradiance/depth blend with hysteresis + border copy (fused) + publish.

Prints ONE JSON line (rank 0).  `value` = primary probe rays per second of whole updates, in
Mrays/s, over `--steps` updates enqueued back to back (no timing events, no host wait between
them); `phases_ms` come from a second pass over the same updates with per-phase HIP events.
`roofline` describes the kernel with the largest share of the step: its algorithmic bytes come from
the traversal's own fetch counters (128-byte 4-wide nodes, 48-byte triangles) of an untimed
counting update, set against the HBM peak; `traffic` (bytes leaving L2 per launch, from the
committed rocprofv3 counter summary) shows how few of them reach HBM, and `cache_ceiling` is the
fastest device copy that stays in L2 / the Infinity Cache, measured in the same run.
`cpu_baseline` times the CPU oracle (FP64 shading, FP32 intersection replica) through its thread
pool on a bounded sample of the same frame (1 warm-up + median of 5, FP32 variant beside it);
the same leg checks the mixed-precision blend workload (`blend_config3`) against the oracle's FP64
blend.
"""
from __future__ import annotations

import argparse
import dataclasses
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
BF16_PEAK_TFLOPS = 2500.0
METRIC = "DDGI update ms + Mrays/s; BF16-MFMA vs FP32 blend @<=1e-3 rel err; bwd/fwd"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--cpu-probes", type=int, default=4096, help="probes in the CPU-baseline sample")
    ap.add_argument("--cpu-threads", type=int, default=16,
                    help="cap on the baseline's threads (the host CPU share of one GPU)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--blend", choices=["fp32", "bf16", "fp16"], default="fp32")
    ap.add_argument("--no-blend-compare", action="store_true")
    ap.add_argument("--no-diff", action="store_true")
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-frame", action="store_true")
    ap.add_argument("--no-dynamic", action="store_true")
    ap.add_argument("--no-sky", action="store_true")
    ap.add_argument("--no-peaks", action="store_true", help="skip the achievable-peak measurements")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="also run 3 elementwise copies of 256 MiB (torch.add(x, 0)) so that profiler byte "
                         "counters can be scaled against a kernel of known traffic")
    return ap.parse_args()


def host_cpu():
    """The host's CPU as the baseline reports it: lscpu's model name and physical cores (cores per
    socket x sockets), std::thread::hardware_concurrency() (= os.cpu_count()) and the CPUs this
    process may run on."""
    info = {"model": None, "physical_cores": None, "hardware_concurrency": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}
    try:
        import subprocess
        kv = {}
        for line in subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout.splitlines():
            if ":" in line:
                k, v = line.split(":", 1)
                kv[k.strip()] = v.strip()
        info["model"] = kv.get("Model name")
        if kv.get("Core(s) per socket") and kv.get("Socket(s)"):
            info["physical_cores"] = int(kv["Core(s) per socket"]) * int(kv["Socket(s)"])
    except (OSError, ValueError):
        pass
    return info


def cpu_baseline(scene, grid, light, sky, R, n_probes, max_threads):
    """The oracle (FP64 shading over the FP32 intersection replica) on the host through its thread
    pool, on a bounded sample of the same frame: `n_probes` probes spread evenly over the grid, each
    traced, shaded, blended and border-copied as in the update.  Threads: hardware_concurrency(),
    capped by the CPUs this process may use and by `max_threads` (the host share of one GPU).  One
    warm-up, then the median of 5 runs, per phase and in total; the oracle's FP32 instantiation is
    timed the same way beside it.  Returns the baseline object and the states of the sample."""
    from oracle import py as orc
    cpu = host_cpu()
    threads = max(1, min(cpu["hardware_concurrency"] or 1, cpu["affinity_cpus"], max_threads))
    osc = orc.OracleScene(scene)
    og = orc.Grid(grid)
    used = orc.set_threads(threads)
    states, _ = osc.init_probes(og, R)
    idx = np.linspace(0, og.count - 1, n_probes).astype(np.uint32)
    n = int(grid["rays_per_probe"])

    def run(fp32):
        res = []
        for k in range(6):  # 1 warm-up + 5 timed
            irr, dep = og.zero_atlases()
            t0 = time.perf_counter()
            r = osc.update(og, light, sky, R, idx, states.copy(), irr, dep, fp32=fp32)
            if k:
                res.append((time.perf_counter() - t0, r["trace_ms"], r["blend_ms"], r["border_ms"]))
        med = np.median(np.array(res), axis=0)
        return {"total_ms": round(med[0] * 1e3, 1), "trace_ms": round(med[1], 1), "blend_ms": round(med[2], 1),
                "border_ms": round(med[3], 2), "Mrays/s": round(n_probes * n / med[0] / 1e6, 3)}

    fp64 = run(False)
    fp32 = run(True)
    rate = fp64["Mrays/s"]
    return {"value": rate, "unit": "Mrays/s", "cores": used, "kind": "port",
            "sample": f"{n_probes} of {og.count} probes (evenly spaced) x {n} rays of one frame: trace + shade + "
                      f"shadow + blend + border, FP64 oracle, {used} threads, 1 warm-up + median of 5 = "
                      f"{fp64['total_ms']:.0f} ms",
            "threads": used, "cpu": cpu, "phases_ms": fp64, "fp32_variant": fp32,
            "update_ms_extrapolated": round(og.count * n / (rate * 1e6) * 1e3, 1)}


PROFILE_SUMMARY = os.path.join(ROOT, "profiles", "ddgi_update_large_rocprof.md")
PROFILE_KERNELS = {"trace_primary": "k_trace_primary<16, false>", "shade_shadow": "k_shade<16, false>",
                   "blend": "k_blend_wide<0u, 8>"}


def profiled_traffic(kernel: str):
    """HBM bytes per launch of `kernel` (FETCH_SIZE + WRITE_SIZE, scaled by a copy of known size) from
    the committed rocprofv3 summary of this workload; None if the summary does not list it."""
    name = PROFILE_KERNELS.get(kernel)
    if name is None or not os.path.exists(PROFILE_SUMMARY):
        return None
    for line in open(PROFILE_SUMMARY):
        parts = line.strip().split(" ")
        if parts[0] == "traffic_per_launch" and " ".join(parts[1:-1]) == name:
            return int(parts[-1])
    return None


def _event_ms(fn, reps=10):
    import torch
    times = []
    for _ in range(reps):
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        a.record()
        fn()
        b.record()
        b.synchronize()
        times.append(a.elapsed_time(b))
    return float(np.median(times))


def blend_compare(probes, R, lib):
    """FP32 / BF16 / FP16 blends of one ray buffer (16,384 probes x 256 rays), 10 times each: the
    blend phase (frame directions, weights, blend) and the error of BF16/FP16 against FP32.  The
    probe list covers the grid (UPDATE_ALL_PROBES), so no atlas is copied forward."""
    import torch
    state = probes.save_state()
    probes.update(R, lib.UPDATE_ALL_PROBES)
    probes.synchronize()
    rays, idx = probes.last_update_rays()
    t_rays = torch.from_numpy(rays).cuda()
    t_idx = torch.from_numpy(idx.view(np.int32).copy()).cuda()
    out = {}
    ref = None
    for mode, flag in (("fp32", 0), ("bf16", lib.UPDATE_BLEND_BF16), ("fp16", lib.UPDATE_BLEND_FP16)):
        ms = []
        for _ in range(10):
            probes.load_state(state)
            probes.blend_rays(t_rays, t_idx, R, lib.UPDATE_ALL_PROBES | flag)
            ms.append(probes.timings()["blend_ms"])
        probes.synchronize()
        irr = probes.irradiance(lib.LAYOUT_TILED)[..., :3].astype(np.float64)
        if ref is None:
            ref = irr
        err = np.abs(irr - ref) / np.maximum(np.abs(ref), 1e-3 * np.abs(ref).max())
        out[mode] = {"blend_ms": round(float(np.median(ms)), 4), "max_rel_err_vs_fp32": float(f"{err.max():.3e}"),
                     "mean_rel_err_vs_fp32": float(f"{err.mean():.3e}")}
    probes.load_state(state)
    out["bf16_speedup_vs_fp32"] = round(out["fp32"]["blend_ms"] / out["bf16"]["blend_ms"], 2)
    out["fp16_speedup_vs_fp32"] = round(out["fp32"]["blend_ms"] / out["fp16"]["blend_ms"], 2)
    return out


def blend_config3(lib):
    """The mixed-precision blend workload: the closed room + 64 boxes (780 triangles), 8x4x8 probes
    x 512 rays, cosine weights [36 x 512] for irradiance and cos^50 weights [196 x 512] for depth.
    One ray buffer from the update is blended by each mode with hysteresis 0 (the new-frame
    estimate), its blend phase (frame directions, weights, blend; the probe list covers the grid)
    timed over 10 blends; the atlases are kept for the error check against the FP64 blend (done in
    the CPU leg)."""
    import torch
    from vulkanexp_amd import IrradianceProbes, scenes
    scene, grid = scenes.config_small(seed=1, rays=512)
    grid = np.array(grid)
    grid["depth_sharpness"] = 50.0
    p = IrradianceProbes(grid, max_rays_per_probe=512)
    p.set_scene(scene)
    p.set_light(scenes.sun_light())
    p.set_sky(scenes.SKY_RADIANCE)
    Rs = scenes.orientations(33, 4)
    for R in Rs[:3]:
        p.update(R, lib.UPDATE_ALL_PROBES)
    p.GridParameters["hysteresis"] = 0.0
    p.update_uniforms()
    state = p.save_state()
    p.update(Rs[3], lib.UPDATE_ALL_PROBES)
    p.synchronize()
    rays, idx = p.last_update_rays()
    p.load_state(state)
    check = {"grid": np.array(p.GridParameters), "R": Rs[3], "rays": rays, "idx": idx, "states": p.states(),
             "irr": p.irradiance()[..., :3].astype(np.float64), "dep": p.depth().astype(np.float64), "out": {}}
    t_rays = torch.from_numpy(np.ascontiguousarray(rays)).cuda()
    t_idx = torch.from_numpy(idx.view(np.int32).copy()).cuda()
    out = {"workload": "closed room + 64 boxes (780 tris), 8x4x8 probes x 512 rays, depthSharpness 50, hysteresis 0"}
    for mode, flag in (("fp32", 0), ("bf16", lib.UPDATE_BLEND_BF16), ("fp16", lib.UPDATE_BLEND_FP16)):
        ms = []
        for _ in range(10):
            p.load_state(state)
            p.blend_rays(t_rays, t_idx, Rs[3], lib.UPDATE_ALL_PROBES | flag)
            ms.append(p.timings()["blend_ms"])
        p.synchronize()
        check["out"][mode] = (p.irradiance()[..., :3].astype(np.float64), p.depth().astype(np.float64))
        out[mode] = {"blend_ms": round(float(np.median(ms)), 4)}
    p.destroy()
    return out, check


def config3_errors(check):
    """Per-texel relative error of each blend mode against the oracle's FP64 blend of the same ray
    buffer: |gpu - fp64| / max(|fp64|, 1e-3 max |fp64|), max and mean, irradiance and depth."""
    from oracle import py as orc
    og = orc.Grid(check["grid"])
    irr, dep = check["irr"].copy(), check["dep"].copy()
    orc.blend(og, orc.frame_dirs(check["R"], int(og.params["rays_per_probe"])), check["rays"].astype(np.float64),
              check["idx"], check["states"].copy(), irr, dep)

    def rel(g, o):
        e = np.abs(g - o) / np.maximum(np.abs(o), 1e-3 * max(np.abs(o).max(), 1e-12))
        return {"max": float(f"{e.max():.3e}"), "mean": float(f"{e.mean():.3e}")}
    return {mode: {"irradiance": rel(gi, irr), "depth": rel(gd, dep)} for mode, (gi, gd) in check["out"].items()}


def frame_bench(probes, scene, width=1920, height=1080):
    """Full-frame consumer on the large scene: G-buffer of a 1920x1080 pinhole camera (back faces
    culled) and the final gather (hard sun shadow ray + sampleProbes + PBR) over it."""
    import torch
    from vulkanexp_amd import gbuffer
    eye = (8.0, 9.0, 6.0)
    rays = torch.from_numpy(gbuffer.camera_rays(width, height, eye=eye, look_at=(30.0, 2.0, 20.0))).cuda()
    gb = probes.gbuffer(rays)
    probes.gather(gb, eye)
    g_ms = _event_ms(lambda: probes.gbuffer(rays))
    f_ms = _event_ms(lambda: probes.gather(gb, eye))
    hit = float((gb[:, 3] > 0).float().mean())
    px = width * height
    # ray-traced passes: jittered shadow rays, and reflections with every hit texel made glossy
    # (roughness 0.25 < 0.4) so that each one traces and shades a reflected ray
    gen = torch.Generator(device="cuda").manual_seed(5)
    noise = torch.rand((px, 2), device="cuda", generator=gen)
    d_ms = _event_ms(lambda: probes.direct_light(gb, noise=noise))
    glossy = gb.clone()
    glossy[:, 11] = 0.25
    probes.reflection(glossy, eye, noise=noise)
    r_ms = _event_ms(lambda: probes.reflection(glossy, eye, noise=noise))
    return {"pixels": px, "hit_fraction": round(hit, 3), "gbuffer_ms": round(g_ms, 4), "gather_ms": round(f_ms, 4),
            "Mpixels/s": round(px / ((g_ms + f_ms) * 1e-3) / 1e6, 1), "shadow_pass_ms": round(d_ms, 4),
            "reflection_pass_ms": round(r_ms, 4), "reflection_rays": int(round(hit * px))}


def dynamic_bench(probes, scene, R, lib):
    """Dynamic scene on the large workload: every instance but the floor turns by a few degrees
    about its own origin and shifts by 0.5 m; one ddgi_update_instances call (device re-transform
    + bottom-up refit) against a host rebuild (ddgi_set_scene), and the probe update's time on the
    refitted tree.  The instances are moved back afterwards."""
    import time
    import torch
    n = len(scene.instances)
    xf0 = scene.instances["transform"][1:].reshape(-1, 3, 4).astype(np.float64)
    a = np.radians(7.0)
    rot = np.array([[np.cos(a), 0, np.sin(a)], [0, 1, 0], [-np.sin(a), 0, np.cos(a)]])
    xf1 = np.empty_like(xf0)
    xf1[:, :, :3] = rot @ xf0[:, :, :3]
    xf1[:, :, 3] = xf0[:, :, 3] + np.array([0.5, 0.0, 0.5])
    xf1 = xf1.astype(np.float32).reshape(-1, 12)
    xf0 = xf0.astype(np.float32).reshape(-1, 12)
    probes.update(R, lib.UPDATE_ALL_PROBES | lib.UPDATE_NO_TIMING)
    probes.synchronize()
    before = _event_ms(lambda: probes.update(R, lib.UPDATE_ALL_PROBES | lib.UPDATE_NO_TIMING), reps=5)
    gpu, wall = [], []
    for k in range(6):
        target = xf1 if k % 2 == 0 else xf0
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        t0 = time.perf_counter()
        e0.record()
        probes.update_instances(1, target)
        e1.record()
        e1.synchronize()
        wall.append((time.perf_counter() - t0) * 1e3)
        gpu.append(e0.elapsed_time(e1))
    probes.update_instances(1, xf1)
    probes.synchronize()
    after = _event_ms(lambda: probes.update(R, lib.UPDATE_ALL_PROBES | lib.UPDATE_NO_TIMING), reps=5)
    moved = dataclasses.replace(scene, instances=scene.instances.copy())
    moved.instances["transform"][1:] = xf1
    t0 = time.perf_counter()
    probes.set_scene(moved)
    probes.synchronize()
    rebuild_s = time.perf_counter() - t0
    rebuilt = _event_ms(lambda: probes.update(R, lib.UPDATE_ALL_PROBES | lib.UPDATE_NO_TIMING), reps=5)
    probes.set_scene(scene)
    probes.synchronize()
    return {"instances_moved": n - 1, "triangles_moved": int(scene.instances["triangle_count"][1:].sum()),
            "update_instances_ms": round(float(np.median(gpu)), 4), "update_instances_wall_ms": round(float(np.median(wall)), 4),
            "host_rebuild_s": round(rebuild_s, 3), "probe_update_ms": {"before_move": round(before, 4),
            "refitted": round(after, 4), "rebuilt": round(rebuilt, 4)}}


def sky_bench(probes, R, lib):
    """The probe update of the large workload with the atmospheric sky (sky.glsl's 64-sample
    scattering integral per missed ray) instead of the constant sky."""
    flags = lib.UPDATE_ALL_PROBES | lib.UPDATE_NO_TIMING
    constant = _event_ms(lambda: probes.update(R, flags), reps=5)
    probes.set_sky_model(lib.SKY_ATMOSPHERE)
    try:
        probes.update(R, flags)
        atmosphere = _event_ms(lambda: probes.update(R, flags), reps=5)
        rays, _ = probes.last_update_rays()
        miss = float((rays[..., 3] == -1).mean())
    finally:
        probes.set_sky_model(lib.SKY_CONSTANT)
    return {"update_ms_constant_sky": round(constant, 4), "update_ms_atmosphere": round(atmosphere, 4),
            "miss_fraction": round(miss, 3)}


def diff_bench(lib):
    """Differentiable probes: room + boxes scene (open top), 8x4x8 probes x 128 rays, a 1920x1080
    G-buffer traced on the GPU, 256 albedo slots + sun; forward and backward time."""
    import torch
    from vulkanexp_amd import IrradianceProbes, gbuffer, scenes
    from vulkanexp_amd.autograd import DiffContext
    scene, grid = scenes.config_small(seed=1, open_top=True)
    light = scenes.sun_light()
    p = IrradianceProbes(grid, max_rays_per_probe=128)
    p.set_scene(scene)
    p.set_light(light)
    p.set_sky(scenes.SKY_RADIANCE)
    Rs = scenes.orientations(41, 4)
    for R in Rs[:3]:
        p.update(R, lib.UPDATE_ALL_PROBES)
    rays = gbuffer.camera_rays(1920, 1080, eye=(2.0, 7.0, 2.0), look_at=(12.0, 1.0, 12.0))
    hits = gbuffer.hits_from_device(p.trace_rays(torch.from_numpy(rays).cuda()).cpu().numpy())
    gb = gbuffer.build(scene, rays, hits)
    t = lambda a, dt=torch.float32: torch.from_numpy(np.ascontiguousarray(a)).cuda().to(dt)
    dctx = DiffContext(p, Rs[3], t(gb["position"]), t(gb["normal"]), t(gb["to_camera"]),
                       t(gb["material"].astype(np.int32), torch.int32), 256)
    albedo, albedo_t, sun, sun_t = gbuffer.perturbed_params(scene.materials["base_color"], light["color"][:3], seed=5)
    A, S = t(albedo), t(sun)
    _, target = dctx.forward(t(albedo_t), t(sun_t), torch.zeros((dctx.N, 3), device="cuda"))
    target = target.clone()
    one = torch.ones(1, device="cuda")
    dctx.forward(A, S, target)
    dctx.backward(A, S, target, one)
    fwd = _event_ms(lambda: dctx.forward(A, S, target))
    bwd = _event_ms(lambda: dctx.backward(A, S, target, one))
    p.destroy()
    # algorithmic bytes per texel: forward 640 B of probe taps (8 cage probes x 4 irradiance and 4
    # depth texels, both atlases) + 32 B of probe states + 28 B of G-buffer + 12 B out = 712 B;
    # backward the same reads plus a 12-byte gradient per irradiance tap (8 x 4 x 12 = 384 B)
    fwd_b, bwd_b = 712 * int(dctx.N), (712 + 384) * int(dctx.N)
    return {"gbuffer_texels": int(dctx.N), "params": 256 * 3 + 3, "forward_ms": round(fwd, 4),
            "backward_ms": round(bwd, 4), "bwd_over_fwd": round(bwd / fwd, 3),
            "forward_algorithmic_GB/s": round(fwd_b / (fwd * 1e-3) / 1e9, 1),
            "backward_algorithmic_GB/s": round(bwd_b / (bwd * 1e-3) / 1e9, 1),
            "bytes_per_texel": {"forward": 712, "backward": 712 + 384}}


def batch_bench(lib, frames=5):
    """Batched volumes: 24 volumes over 6 scenes (780 .. 262,144 triangles), grids 4x4x4 ..
    32x16x32, 64..512 rays per probe (83,859 probes, 24.1 M rays per update).  One graph launch per
    update (after the autotuner picked its configuration) against the same updates issued as a
    sequence of per-volume ddgi_update calls on one stream."""
    import torch
    from vulkanexp_amd import IrradianceProbes, ProbeBatch, scenes
    scene_list, specs = scenes.batched_config(seed=5)
    light = scenes.sun_light()
    vols = []
    for si, g in specs:
        p = IrradianceProbes(g, max_rays_per_probe=int(g["rays_per_probe"]))
        p.set_scene(scene_list[si])
        p.set_light(light)
        p.set_sky(scenes.SKY_RADIANCE)
        vols.append(p)
    for k, p in enumerate(vols):
        p.init_probes(scenes.orientations(200 + k, 1)[0])
    nv = len(vols)
    rays = sum(p.get_probe_count() * int(p.GridParameters["rays_per_probe"]) for p in vols)
    Rs = [np.stack([scenes.orientations(10_000 * f + v, 1)[0] for v in range(nv)]) for f in range(frames + 1)]
    flags = lib.UPDATE_ALL_PROBES | lib.UPDATE_NO_TIMING
    for p, R in zip(vols, Rs[0]):
        p.update(R, flags)
    torch.cuda.synchronize()
    seq_total, seq_dur, seq_done = [], [], []
    for f in range(frames):
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(nv + 1)]
        ev[0].record()
        for v, (p, R) in enumerate(zip(vols, Rs[1 + f])):
            p.update(R, flags)
            ev[v + 1].record()
        ev[-1].synchronize()
        seq_dur += [ev[v].elapsed_time(ev[v + 1]) for v in range(nv)]
        seq_done += [ev[0].elapsed_time(ev[v + 1]) for v in range(nv)]
        seq_total.append(ev[0].elapsed_time(ev[-1]))
    batch = ProbeBatch(vols)
    batch.update(Rs[0])
    torch.cuda.synchronize()
    default_ms = []
    for f in range(frames):
        batch.update(Rs[1 + f])
        default_ms.append(batch.timings()[0])
    t0 = time.perf_counter()
    best, records = batch.autotune(Rs[0], iterations=3)
    tune_s = time.perf_counter() - t0
    batch.update(Rs[0])
    torch.cuda.synchronize()
    b_total, b_done = [], []
    for f in range(frames):
        batch.update(Rs[1 + f])
        total, per = batch.timings()
        b_total.append(total)
        b_done += list(per)
    batch.destroy()
    for p in vols:
        p.destroy()
    pct = lambda a, q: round(float(np.percentile(np.asarray(a), q)), 4)
    seq_ms, bat_ms = float(np.median(seq_total)), float(np.median(b_total))
    return {
        "volumes": nv, "probes": int(sum(p for p in [int(np.prod(g["resolution"])) for _, g in specs])),
        "rays_per_update": int(rays), "triangles": [s.triangle_count for s in scene_list],
        "sequential": {"total_ms": round(seq_ms, 4), "Mrays/s": round(rays / seq_ms / 1e3, 1),
                       "volume_update_ms_p50": pct(seq_dur, 50), "volume_update_ms_p99": pct(seq_dur, 99),
                       "volume_done_ms_p50": pct(seq_done, 50), "volume_done_ms_p99": pct(seq_done, 99)},
        "batched": {"total_ms": round(bat_ms, 4), "Mrays/s": round(rays / bat_ms / 1e3, 1),
                    "volume_done_ms_p50": pct(b_done, 50), "volume_done_ms_p99": pct(b_done, 99),
                    "default_config_total_ms": round(float(np.median(default_ms)), 4)},
        "speedup_vs_sequential": round(seq_ms / bat_ms, 3),
        "autotune": {"best": best, "seconds": round(tune_s, 2),
                     "records": [[r["trace_block"], r["trace_item_blocks"], r["blend_probes"], round(r["ms"], 4)]
                                 for r in records]},
    }


def achievable_peaks(torch):
    """Peaks measured on this GPU, for the roofline fractions beside the published ones: a device
    copy of 2 GiB (beyond the 256 MB Infinity Cache; read + written bytes per second) and an FP32
    GEMM of 8192^3 through torch (hipBLASLt / rocBLAS on the matrix cores)."""
    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e-3
    src = torch.ones(512 * 1024 * 1024, dtype=torch.float32, device="cuda")
    dst = torch.empty_like(src)
    t_copy = timed(lambda: dst.copy_(src), 10)
    copy_gbs = 2 * src.numel() * 4 / t_copy / 1e9
    del src, dst
    cache = {}
    for mib in (8, 32, 64):  # copies that stay in L2 (8 x 4 MB) and in the 256 MB Infinity Cache
        src = torch.ones(mib * 1024 * 1024 // 4, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        t = timed(lambda: dst.copy_(src), 200)
        cache[f"{mib}MiB"] = round(2 * src.numel() * 4 / t / 1e9, 1)
        del src, dst
    a = torch.randn(8192, 8192, device="cuda", dtype=torch.float32)
    b = torch.randn(8192, 8192, device="cuda", dtype=torch.float32)
    t_mm = timed(lambda: torch.mm(a, b), 5)
    gemm_tf = 2 * 8192 ** 3 / t_mm / 1e12
    del a, b
    torch.cuda.empty_cache()
    return {"copy_GB/s": round(copy_gbs, 1), "fp32_gemm_TFLOP/s": round(gemm_tf, 1),
            "cache_copy_GB/s": cache, "cache_ceiling_GB/s": max(cache.values()),
            "copy": "2 GiB device-to-device torch copy_, read + written bytes",
            "cache_copy": "device-to-device copies of 8 / 32 / 64 MiB repeated 200 times (resident in L2 or the "
                          "Infinity Cache), read + written bytes; the fastest is the cache-level ceiling",
            "gemm": "torch.mm 8192x8192x8192 float32"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "the library and this benchmark run on one GPU; multi-GPU sharding is designed but not built "
                          "(probe z-slabs per GPU with an atlas all-gather), so --gpus must be 1"}))
        sys.exit(2)
    import torch

    from vulkanexp_amd import IrradianceProbes, scenes
    from vulkanexp_amd import _lib

    torch.cuda.set_device(0)
    scene, grid = scenes.config_large(seed=2)
    light = scenes.sun_light()
    sky = scenes.SKY_RADIANCE
    n = int(grid["rays_per_probe"])
    probes = IrradianceProbes(grid, max_rays_per_probe=n)
    t0 = time.perf_counter()
    probes.set_scene(scene)
    bvh_s = time.perf_counter() - t0
    probes.set_light(light)
    probes.set_sky(sky)
    Rs = scenes.orientations(1234, args.warmup + args.steps + 2)
    probes.init_probes(Rs[0])
    flags = _lib.UPDATE_ALL_PROBES
    blend_flag = {"fp32": 0, "bf16": _lib.UPDATE_BLEND_BF16, "fp16": _lib.UPDATE_BLEND_FP16}[args.blend]
    flags |= blend_flag
    P = probes.get_probe_count()
    stream = torch.cuda.current_stream()

    for k in range(args.warmup):
        probes.update(Rs[1 + k], flags, stream)
    torch.cuda.synchronize()
    # timed region: the updates enqueued back to back, as a renderer submits frames (no timing
    # events, no host wait between updates)
    t0 = time.perf_counter()
    for k in range(args.steps):
        probes.update(Rs[1 + args.warmup + k], flags | _lib.UPDATE_NO_TIMING, stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = P * n / (ms_per_step * 1e-3) / 1e6
    # per-phase HIP-event times of the same updates, in a separate untimed pass
    phase = {"select_ms": [], "primary_ms": [], "shade_ms": [], "blend_ms": [], "total_ms": []}
    for k in range(args.steps):
        probes.update(Rs[1 + args.warmup + k], flags, stream)
        tm = probes.timings()  # waits for this update's events
        for key in phase:
            phase[key].append(tm[key])
    torch.cuda.synchronize()
    avg = {k: float(np.mean(v)) for k, v in phase.items()}

    # untimed counting update for the algorithmic bytes of the traversal kernels
    probes.update(Rs[-1], flags | _lib.UPDATE_COUNT_VISITS, stream)
    torch.cuda.synchronize()
    st = probes.trace_stats()
    rays = P * n
    prim_bytes = st["primary_nodes"] * _lib.NODE_BYTES + st["primary_tris"] * _lib.TRIANGLE_BYTES + rays * 16
    shade_bytes = (st["shadow_nodes"] * _lib.NODE_BYTES + st["shadow_tris"] * _lib.TRIANGLE_BYTES
                   + st["shadow_rays"] * (2 * 640 + 3 * 32 + 48 + 48) + rays * 32)
    blend_bytes = P * n * 16 + P * (36 * 12 + 196 * 8) * 2
    blend_flops = 2 * P * n * (36 * 3 + 196 * 2)
    kernels = {
        "trace_primary": {"ms": avg["primary_ms"], "algorithmic_bytes": prim_bytes},
        "shade_shadow": {"ms": avg["shade_ms"], "algorithmic_bytes": shade_bytes},
        "blend": {"ms": avg["blend_ms"], "algorithmic_bytes": blend_bytes, "flops": blend_flops},
    }
    dom = max(kernels, key=lambda k: kernels[k]["ms"])
    dk = kernels[dom]
    achieved = dk["algorithmic_bytes"] / (dk["ms"] * 1e-3) / 1e9
    traffic = profiled_traffic(dom)
    roofline = {"kernel": dom, "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                "traffic_source": os.path.relpath(PROFILE_SUMMARY, ROOT),
                "algorithmic_bytes_per_launch": int(dk["algorithmic_bytes"]), "avg_ms": round(dk["ms"], 4),
                "algorithmic_bytes_definition": "per primary ray: 128 B per 4-wide node fetched + 48 B per triangle "
                                                "fetched (the traversal's own fetch counters) + the 16 B hit record"}
    for k, v in kernels.items():
        v["algorithmic_GB/s"] = round(v["algorithmic_bytes"] / (v["ms"] * 1e-3) / 1e9, 1)
        t = profiled_traffic(k)
        if t:
            v["traffic_bytes"] = t
            v["traffic_GB/s"] = round(t / (v["ms"] * 1e-3) / 1e9, 1)
        v["ms"] = round(v["ms"], 4)
    kernels["blend"]["TFLOP/s"] = round(blend_flops / (avg["blend_ms"] * 1e-3) / 1e12, 2)
    peaks = None if args.no_peaks else achievable_peaks(torch)
    if traffic:
        # what leaves L2 (rocprofv3 FETCH_SIZE + WRITE_SIZE, MALL hits included) is the HBM-side load;
        # the algorithmic bytes above are mostly served by L2 and the Infinity Cache
        tr_gbs = traffic / (dk["ms"] * 1e-3) / 1e9
        roofline["traffic_GB/s"] = round(tr_gbs, 1)
        roofline["traffic_frac_of_peak"] = round(tr_gbs / HBM_PEAK_GBS, 4)
        roofline["fetched_bytes_beyond_L2"] = round(traffic / dk["algorithmic_bytes"], 4)
    if dom == "trace_primary":
        roofline["limiter"] = ("latency of dependent node fetches under lane divergence: a wavefront runs until its "
                               "longest ray ends (VALU lane utilisation in the profile summary), not HBM bandwidth")
        roofline["frac_note"] = ("achieved counts every lane's node and triangle fetches, most of them served by L1, L2 "
                                 "and the Infinity Cache, so frac can pass 1; traffic_frac_of_peak is the HBM share")
    if peaks:
        roofline["cache_ceiling"] = peaks["cache_ceiling_GB/s"]
        roofline["frac_of_cache_ceiling"] = round(achieved / peaks["cache_ceiling_GB/s"], 4)
        if traffic:
            roofline["traffic_frac_of_achievable_hbm"] = round(roofline["traffic_GB/s"] / peaks["copy_GB/s"], 4)
        kernels["blend"]["frac_of_fp32_peak"] = round(kernels["blend"]["TFLOP/s"] / FP32_PEAK_TFLOPS, 4)
        kernels["blend"]["frac_of_fp32_gemm"] = round(kernels["blend"]["TFLOP/s"] / peaks["fp32_gemm_TFLOP/s"], 4)
    trace_stats = {"nodes_per_primary_ray": st["primary_nodes"] / rays, "tris_per_primary_ray": st["primary_tris"] / rays,
                   "shadow_rays": st["shadow_rays"], "nodes_per_shadow_ray": st["shadow_nodes"] / max(1, st["shadow_rays"])}

    frame = None if args.no_frame else frame_bench(probes, scene)
    blend = None if args.no_blend_compare else blend_compare(probes, Rs[-1], _lib)
    config3, config3_check = (None, None) if args.no_blend_compare else blend_config3(_lib)
    diff = None if args.no_diff else diff_bench(_lib)
    batched = None if args.no_batch else batch_bench(_lib)
    dynamic = None if args.no_dynamic else dynamic_bench(probes, scene, Rs[-1], _lib)
    atmosphere = None if args.no_sky else sky_bench(probes, Rs[-1], _lib)

    if args.calibrate_copy:
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.add(src, 0.0, out=dst)
        torch.cuda.synchronize()

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(scene, grid, light, sky.astype(np.float64), Rs[1], args.cpu_probes, args.cpu_threads)
        if config3 is not None:  # the oracle's FP64 blend as the checker of the config-3 blends
            for mode, err in config3_errors(config3_check).items():
                config3[mode]["rel_err_vs_fp64"] = err

    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "Mrays/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32" if args.blend == "fp32" else f"fp32 trace, {args.blend} blend",
        "data": "synthetic (seeded procedural scene)",
        "config": {"workload": "plain DDGI large: 262,144-tri clustered scene, 32x16x32 probes x 256 rays, "
                               "8x8/16x16 FP32 atlases, all probes per update",
                   "probes": P, "rays_per_probe": n, "triangles": scene.triangle_count,
                   "bvh": probes.bvh_info(), "bvh_build_s": round(bvh_s, 3)},
        "phases_ms": {k: round(v, 4) for k, v in avg.items()},
        "phases_ms_median": {k: round(float(np.median(v)), 4) for k, v in phase.items()},
        "kernels": kernels, "trace_stats": trace_stats,
        "roofline": roofline, "cpu_baseline": cpu,
        "full_frame": frame, "blend_precision": blend, "blend_config3": config3, "differentiable": diff, "batched_volumes": batched,
        "dynamic_scene": dynamic, "atmospheric_sky": atmosphere, "achievable_peaks": peaks,
    }
    if cpu:
        out["speedup_vs_cpu"] = round(value / cpu["value"], 1)
    probes.destroy()
    print(json.dumps(out))


if __name__ == "__main__":
    main()
