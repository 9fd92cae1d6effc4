#!/usr/bin/env python3
"""Benchmark: paths/s of the wavefront path tracer at 1080p on one MI355X.

Workload (the reference's plain-method procedural-spheres configuration):
scenes/weekend_cover.scene (484 spheres: ground r=1000, three r=1 spheres,
480 seeded small spheres; Lambertian/GGX metal/dielectric), rendered at
1920x1080, 256 spp, max depth 50. A step is one batch of --passes-per-step
samples of every pixel (16 by default); --steps 16 covers the full 256 spp
render, each step adding to the previous ones' sums as a progressive
render does. Scene data is resident in HBM before the timed region.

The JSON line also carries
  roofline      the dominant stage kernel, its algorithmic bytes per launch
                (rays x fixed ray/hit bytes + BVH node bytes + primitive
                bytes, counted by the instrumented kernels on the same
                paths) over its HIP-event-timed average launch duration,
                and its measured HBM traffic per launch (profiles/);
  cpu_baseline  the reference renderer itself (oracle/_ref, compiled from the
                reference's sources with its Release flags) timed on this
                host's cores on a bounded sample of the same workload (and
                the whole Cornell workload), or the oracle when the reference
                build is absent;
  c1            the Cornell box, 512x512, 64 spp, depth 16 (GPU, and the
                reference's time for it in cpu_baseline.c1);
  c3            the batch of 8 generated terrain + instance scenes at uneven
                resolutions, 32 spp, depth 12, rendered in one batch of
                launches (mpt_render_batch), with its own roofline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "paths/s at 1080p; equal-time relMSE vs 4096spp ref (learned vs plain)"
HBM_PEAK_GBS = 8000.0
# per extension ray: queue index 4 + ray 48 read, class-queue append 4 (the 64-byte hit record
# written for hits is not counted)
EXT_RAY_BYTES = 56
# per shadow ray: queue 4 + origin 24 + to_light 24 + key 8 + f 24 + weight 8 + beta 24 + radiance 24 read,
# radiance 24 written
SHADOW_RAY_BYTES = 164


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def synchronize():
    import mypbr_amd as m
    m.synchronize()


def cpu_baseline(scene_path, cam, seconds, threads, c1_path=None, c1_cam=None):
    """Times the reference renderer (its Release-flags build, oracle/_ref) on
    this host's cores: a bounded sample of the spheres workload (the full
    1080p frame at a few spp, chosen to take about `seconds`) and, given
    c1_path, the whole Cornell workload. Falls back to the oracle when no
    reference build is present."""
    import numpy as np
    import oracle_lib as ol
    w, h = cam.image_width, cam.height()
    cpu = ol.cpu_description()
    tr = ol.timing_reference()
    if tr is not None or ol.reference_available():
        lib, march, flags = tr if tr is not None else (ol.reference(), "x86-64", "-O3 -ffp-contract=off")
        r = ol.RefScene(scene_path, lib)
        rad = np.zeros((h, w, 3))
        small = np.zeros((h // 8 + 1, w // 8 + 1, 3))
        r.lib.ref_render(r.h, max(1, w // 8), 1, cam.max_depth, threads, ol._d(small))  # builds the GGX table once
        t0 = time.time()
        r.lib.ref_render(r.h, w, 1, cam.max_depth, threads, ol._d(rad))
        one = time.time() - t0
        spp = max(1, min(64, int(seconds / max(one, 1e-3))))
        t0 = time.time()
        r.lib.ref_render(r.h, w, spp, cam.max_depth, threads, ol._d(rad))
        dt = time.time() - t0
        out = {"value": w * h * spp / dt, "unit": "paths/s", "cores": threads, "kind": "reference",
               "sample": f"reference camera::render, {w}x{h} at {spp} spp (same scene, depth {cam.max_depth})",
               "seconds": dt, "cpu": cpu["model"], "physical_cores": cpu["physical_cores"],
               "logical_cpus": cpu["logical_cpus"], "threads": threads,
               "compiler": "g++ 11.4 (the build container's), the reference's sources in place",
               "flags": flags}
        if c1_path:
            rc = ol.RefScene(c1_path, lib)
            w1, h1 = c1_cam.image_width, c1_cam.height()
            rad1 = np.zeros((h1, w1, 3))
            t0 = time.time()
            rc.lib.ref_render(rc.h, w1, c1_cam.samples_per_pixel, c1_cam.max_depth, threads, ol._d(rad1))
            dt1 = time.time() - t0
            out["c1"] = {"value": w1 * h1 * c1_cam.samples_per_pixel / dt1, "unit": "paths/s", "seconds": dt1,
                         "sample": f"reference camera::render, the whole Cornell workload: {w1}x{h1}, "
                                   f"{c1_cam.samples_per_pixel} spp, depth {c1_cam.max_depth}"}
        return out
    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy
    d = m.describe(scene_path)
    o = ol.OracleScene(d.desc)
    c1 = camera_copy(cam)
    t0 = time.time()
    o.render_sums(c1, seed=11, first_pass=0, passes=1, threads=threads)
    one = time.time() - t0
    spp = max(1, min(64, int(seconds / max(one, 1e-3))))
    t0 = time.time()
    o.render_sums(c1, seed=11, first_pass=0, passes=spp, threads=threads)
    dt = time.time() - t0
    return {"value": w * h * spp / dt, "unit": "paths/s", "cores": threads, "kind": "port",
            "sample": f"CPU oracle (restatement) {w}x{h} passes 0..{spp - 1} of the 256-spp camera", "seconds": dt,
            "cpu": cpu["model"], "physical_cores": cpu["physical_cores"], "threads": threads}


def cornell_c1(m, camera_copy, args):
    """The Cornell workload on the GPU: 512x512, 64 spp, depth 16, NEE + MIS,
    one whole render (after a one-pass warm-up), scene resident in HBM."""
    path = os.path.join(ROOT, "scenes", "cornell.scene")
    d = m.describe(path)
    cam = camera_copy(d.camera, image_width=512, samples_per_pixel=64, max_depth=16)
    s = m.Scene.from_desc(d.desc, d.camera)
    s.render(cam, seed=1, pass_count=1, read=False)
    synchronize()
    t0 = time.perf_counter()
    f = s.render(cam, seed=3, read=False)
    synchronize()
    dt = time.perf_counter() - t0
    st = f.stats
    out = {"workload": "Cornell box (18 quads, one area light), 512x512, 64 spp, max depth 16, NEE + MIS",
           "value": st.paths / dt, "unit": "paths/s", "seconds": dt, "paths": st.paths,
           "rays_per_s": (st.extension_rays + st.shadow_rays) / dt,
           "extension_rays_per_path": st.extension_rays / st.paths, "shadow_rays_per_path": st.shadow_rays / st.paths}
    log(f"cornell: {st.paths} paths in {dt:.3f} s = {st.paths / dt / 1e6:.1f} M paths/s")
    return out, path, cam


def _counter_rows(d):
    """(kernel, dispatch, counter, value) rows of the rocprofv3 counter CSVs under d."""
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(d, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                rows.append((r["Kernel_Name"], int(r["Dispatch_Id"]), r["Counter_Name"], float(r["Counter_Value"])))
    return rows


def traffic_summary(root):
    """Writes profiles/hbm_traffic.json from rocprofv3 --pmc runs under root:
    {c2,batch8}_{FETCH,WRITE}_SIZE (the workloads, one counter per run) and
    cal_{FETCH,WRITE}_SIZE (bench.py --copy-calibration 1024). Prints the
    table for the profiles/ summary."""
    import re
    scale = {}
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        vals = [v for _, _, name, v in _counter_rows(os.path.join(root, f"cal_{c}")) if name.startswith(c)]
        big = [v for v in vals if v > 0.5 * max(vals)]
        scale[c] = (1 << 30) / 1024.0 / (sum(big) / len(big))
    out = {"_sources_digest": kernel_sources_digest(),
           "_calibration": {"copy_bytes": 1 << 30, "fetch_scale": scale["FETCH_SIZE"],
                            "write_scale": scale["WRITE_SIZE"]}}
    lines = [f"Calibration: a 1 GiB device copy scales FETCH_SIZE by {scale['FETCH_SIZE']:.3f} and WRITE_SIZE "
             f"by {scale['WRITE_SIZE']:.3f}.", "", "| Workload | Kernel | Dispatches | FETCH_SIZE KiB/dispatch | "
             "WRITE_SIZE KiB/dispatch | HBM MB/dispatch (scaled) |", "|---|---|---|---|---|---|"]
    for wl, suffix in (("c2", ""), ("batch8", "_batch8")):
        per = {}
        for c in ("FETCH_SIZE", "WRITE_SIZE"):
            for kern, disp, name, v in _counter_rows(os.path.join(root, f"{wl}_{c}")):
                if name.startswith(c):
                    per.setdefault(kern, {}).setdefault(c, {})[disp] = per.get(kern, {}).get(c, {}).get(disp, 0) + v
        best = {}
        for kern, cs in per.items():
            if "FETCH_SIZE" not in cs or "WRITE_SIZE" not in cs:
                continue
            short = re.search(r"(k_[a-z_]+)", kern).group(1)
            f = sum(cs["FETCH_SIZE"].values()) / len(cs["FETCH_SIZE"])
            w = sum(cs["WRITE_SIZE"].values()) / len(cs["WRITE_SIZE"])
            rec = {"kernel": kern, "dispatches": len(cs["FETCH_SIZE"]), "fetch_kib_per_dispatch": f,
                   "write_kib_per_dispatch": w, "fetch_scale": scale["FETCH_SIZE"], "write_scale": scale["WRITE_SIZE"]}
            mb = 1024 * (f * scale["FETCH_SIZE"] + w * scale["WRITE_SIZE"]) / 1e6
            lines.append(f"| {wl} | `{kern}` | {rec['dispatches']} | {f:.0f} | {w:.0f} | {mb:.1f} |")
            if short + suffix not in best or f > best[short + suffix]["fetch_kib_per_dispatch"]:
                best[short + suffix] = rec
        out.update(best)
    with open(os.path.join(ROOT, "profiles", "hbm_traffic.json"), "w") as f:
        json.dump(out, f, indent=1)
    print("\n".join(lines))


def copy_calibration(mib):
    import torch
    n = mib << 20
    src = torch.ones(n, dtype=torch.uint8, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(4):
        dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "dispatches": 4}), flush=True)


# the sources whose kernels the committed counter summaries measured: a summary taken on other sources is stale
KERNEL_SOURCES = ["kernels.hip", "dev_trace.hpp", "dev_shade.hpp", "microfacet.hpp", "geom.hpp", "rng.hpp",
                  "scene_host.hpp"]


def kernel_sources_digest():
    import hashlib
    h = hashlib.sha256()
    for name in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "mypbr_amd", "csrc", name), "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def traffic_per_launch(kernel):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 counter
    summary (profiles/hbm_traffic.json): FETCH_SIZE + WRITE_SIZE (KiB per
    dispatch, averaged over the kernel's dispatches), each scaled by the ratio
    of the known bytes of a plain copy kernel to what the same counter reads
    for it. None when the summary was taken on other kernel sources."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            table = json.load(f)
        t = table[kernel]
    except (OSError, KeyError, ValueError):
        return None
    if table.get("_sources_digest") != kernel_sources_digest():
        return None
    return 1024.0 * (t["fetch_kib_per_dispatch"] * t["fetch_scale"] + t["write_kib_per_dispatch"] * t["write_scale"])


def roofline(kernel, bytes_per_launch, avg_s, traffic_key):
    """The roofline object of one kernel: algorithmic bytes per launch over
    its HIP-event-timed average launch, against the HBM peak, and the same
    for the measured HBM bytes (traffic). An algorithmic fraction above 1
    means the data it counts is served from caches, not HBM."""
    achieved = bytes_per_launch / avg_s / 1e9 if avg_s > 0 else 0.0
    traffic = traffic_per_launch(traffic_key)
    r = {"kernel": kernel, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
         "frac": achieved / HBM_PEAK_GBS, "traffic": traffic, "bytes_per_launch": bytes_per_launch,
         "avg_launch_us": avg_s * 1e6}
    if traffic is not None and avg_s > 0:
        r["hbm_achieved"] = traffic / avg_s / 1e9
        r["hbm_frac"] = r["hbm_achieved"] / HBM_PEAK_GBS
    return r


def batch8(args, built=None):
    """The batched workload: 8 generated scenes (2.09 M-triangle fBm terrain,
    500 affine icosphere instances, sun + sky environment, fog) at uneven
    resolutions, 32 spp, max depth 12, every scene resident in HBM. By default
    one mpt_render_batch call renders all 8: every stage launch serves the
    paths of every scene. --batch-mode sequential renders them back to back
    instead (one render per scene). Returns the JSON object of the workload."""
    import mypbr_amd as m
    from mypbr_amd import workloads as W
    scenes = []
    t0 = time.perf_counter()
    for b, cam in W.batch(seed=1):
        d, keep = b.build()
        scenes.append((m.Scene.from_desc(d, cam, keepalive=keep), cam))
    log(f"built 8 scenes in {time.perf_counter() - t0:.1f} s")
    ss = [s for s, _ in scenes]
    cams = [c for _, c in scenes]
    one = args.batch_mode == "one-launch"
    # warm-up: one pass of every scene
    if one:
        m.render_batch(ss, cams, seed=1, pass_count=1, read=False)
    else:
        for s, cam in scenes:
            s.render(cam, seed=1, pass_count=1, read=False)
    synchronize()
    t0 = time.perf_counter()
    per = []
    if one:
        _, st = m.render_batch(ss, cams, seed=2, read=False)
        paths, ext, sh = st.paths, st.extension_rays, st.shadow_rays
    else:
        paths = ext = sh = 0
        for s, cam in scenes:
            t1 = time.perf_counter()
            f = s.render(cam, seed=2, read=False)
            per.append({"width": cam.image_width, "height": cam.height(), "seconds": time.perf_counter() - t1})
            paths += f.stats.paths
            ext += f.stats.extension_rays
            sh += f.stats.shadow_rays
    synchronize()
    dt = time.perf_counter() - t0
    log(f"batch ({args.batch_mode}): {paths} paths in {dt:.3f} s = {paths / dt / 1e6:.1f} M paths/s, "
        f"{(ext + sh) / dt / 1e6:.1f} M rays/s")
    out = {"metric": "paths/s, batch of 8 generated scenes", "value": paths / dt, "unit": "paths/s",
           "seconds": dt, "mode": args.batch_mode, "paths": paths,
           "workload": "8 scenes: fBm terrain 1024^2 grid (2,093,058 tris) + 500 icosphere-4 instances, principled "
                       "GGX, sun+sky 2048x1024 environment, fog 0.02; 256^2 .. 1920x1080, 32 spp, max depth 12",
           "rays_per_s": (ext + sh) / dt, "extension_rays": ext, "shadow_rays": sh}
    if per:
        out["scenes"] = per
    if not args.no_profile and one:
        # 4 of the 32 passes again, with per-stage event timing, then with BVH work counters
        _, ft = m.render_batch(ss, cams, seed=1, pass_count=4, profile=1, read=False, tail_paths=-1)
        _, fc = m.render_batch(ss, cams, seed=1, pass_count=4, profile=2, read=False, tail_paths=-1)
        names = ["extend", "shade", "shadow", "camera+accumulate", "finish"]
        stage = {names[i]: {"ms": ft.stage_ms[i], "launches": ft.stage_launches[i]} for i in range(5)}
        ext_bytes = fc.extension_rays * EXT_RAY_BYTES + fc.node_bytes[0] + fc.prim_bytes[0]
        sh_bytes = fc.shadow_rays * SHADOW_RAY_BYTES + fc.node_bytes[1] + fc.prim_bytes[1]
        stage["extend"].update(algorithmic_bytes=ext_bytes, rays=fc.extension_rays,
                               node_bytes_per_ray=fc.node_bytes[0] / max(fc.extension_rays, 1),
                               prim_bytes_per_ray=fc.prim_bytes[0] / max(fc.extension_rays, 1))
        stage["shadow"].update(algorithmic_bytes=sh_bytes, rays=fc.shadow_rays,
                               node_bytes_per_ray=fc.node_bytes[1] / max(fc.shadow_rays, 1),
                               prim_bytes_per_ray=fc.prim_bytes[1] / max(fc.shadow_rays, 1))
        out["stages"] = stage
        log("batch stages:", json.dumps(stage))
        roof = {}
        for k in ("extend", "shadow"):
            launches = max(stage[k]["launches"], 1)
            roof[k] = roofline(f"k_{k} (batch of 8 scenes, 4 passes)", stage[k]["algorithmic_bytes"] / launches,
                               stage[k]["ms"] / launches / 1e3, f"k_{k}_batch8")
        out["roofline"] = roof["extend"]
        out["roofline_shadow"] = roof["shadow"]
    return out


def display_bench(m, radiance, reps=50):
    """The display pipeline at the frame's size, device-resident: bloom 0.25,
    vignette 0.25, ACES (the reference's voxel-style preset, scene.h:631-635)
    to sRGB bytes, and postprocess alone. Wall time per frame over `reps`."""
    import ctypes as C
    import numpy as np
    h, w = radiance.shape[:2]
    d = m.display_settings(tonemap="aces", bloom=0.25, vignette=0.25)
    src = m.DeviceArray.of(radiance)
    ws = m.DeviceArray((max(m.lib.mpt_display_workspace_bytes(w, h, C.byref(d)), 1),), np.uint8)
    rgb = m.DeviceArray((h, w, 3), np.uint8)
    out = m.DeviceArray((h, w, 3), np.float64)
    res = {"width": w, "height": h, "settings": "bloom 0.25, vignette 0.25, tonemap aces, exposure 0"}
    for name, call in (("to_bytes", lambda: m.lib.mpt_display_bytes_dev(src.handle, w, h, C.byref(d), 1, rgb.handle,
                                                                          ws.handle, None)),
                       ("postprocess", lambda: m.lib.mpt_postprocess_dev(src.handle, w, h, C.byref(d), out.handle,
                                                                         ws.handle, None))):
        for _ in range(3):
            m._check(call())
        synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            m._check(call())
        synchronize()
        dt = (time.perf_counter() - t0) / reps
        # algorithmic bytes: the radiance read once and the output written once
        nbytes = w * h * (24 + (3 if name == "to_bytes" else 24))
        res[name] = {"ms_per_frame": dt * 1e3, "mpix_per_s": w * h / dt / 1e6,
                     "algorithmic_bytes": nbytes, "achieved_gbs": nbytes / dt / 1e9}
    log("display:", json.dumps(res))
    return res


def network_bench(m, s, cam, reps=20):
    """The learned denoiser's network at 1080p on the matrix cores: the
    six-layer, 64-channel KPCNN's forward pass (implicit-GEMM 3x3
    convolutions, k_conv3x3) on a frame already in HBM, and one training
    step (forward, backward, Adam) on a batch of 16 crops of 256^2. FLOPs
    are the convolutions' multiply-adds x 2 as computed (padded channels
    included); the kernel application is not counted."""
    from mypbr_amd.learn import KPCNN
    from mypbr_amd.abi import camera_copy
    import ctypes as C
    import numpy as np
    c = camera_copy(cam, samples_per_pixel=4, collect_features=1)
    s.render(c, seed=5, read=False)
    dev = s.frame_dev(c)
    net = KPCNN(layers=6, hidden=64, seed=1)
    cin, cout, _, _ = net.layout()
    h, w = c.height(), c.image_width
    flops = sum(2.0 * w * h * 9 * ci * co for ci, co in zip(cin, cout))
    out = m.DeviceArray((h, w, 3), np.float32)
    loss = C.c_double(0)
    vp = lambda t: t.handle if t is not None else None  # noqa: E731
    run = lambda: m._check(m.lib.mpt_kpcnn_run(net._h, *[vp(dev[k]) for k in ("radiance", "albedo", "normal",  # noqa: E731
                                                                              "variance", "depth")],
                                                1, h, w, None, 0.0, out.handle, C.byref(loss)))
    for _ in range(3):
        run()
    synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        run()
    synchronize()
    dt = (time.perf_counter() - t0) / reps
    # one training step on a batch of 16 crops of 256^2 (synthetic frames): forward, backward, Adam
    from mypbr_amd.learn import to_device
    rng = np.random.default_rng(3)
    B, ch = 16, 256
    frames = [dict(radiance=rng.uniform(0, 2, (ch, ch, 3)), albedo=rng.uniform(0.05, 0.95, (ch, ch, 3)),
                   normal=rng.uniform(-1, 1, (ch, ch, 3)), variance=rng.exponential(0.05, (ch, ch)),
                   depth=rng.uniform(0, 20, (ch, ch))) for _ in range(B)]
    tdev = to_device(frames)
    tgt = m.DeviceArray.of(rng.uniform(0, 2, (B, ch, ch, 3)), np.float32)
    for _ in range(2):
        net.run(tdev, target=tgt, lr=1e-4)
    t0 = time.perf_counter()
    for _ in range(reps):
        net.run(tdev, target=tgt, lr=1e-4)
    step = (time.perf_counter() - t0) / reps
    train_flops = 3 * sum(2.0 * B * ch * ch * 9 * ci * co for ci, co in zip(cin, cout))
    res = {"network": "KPCNN 6 layers x 64 channels, 3x3 convolutions, %d parameters" % net.nparams,
           "width": w, "height": h, "forward_ms": dt * 1e3, "forward_flops": flops,
           "train_step_ms": step * 1e3, "train_step_batch": f"{B} x {ch}x{ch}",
           "train_step_tflops": train_flops / step / 1e12,
           "roofline": {"kernel": "k_conv3x3 (whole forward pass)", "bound": "mfma", "achieved": flops / dt / 1e12,
                        "peak": 2500.0, "unit": "TFLOP/s", "frac": flops / dt / 1e12 / 2500.0, "traffic": None}}
    log("network:", json.dumps(res))
    return res


def main():
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--passes-per-step", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-profile", action="store_true")
    ap.add_argument("--no-c1", action="store_true", help="skip the Cornell 512x512, 64 spp, depth 16 line")
    ap.add_argument("--no-c3", action="store_true",
                    help="skip the batch of 8 generated terrain + instance scenes (one batch of launches)")
    ap.add_argument("--workload", choices=["spheres", "batch8"], default="spheres",
                    help="spheres: the 1080p procedural-spheres render (the headline metric); batch8: the "
                         "8 generated terrain + instance scenes of uneven resolutions, 32 spp, depth 12")
    ap.add_argument("--batch-mode", choices=["one-launch", "sequential"], default="one-launch",
                    help="batch8: all 8 scenes in one batch of launches (default), or one render per scene")
    ap.add_argument("--no-display", action="store_true",
                    help="skip timing the display pipeline (bloom, vignette, tonemap, sRGB bytes) at 1080p")
    ap.add_argument("--no-network", action="store_true", help="skip the learned denoiser's network timing")
    ap.add_argument("--no-quality", action="store_true",
                    help="skip the equal-time relMSE study of the learned forms against the plain tracer")
    ap.add_argument("--use-graph", type=int, default=0, choices=[-1, 0, 1],
                    help="bounce kernels replayed as hipGraphs (1), launched directly (-1), or the tuned choice (0)")
    ap.add_argument("--sort", type=int, default=0, choices=[-1, 0, 1],
                    help="bin hits by material into per-material shade kernels (1), one kernel per batch (-1), auto (0)")
    ap.add_argument("--bpc", type=int, default=0, help="persistent grid blocks per CU (0 = auto)")
    ap.add_argument("--tail-paths", type=int, default=0,
                    help="live paths at which a batch's tail runs as one launch (0 = the library's default, -1 = never)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device to device 4 times (a plain copy kernel of known size, "
                         "the scale for rocprofv3 FETCH_SIZE/WRITE_SIZE) and exit")
    ap.add_argument("--traffic-summary", metavar="DIR", default="",
                    help="only write profiles/hbm_traffic.json from the rocprofv3 --pmc runs under DIR and exit")
    args = ap.parse_args()
    if args.traffic_summary:
        traffic_summary(args.traffic_summary)
        return
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported in this build "
              "(paths shard by pass or tile with one sum-reduce; see the design notes)", file=sys.stderr)
        sys.exit(2)

    import mypbr_amd as m
    from mypbr_amd.abi import camera_copy

    if args.workload == "batch8":
        b = batch8(args)
        line = {"metric": b.pop("metric"), "value": b.pop("value"), "unit": "paths/s", "n_gpus": 1, "steps": 1,
                "warmup": 1, "ms_per_step": b["seconds"] * 1e3, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "f64", "data": "synthetic: mypbr_amd/workloads.py, seed 1",
                "config": {"workload": b.pop("workload"), "mode": b["mode"]}}
        line.update(b)
        print(json.dumps(line), flush=True)
        return

    scene_path = os.path.join(ROOT, "scenes", "weekend_cover.scene")
    d = m.describe(scene_path)
    cam = camera_copy(d.camera, image_width=1920, samples_per_pixel=256, max_depth=50)
    assert cam.height() == 1080
    s = m.Scene.from_desc(d.desc, d.camera)
    info = m.device_info()
    log("device:", info)
    pps = args.passes_per_step
    total_passes = args.steps * pps
    sched = dict(use_graph=args.use_graph, sort_by_material=args.sort, blocks_per_cu=args.bpc)

    for k in range(args.warmup):
        s.render(cam, seed=1, first_pass=(k * pps) % 256, pass_count=pps, read=False, tail_paths=args.tail_paths, **sched)
    synchronize()
    t0 = time.perf_counter()
    ext = sh = 0
    for k in range(args.steps):
        f = s.render(cam, seed=1, first_pass=(k * pps) % 256, pass_count=pps, accumulate=k > 0, read=False, tail_paths=args.tail_paths, **sched)
        ext += f.stats.extension_rays
        sh += f.stats.shadow_rays
        log(f"step {k + 1}/{args.steps}: {f.stats.seconds * 1e3:.1f} ms")
    synchronize()
    dt = time.perf_counter() - t0
    paths = cam.image_width * cam.height() * total_passes
    value = paths / dt
    log(f"{paths} paths in {dt:.3f} s = {value / 1e6:.1f} M paths/s, {(ext + sh) / dt / 1e9:.2f} G rays/s")

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "paths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic: scenes/weekend_cover.scene, the reference's seeded procedural-spheres scene",
        "config": {"workload": "procedural spheres (484), 1920x1080, 256 spp, max depth 50, plain wavefront tracer",
                   "width": 1920, "height": 1080, "spp": 256, "max_depth": 50,
                   "passes_per_step": pps, "passes_timed": total_passes},
        "rays_per_s": (ext + sh) / dt,
        "extension_rays_per_path": ext / paths,
        "shadow_rays_per_path": sh / paths,
        "device": info["name"],
    }

    if not args.no_profile:
        # one step again with per-stage event timing, then with BVH work counters
        ft = s.render(cam, seed=1, first_pass=0, pass_count=pps, profile=1, read=False, tail_paths=args.tail_paths, **sched).stats
        fc = s.render(cam, seed=1, first_pass=0, pass_count=pps, profile=2, read=False, tail_paths=args.tail_paths, **sched).stats
        names = ["extend", "shade", "shadow", "camera+accumulate", "finish"]
        stage = {names[i]: {"ms": ft.stage_ms[i], "launches": ft.stage_launches[i]} for i in range(5)}
        # rays of the stage kernels only: the batch tails (k_finish) neither count their BVH work
        # nor are timed as part of k_extend / k_shadow
        ext_rays = fc.extension_rays - fc.tail_rays[0]
        sh_rays = fc.shadow_rays - fc.tail_rays[1]
        ext_bytes = ext_rays * EXT_RAY_BYTES + fc.node_bytes[0] + fc.prim_bytes[0]
        sh_bytes = sh_rays * SHADOW_RAY_BYTES + fc.node_bytes[1] + fc.prim_bytes[1]
        stage["extend"]["algorithmic_bytes"] = ext_bytes
        stage["shadow"]["algorithmic_bytes"] = sh_bytes
        stage["extend"]["rays"] = ext_rays
        stage["extend"]["node_bytes_per_ray"] = fc.node_bytes[0] / max(ext_rays, 1)
        stage["extend"]["prim_bytes_per_ray"] = fc.prim_bytes[0] / max(ext_rays, 1)
        line["stages"] = stage
        dom = max(("extend", "shadow"), key=lambda k: stage[k]["ms"])
        launches = stage[dom]["launches"]
        line["roofline"] = roofline(f"k_{dom}", stage[dom]["algorithmic_bytes"] / launches,
                                    stage[dom]["ms"] / launches / 1e3, f"k_{dom}")
        line["roofline"]["note"] = ("algorithmic bytes count every BVH node and sphere read; the 484-sphere tree "
                                    "stays in the caches, so frac can approach or pass 1 while hbm_frac, from the "
                                    "measured HBM bytes, stays low")
        log("stages:", json.dumps(stage))

    c1 = c1_path = c1_cam = None
    if not args.no_c1:
        c1, c1_path, c1_cam = cornell_c1(m, camera_copy, args)
        line["c1"] = c1
    if not args.no_c3:
        line["c3"] = batch8(args)

    if not args.no_display:
        f = s.render(cam, seed=7, first_pass=0, pass_count=4)
        line["display"] = display_bench(m, f.radiance)

    if not args.no_network:
        line["network"] = network_bench(m, s, cam)

    if not args.no_quality:
        # the metric's second half: equal-time relMSE (eps 0.01) against a 4096-spp render of the plain tracer
        from mypbr_amd.quality import equal_time_study
        line["quality"] = equal_time_study(scene_path, box=((-14, -1, -14), (14, 4, 14)), log=log)

    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        cb = cpu_baseline(scene_path, cam, args.cpu_seconds, threads, c1_path, c1_cam)
        # the same integrator traces the same rays per path in expectation: CPU rays/s from the GPU's count
        cb["rays_per_s"] = cb["value"] * line["rays_per_s"] / value
        cb["rays_per_s_note"] = "paths/s times the rays per path the GPU counted on the same workload"
        line["cpu_baseline"] = cb
        line["gpu_over_cpu"] = value / cb["value"]
        if c1 is not None and "c1" in cb:
            cb["c1"]["rays_per_s"] = cb["c1"]["value"] * c1["rays_per_s"] / c1["value"]
            c1["gpu_over_cpu"] = c1["value"] / cb["c1"]["value"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
