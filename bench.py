#!/usr/bin/env python3
"""Throughput of the per-sample radiance path on one MI355X.

A step renders one full frame of the workload (every pixel, every sample per
pixel) with scene, photon map and frame accumulator resident in HBM.
`value` = samples / s of the wavefront renderer over exactly K timed steps;
the plain megakernel is timed on the same frames beside it.

Default workload: the QBVH-bound mesh scene (a 1,310,720-triangle displaced
icosphere in rough GGX glass under a 2048x1024 HDR environment, 1280x720,
64 spp, depth 12).  The gem-with-caustics scene (1280x720, 256 spp, depth 24,
2M photons), the uneven batch (36 gem spin stills of mixed resolution plus
a heterogeneous-volume glass box, one launch) and the adjoint (gradient
evaluations/s of a 20-parameter L2 image loss on a 512x512 rough gem, with a
finite-difference check) are measured after it and reported under "also",
with the RMSE of 64-spp gem frames against a 4096-spp host image ("at fixed
RMSE": GPU and host estimators agree, so samples/s compare at equal spp).

Beside the throughput, on the same line:
  roofline     — algorithmic HBM bytes of the closest-hit traversal kernel
                 (k_wf_extend) per launch — BVH node records x 256 B +
                 primitive records, counted by an untimed instrumented pass —
                 over its average launch time from HIP events on the library's
                 stream, against 8 TB/s;
  cpu_baseline — the CPU restatement of render_tiles (oracle/, the
                 reference's row-sequential SmallRng streams) on this host's
                 cores, on a bounded sample of the same frame.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# bytes of one primitive record read by a test: triangle v0/e1/e2 48 B,
# quad 80 B, convex gem 57 half-spaces x 16 B
PRIM_BYTES = {"box": 80.0, "mesh": 48.0, "gem": 16.0 * 57}


BENCH_TOML = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "bench_scene", "bench.toml")


def reference_bench_scene():
    """The reference's own --bench workload (run_bench, src/main.rs:84-117):
    scenes/bench.toml (room, OBJ mesh, [[gems]] diamond and sapphire, dusk
    sun, blackbody panel) at 1200 x 800, 128 samples, depth 50 (ray_color's
    MAX_DEPTH), the photon map of rebuild_caustics."""
    from rustracer_amd import scene_file
    scene, camera, info = scene_file.load(BENCH_TOML)
    W, H = 1200, 800
    return scene, camera(W / H), dict(width=W, height=H, spp=128, max_depth=50, caustics=info)


def gem_workload():
    """The dispersive-gem caustic workload as configured: Sellmeier diamond,
    2M-photon k = 64 map, 1280x720, 256 spp, depth 24, four hero wavelengths
    over [380, 780) nm (camera paths and photons alike)."""
    from rustracer_amd import scenes
    sc, cam, cfg = scenes.gem_caustics(1280, 720, **scenes.GEM_BENCH)
    cfg["wavelengths"] = 4
    return sc, cam, cfg


def workloads():
    from rustracer_amd import scenes
    return {
        "box": lambda: scenes.box_room(256, 256),
        "mesh": lambda: scenes.mesh_envmap(1280, 720, level=8),
        "gem": gem_workload,
        "bench_toml": reference_bench_scene,
    }


def log(msg):
    print(msg, file=sys.stderr, flush=True)


# Path-queue bytes the wavefront stages move per ray (implementation overhead,
# reported beside the algorithmic bytes): extend reads the queue entry (4 B)
# and the ray of the 128-B path record (36 B) and writes the hit (12 B) and
# the bucket entry (4 B); shade reads the bucket entry and the record and
# writes the record and the next-queue entry; each shadow request and its
# 4-B entry in the request list are written by shade and read by shadow.
QUEUE_BYTES = {"extend": 4 + 36 + 12 + 4, "shade": 4 + 128 + 128 + 4, "shadow_request": 64 + 4}


def stage_figures(st, cnt, samples):
    """Per-stage figures of a timed wavefront pass (graphs off, one stream):
    average ms per launch of each stage, the path-queue bytes it moves (per
    sample and as GB/s over the stage's time) and, with a k-NN photon map,
    the k-NN stage's algorithmic bytes (photon records read by the reference
    algorithm's gathers, from the counting pass)."""
    ns = max(cnt["samples"], 1)
    rays, srays = cnt["rays"] / ns, cnt["shadow_rays"] / ns
    qb = {"extend": rays * QUEUE_BYTES["extend"],
          "shade": rays * QUEUE_BYTES["shade"] + srays * QUEUE_BYTES["shadow_request"],
          "shadow": srays * QUEUE_BYTES["shadow_request"]}
    out = {}
    for k, (ms, n) in st.items():
        if not n:
            continue
        f = {"avg_launch_ms": ms / n, "launches": n, "total_ms": ms}
        if k in qb:
            f["queue_bytes_per_sample"] = qb[k]
            f["queue_gbs"] = qb[k] * samples / (ms / 1e3) / 1e9
        if k == "knn":
            pbs = cnt["photon_bytes"] / ns
            bpl = pbs * samples / n
            f.update({"photon_bytes_per_sample": pbs, "bytes_per_launch": bpl,
                      "achieved_gbs": bpl / (ms / n / 1e3) / 1e9,
                      "note": "photon map is cache-resident (L2 / MALL); latency-bound walks, not an HBM figure"})
        out[k] = f
    return out


def tune_plain(r, step_plain, log_name):
    """Register budget of the plain megakernel (0: the compiler's 256 VGPRs, 2:
    two waves per SIMD), picked on one untimed frame each."""
    import torch
    best, best_ms = 0, None
    for occ in (0, 2):
        r.set_megakernel_occupancy(occ)
        step_plain(50)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step_plain(51)
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) * 1e3
        log(f"[{log_name}] plain megakernel, occupancy {occ}: {ms:.1f} ms per frame")
        if best_ms is None or ms < best_ms:
            best, best_ms = occ, ms
    r.set_megakernel_occupancy(best)
    return best


def measure(name, args, with_cpu):
    import numpy as np
    import torch
    from rustracer_amd import Renderer
    from rustracer_amd.autotune import tune_wavefront

    t0 = time.time()
    scene, cam, cfg = workloads()[name]()
    build_s = time.time() - t0
    W, H, spp, depth = cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"]
    wl = cfg.get("wavelengths", 0)
    strata = int(spp ** 0.5)
    r = Renderer(0)
    t0 = time.time()
    r.upload(scene)
    upload_s = time.time() - t0
    stats = r.stats()
    photons = "num_photons" in cfg or "caustics" in cfg
    ph = None
    if "caustics" in cfg:
        t0 = time.time()
        stored = r.rebuild_caustics(cfg["caustics"], wavelengths=wl)
        photons = stored > 0
        ph = {"traced": cfg["caustics"]["num_photons"], "stored": stored, "build_s": round(time.time() - t0, 3),
              "emitter": "rebuild_caustics"}
    elif photons:
        lq = cfg["light_quad"]
        t0 = time.time()
        stored = r.photon_build_from_quad(lq["q"], lq["u"], lq["v"], cfg["light_color"], cfg["num_photons"],
                                          cfg["gather_radius"], knn=cfg.get("knn", 0), wavelengths=wl)
        ph = {"traced": cfg["num_photons"], "stored": stored, "build_s": round(time.time() - t0, 3)}
    acc = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    log(f"[{name}] scene {build_s:.1f}s upload {upload_s:.1f}s, {stats['nodes']} nodes, {stats['triangles']} tris")

    tuned, tune_info = (json.loads(args.wf_opts) if args.wf_opts else {}), None
    if not args.no_tune and not args.wf_opts:
        tuned, tune_info = tune_wavefront(r, acc, cam, W, H, spp, depth, photons, knn=bool(cfg.get("knn", 0)),
                                          wavelengths=wl)
        log(f"[{name}] autotuned wavefront: {tuned} ({tune_info['seconds']}s)")

    def step_plain(k):
        r.render_accumulate(acc, cam, W, H, spp, strata=strata, max_depth=depth, seed=k, photon_map=photons,
                            wavelengths=wl)

    def step_wf(k, graph=None, streams=None, wavelengths=None):
        o = dict(tuned) if tuned else {}
        if graph is not None:
            o["use_graph"] = graph
        if streams is not None:
            o["streams"] = streams
        r.render_wavefront(acc, cam, W, H, spp, strata=strata, max_depth=depth, seed=k, photon_map=photons,
                           wavelengths=wl if wavelengths is None else wavelengths, **o)

    def timed(step):
        for k in range(args.warmup):
            step(k)
        torch.cuda.synchronize()
        r.set_timing(True)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            step(1000 + k)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        kms, nl = r.timing()
        r.set_timing(False)
        img = acc.cpu().numpy() / spp
        assert np.isfinite(img).all(), "non-finite radiance"
        frames.append(img)
        return el, kms, nl

    frames = []

    # counting pass (untimed): traversal work per sample (plain megakernel, same streams)
    r.set_counting(True)
    step_plain(0)
    torch.cuda.synchronize()
    cnt = r.counters()
    r.set_counting(False)

    samples = W * H * spp
    res = {}
    plain_occ = None
    if args.renderer in ("both", "plain"):
        plain_occ = tune_plain(r, step_plain, name)
        res["plain"] = timed(step_plain)
        log(f"[{name}] plain {samples * args.steps / res['plain'][0] / 1e6:.1f} M samples/s")
    if args.renderer in ("both", "wavefront"):
        res["wavefront"] = timed(step_wf)
        log(f"[{name}] wavefront {samples * args.steps / res['wavefront'][0] / 1e6:.1f} M samples/s")
    head = "wavefront" if "wavefront" in res else "plain"
    rates = {k: samples * args.steps / v[0] for k, v in res.items()}
    same_frame = None
    if len(frames) == 2:
        # both renderers drew the same per-pixel streams for the last timed
        # step: the frames agree up to the order of float additions
        same_frame = float(np.abs(frames[0] - frames[1]).max() / max(np.abs(frames[0]).max(), 1e-30))
        assert same_frame < 1e-3, f"wavefront frame differs from the megakernel frame ({same_frame})"

    ns = max(cnt["samples"], 1)
    pb = PRIM_BYTES.get(name, 64.0)
    closest_bps = (cnt["node_tests"] * 256.0 + cnt["prim_tests"] * pb) / ns
    shadow_bps = (cnt["shadow_node_tests"] * 256.0 + cnt["shadow_prim_tests"] * pb) / ns
    roof = {}
    if "plain" in res:
        el, kms, nl = res["plain"]
        spl = samples * args.steps / max(nl, 1)
        roof["plain"] = {"kernel": "k_render", "avg_launch_ms": kms / max(nl, 1),
                         "bytes_per_launch": (closest_bps + shadow_bps) * spl,
                         "achieved": (closest_bps + shadow_bps) * spl / (kms / 1e3 / max(nl, 1)) / 1e9}
    if "wavefront" in res:
        # per-launch HIP events around k_wf_extend: the non-graph launch path on
        # one stream (with two, a launch's events would also time the other
        # stream's kernels sharing the GPU)
        r.set_timing(True)
        nfr = max(1, args.steps)
        for k in range(nfr):
            step_wf(2000 + k, graph=False, streams=1)
        torch.cuda.synchronize()
        ems, en = r.stage_timing()
        stages = stage_figures(r.stage_times(), cnt, samples * nfr)
        r.set_timing(False)
        bpl = closest_bps * samples * nfr / max(en, 1)
        roof["wavefront"] = {"kernel": "k_wf_extend", "avg_launch_ms": ems / max(en, 1), "bytes_per_launch": bpl,
                             "achieved": bpl / (ems / 1e3 / max(en, 1)) / 1e9, "launches": en, "stages": stages}
    rh = roof[head]
    one_wl = None
    if wl and "wavefront" in res:
        # the reference's single wavelength over [380, 700) on the same map, beside it
        for k in range(max(1, args.warmup)):
            step_wf(3000 + k, wavelengths=0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for k in range(args.steps):
            step_wf(3100 + k, wavelengths=0)
        torch.cuda.synchronize()
        one_wl = samples * args.steps / (time.perf_counter() - t0)
        log(f"[{name}] wavefront, one wavelength over [380, 700): {one_wl / 1e6:.1f} M samples/s")

    cpu = cpu_baseline(scene, cam, cfg, args.cpu_seconds, photons) if with_cpu else None
    r.close()
    traffic, traffic_note = None, None
    if head == "wavefront":
        traffic, traffic_note = measured_traffic(scene.name, tuned, rh["bytes_per_launch"])
    el = res[head][0]
    return {
        "value": rates[head], "ms_per_step": el * 1e3 / args.steps,
        "config": {"workload": scene.name, "width": W, "height": H, "spp": spp, "max_depth": depth,
                   "renderer": head, "samples_per_s": rates,
                   "wavelengths": ("four hero wavelengths over [380, 780) nm" if wl == 4 else
                                   "one over [380, 780) nm" if wl == 1 else "one hero wavelength over [380, 700) nm"),
                   "wavefront_samples_per_s_one_wavelength_380_700": one_wl,
                   "wavefront_over_plain": (rates["wavefront"] / rates["plain"]) if len(rates) == 2 else None,
                   "plain_megakernel_occupancy": plain_occ,
                   "frame_max_rel_diff_wavefront_vs_plain": same_frame,
                   "wavefront_opts": tuned,
                   "autotune": tune_info and {k: tune_info[k] for k in ("probe_spp", "probes", "seconds")},
                   "bvh_nodes": stats["nodes"], "triangles": stats["triangles"], "upload_s": round(upload_s, 3),
                   "photons": ph},
        "roofline": {"bound": "hbm", "achieved": rh["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": rh["achieved"] / HBM_PEAK_GBS, "traffic": traffic, "traffic_source": traffic_note,
                     "kernel": rh["kernel"], "avg_launch_ms": rh["avg_launch_ms"], "bytes_per_launch": rh["bytes_per_launch"],
                     "closest_bytes_per_sample": closest_bps, "shadow_bytes_per_sample": shadow_bps,
                     "node_tests_per_sample": cnt["node_tests"] / ns, "prim_tests_per_sample": cnt["prim_tests"] / ns,
                     "rays_per_sample": cnt["rays"] / ns, "shadow_rays_per_sample": cnt["shadow_rays"] / ns,
                     "photon_bytes_per_sample": cnt["photon_bytes"] / ns,
                     "stages": roof.get("wavefront", {}).get("stages"),
                     "per_renderer": roof},
        "cpu_baseline": cpu,
    }


def measure_batch(args, with_cpu):
    """Uneven batch: 36 gem spin stills (10 degree steps, resolutions drawn
    from 256x256 / 512x384 / 960x540, 64 spp, depth 24, 2M-photon caustic map
    each) plus the 640x480 volume glass box (64 spp, depth 24), all in one
    rt_render_batch launch per step.  The gem stills draw four hero
    wavelengths over [380, 780) nm (the gem workload's mode, photons alike).
    The same views rendered one rt_render_wavefront call after another are
    timed beside it."""
    import numpy as np
    import torch
    from rustracer_amd import Renderer, scenes
    from rustracer_amd.autotune import tune_batch

    t0 = time.time()
    items = scenes.spin_batch()
    r = Renderer(0)
    views, photons = [], 0
    for sc, cam, cfg in items:
        h = r.upload_extra(sc)
        use_ph = "num_photons" in cfg
        if use_ph:
            lq = cfg["light_quad"]
            photons += r.photon_build_from_quad(lq["q"], lq["u"], lq["v"], cfg["light_color"], cfg["num_photons"],
                                                cfg["gather_radius"], scene=h, knn=cfg.get("knn", 0), wavelengths=4)
        W, H = cfg["width"], cfg["height"]
        views.append(dict(scene=h, camera=cam, width=W, height=H, spp=cfg["spp"], max_depth=cfg["max_depth"],
                          photon_map=use_ph, strata=int(cfg["spp"] ** 0.5), wavelengths=4 if use_ph else 0,
                          out=torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")))
    setup_s = time.time() - t0
    samples = sum(v["width"] * v["height"] * v["spp"] for v in views)
    log(f"[batch] {len(views)} views, {samples / 1e6:.1f} M samples, setup {setup_s:.1f}s")
    tuned, info = (json.loads(args.wf_opts) if args.wf_opts else {}), None
    if not args.no_tune and not args.wf_opts:
        tuned, info = tune_batch(r, views)
        log(f"[batch] autotuned: {tuned} ({info['seconds']}s)")

    def step_batch(k):
        for v in views:
            v["seed"] = k
        r.render_batch(views, **tuned)

    def step_seq(k):
        o = {key: val for key, val in tuned.items()}
        for v in views:
            r.render_wavefront(v["out"], v["camera"], v["width"], v["height"], v["spp"], strata=v["strata"],
                               max_depth=v["max_depth"], photon_map=v["photon_map"], seed=k, scene=v["scene"],
                               wavelengths=v["wavelengths"], **o)

    def timed(step):
        for k in range(args.warmup):
            step(k)
        torch.cuda.synchronize()
        r.set_timing(True)
        t0 = time.perf_counter()
        for k in range(args.steps):
            step(1000 + k)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        r.set_timing(False)
        for v in views:
            assert torch.isfinite(v["out"]).all().item(), "non-finite radiance"
        return el

    el_b = timed(step_batch)
    el_s = timed(step_seq)
    rate_b, rate_s = samples * args.steps / el_b, samples * args.steps / el_s
    log(f"[batch] one launch {rate_b / 1e6:.1f} M samples/s, view by view {rate_s / 1e6:.1f} M samples/s")
    # k_wf_extend launch times on the non-graph path for the roofline
    r.set_timing(True)
    tb = dict(tuned, use_graph=False)
    for v in views:
        v["seed"] = 2000
    r.render_batch(views, **tb)
    torch.cuda.synchronize()
    ems, en = r.stage_timing()
    st_times = r.stage_times()
    r.set_timing(False)
    # traversal work per sample from a counting pass of the plain megakernel
    # on every view at 2 spp
    r.set_counting(True)
    cs = 0
    for v in views:
        r.render_accumulate(v["out"], v["camera"], v["width"], v["height"], 2, max_depth=v["max_depth"],
                            photon_map=v["photon_map"], scene=v["scene"])
        cs += v["width"] * v["height"] * 2
    torch.cuda.synchronize()
    cnt = r.counters()
    r.set_counting(False)
    gem_views = sum(1 for v in views if v["photon_map"])
    pb = (PRIM_BYTES["gem"] * gem_views + PRIM_BYTES["box"] * (len(views) - gem_views)) / len(views)
    closest_bps = (cnt["node_tests"] * 256.0 + cnt["prim_tests"] * pb) / max(cnt["samples"], 1)
    bpl = closest_bps * samples / max(en, 1)
    achieved = bpl / (ems / 1e3 / max(en, 1)) / 1e9
    stages = stage_figures(st_times, cnt, samples)
    cpu = None
    if with_cpu:
        sc, cam, cfg = items[-1]
        cpu = cpu_baseline(sc, cam, cfg, args.cpu_seconds / 2)
        sc, cam, cfg = items[0]
        g = cpu_baseline(sc, cam, cfg, args.cpu_seconds / 2, photons=True)
        cpu["sample"] = "volume: " + cpu["sample"] + "; gem still: " + g["sample"]
        # whole-batch CPU rate: time-weighted over the two scene kinds
        gem_s = sum(v["width"] * v["height"] * v["spp"] for v in views if v["photon_map"]) / g["value"]
        vol_s = sum(v["width"] * v["height"] * v["spp"] for v in views if not v["photon_map"]) / cpu["value"]
        cpu["value"] = samples / (gem_s + vol_s)
    r.close()
    return {"value": rate_b, "ms_per_step": el_b * 1e3 / args.steps,
            "config": {"workload": "spin_batch_36_gems_plus_volume_box", "views": len(views), "samples": samples,
                       "samples_per_s": {"batch": rate_b, "view_by_view": rate_s},
                       "batch_over_view_by_view": rate_b / rate_s, "wavefront_opts": tuned,
                       "autotune": info and {k: info[k] for k in ("probe_spp", "probes", "seconds")},
                       "photons_stored": photons, "setup_s": round(setup_s, 2)},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "kernel": "k_wf_extend", "avg_launch_ms": ems / max(en, 1),
                         "bytes_per_launch": bpl, "launches": en, "stages": stages},
            "cpu_baseline": cpu}


def measure_adjoint(args, with_cpu):
    """Adjoint material gradients on the 512x512 gem: one gradient evaluation
    = the primal frame at 64 spp (wavefront renderer) + dLoss/dI of the L2
    image loss against a target rendered at the known parameters + the
    path-replay backward over the 20 parameters (rt_backward).  Untimed
    beside it: the oracle's host gradient of the same loss on the same
    per-pixel streams at full size (a different algorithm: forward score
    accumulation in double), and central finite differences with common
    random numbers of the same loss through the detached renderer at
    h = 1e-3 and 5e-4 relative, Richardson-extrapolated
    ((4 D(h/2) - D(h)) / 3).  For n_d one seed's difference quotient does
    not settle as h shrinks (nearly specular facets, alpha down to 0.005),
    so n_d is also compared as means over 24 seeds of the central difference
    and of the adjoint, with the standard error of the gap."""
    import numpy as np
    import torch
    from rustracer_amd import Renderer, scenes

    sc, cam, cfg = scenes.gem_adjoint(512, 512)
    W, H, spp, depth, P = cfg["width"], cfg["height"], cfg["spp"], cfg["max_depth"], cfg["params"]
    r = Renderer(0).upload(sc)
    opts = dict(capacity=1 << 23, grid=8192, sort=False, use_graph=True)
    target = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    r.render_wavefront(target, cam, W, H, spp, max_depth=depth, seed=99, **opts)
    target /= spp
    theta = cfg["theta_start"]
    r.set_params(P, theta)
    img = torch.zeros_like(target)
    grad = torch.zeros(len(P), dtype=torch.float32, device="cuda")
    n = float(img.numel())

    def step_with(o, k):
        r.render_wavefront(img, cam, W, H, spp, max_depth=depth, seed=k, **o)
        dl = (2.0 / n) * (img / spp - target)
        r.backward(dl, cam, W, H, spp, P, max_depth=depth, seed=k, grad=grad, wavefront=True,
                   capacity=o["capacity"], grid=o["grid"], sort=o["sort"], streams=o["streams"])

    # chunk capacity x streams: a few probes of one evaluation each
    probes = []
    for cap, ns, grid in ((1 << 24, 1, 16384), (1 << 23, 2, 16384), (1 << 23, 1, 8192), (1 << 22, 2, 8192),
                          (1 << 22, 3, 8192), (1 << 21, 3, 8192), (1 << 21, 4, 8192)):
        o = dict(opts, capacity=cap, streams=ns, grid=grid)
        step_with(o, 500)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        step_with(o, 501)
        torch.cuda.synchronize()
        probes.append((time.perf_counter() - t0, o))
    opts = min(probes, key=lambda x: x[0])[1]
    bopts = dict(capacity=opts["capacity"], grid=opts["grid"], sort=opts["sort"], streams=opts["streams"])
    log(f"[adjoint] options {opts} ({', '.join(f'{t * 1e3:.0f}' for t, _ in probes)} ms per probe)")

    def step(k):
        step_with(opts, k)

    for k in range(args.warmup):
        step(k)
    torch.cuda.synchronize()
    r.set_timing(True)
    t0 = time.perf_counter()
    for k in range(args.steps):
        step(1000 + k)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    kms, nl = r.timing()
    r.set_timing(False)
    rate = args.steps / el
    g = grad.cpu().numpy().astype(np.float64)
    assert np.isfinite(g).all(), "non-finite gradient"
    log(f"[adjoint] {rate:.2f} gradient evaluations/s ({el * 1e3 / args.steps:.1f} ms each)")
    # backward alone, HIP events on the library stream
    img.zero_()
    r.render_wavefront(img, cam, W, H, spp, max_depth=depth, seed=7, **opts)
    dl = (2.0 / n) * (img / spp - target)
    torch.cuda.synchronize()
    r.set_timing(True)
    g7 = r.backward(dl, cam, W, H, spp, P, max_depth=depth, seed=7, wavefront=True, **bopts).astype(np.float64)
    bms, bn = r.timing()
    r.set_timing(False)
    r.set_timing(True)
    g7m = r.backward(dl, cam, W, H, spp, P, max_depth=depth, seed=7).astype(np.float64)
    mms, mn = r.timing()
    r.set_timing(False)
    log(f"[adjoint] backward: wavefront {bms:.1f} ms, megakernel {mms:.1f} ms")
    # finite differences of the same loss, same streams (seed 7)
    tgt = target.double()

    def loss(th):
        out = torch.zeros_like(img)
        r.render_detached(out, cam, W, H, spp, P, th, max_depth=depth, seed=7)
        return float(((out.double() / spp - tgt) ** 2).mean())

    def central(j, hrel):
        h = hrel * abs(float(theta[j]))
        tp, tm = theta.astype(np.float64).copy(), theta.astype(np.float64).copy()
        tp[j] += h
        tm[j] -= h
        return (loss(tp) - loss(tm)) / (float(np.float32(tp[j])) - float(np.float32(tm[j])))

    fd1, fd2 = np.zeros(len(P)), np.zeros(len(P))
    t0 = time.time()
    for j in range(len(P)):
        fd1[j], fd2[j] = central(j, 1e-3), central(j, 5e-4)
    fd = (4.0 * fd2 - fd1) / 3.0
    fd_s = time.time() - t0
    rel = np.abs(g7 - fd) / np.maximum(np.abs(fd), 1e-30)
    rel1 = np.abs(g7 - fd1) / np.maximum(np.abs(fd1), 1e-30)
    within = rel <= 2e-2
    log(f"[adjoint] finite differences (Richardson, h = 1e-3 and 5e-4): {int(within.sum())}/{len(P)} within 2%, "
        f"max relative error {float(rel.max()):.2e} (plain central h = 1e-3: max {float(rel1.max()):.2e}) ({fd_s:.1f}s)")
    # n_d over many seeds: the detached frame's per-seed dependence on n_d is
    # too rough for one seed's difference quotient to converge (h and h / 2
    # scatter by 10-20 % at 64 spp), so compare means over seeds instead
    nd_seeds = list(range(300, 324))
    gs, fs = [], []
    t0 = time.time()
    for sd in nd_seeds:
        r.render_wavefront(img, cam, W, H, spp, max_depth=depth, seed=sd, **opts)
        dls = (2.0 / n) * (img / spp - target)
        gs.append(float(r.backward(dls, cam, W, H, spp, P, max_depth=depth, seed=sd, wavefront=True, **bopts)[0]))

        def loss_s(th):
            out = torch.zeros_like(img)
            r.render_detached(out, cam, W, H, spp, P, th, max_depth=depth, seed=sd)
            return float(((out.double() / spp - tgt) ** 2).mean())
        h = 1e-3 * abs(float(theta[0]))
        tp, tm = theta.astype(np.float64).copy(), theta.astype(np.float64).copy()
        tp[0] += h
        tm[0] -= h
        fs.append((loss_s(tp) - loss_s(tm)) / (float(np.float32(tp[0])) - float(np.float32(tm[0]))))
    gs, fs = np.array(gs), np.array(fs)
    dif = fs - gs
    gm = float(gs.mean())
    nd_stats = {"seeds": len(nd_seeds), "h_rel": 1e-3, "adjoint_mean": gm, "fd_mean": float(fs.mean()),
                "mean_rel_gap": float(dif.mean() / abs(gm)),
                "se_rel_gap": float(dif.std(ddof=1) / np.sqrt(len(dif)) / abs(gm)),
                "per_seed_rel_gap_abs_median": float(np.median(np.abs(dif / gs)))}
    nd_stats["ci95_rel_gap"] = [nd_stats["mean_rel_gap"] - 1.96 * nd_stats["se_rel_gap"],
                                nd_stats["mean_rel_gap"] + 1.96 * nd_stats["se_rel_gap"]]
    log(f"[adjoint] n_d over {len(nd_seeds)} seeds ({time.time() - t0:.1f}s): mean FD {fs.mean():.5g}, mean adjoint "
        f"{gm:.5g}, gap {nd_stats['mean_rel_gap'] * 100:.2f} % +- {nd_stats['se_rel_gap'] * 100:.2f} % (SE)")
    # the oracle's host gradient of the same loss, same per-pixel streams, full size
    from oracle import oracle as O
    orc = O.OracleScene(sc)
    orc.set_params(P, theta)
    t0 = time.time()
    gh = orc.backward(cam, Renderer.params(W, H, max_depth=depth, seed=7), spp, P, dl.cpu().numpy(), mode=1,
                      nthreads=O.threads()).astype(np.float64)
    host_s = time.time() - t0
    rel_h = np.abs(g7 - gh) / np.maximum(np.abs(gh), 1e-30)
    log(f"[adjoint] host replay of the same streams ({W}x{H}x{spp}, {host_s:.1f}s): max relative difference "
        f"{float(rel_h.max()):.2e}, median {float(np.median(rel_h)):.2e}")
    # traversal work of the backward: two traces of every path (primal + replay)
    r.set_counting(True)
    r.render_accumulate(img, cam, W, H, 4, max_depth=depth, seed=7)
    torch.cuda.synchronize()
    cnt = r.counters()
    r.set_counting(False)
    pb = PRIM_BYTES["gem"]
    bps = ((cnt["node_tests"] + cnt["shadow_node_tests"]) * 256.0 + (cnt["prim_tests"] + cnt["shadow_prim_tests"]) * pb) \
        / max(cnt["samples"], 1)
    # one "launch" = one chunk of paths through the primal and replay wavefronts
    bpl = 2.0 * bps * W * H * spp / max(bn, 1)
    achieved = bpl / (bms / 1e3 / max(bn, 1)) / 1e9
    cpu = None
    if with_cpu:
        cpu = cpu_adjoint_baseline(sc, cam, cfg, args.cpu_seconds)
    r.close()
    return {"value": rate, "unit": "gradient evaluations/s", "ms_per_step": el * 1e3 / args.steps,
            "config": {"workload": "gem_adjoint_512x512x64_20_params", "width": W, "height": H, "spp": spp,
                       "max_depth": depth, "params": len(P), "forward": "wavefront", "wavefront_opts": opts,
                       "backward": "path replay, wavefront", "backward_ms": bms, "backward_megakernel_ms": mms,
                       "wavefront_vs_megakernel_grad_max_rel": float(np.abs(g7 - g7m).max() / max(np.abs(g7m).max(), 1e-30)),
                       "fd": "central, common random numbers, Richardson over h = 1e-3 and 5e-4 relative",
                       "fd_within_2pct": int(within.sum()), "fd_rel_err": [float(x) for x in rel],
                       "fd_rel_err_plain_h1e-3": [float(x) for x in rel1],
                       "fd_n_d_over_seeds": nd_stats,
                       "host_replay_max_rel": float(rel_h.max()), "host_replay_rel": [float(x) for x in rel_h],
                       "host_replay_seconds": round(host_s, 1),
                       "gradient": [float(x) for x in g7], "finite_difference": [float(x) for x in fd],
                       "host_replay_gradient": [float(x) for x in gh]},
            "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "kernel": "rt_backward_wavefront (primal + replay stages)",
                         "avg_launch_ms": bms / max(bn, 1),
                         "bytes_per_launch": bpl},
            "cpu_baseline": cpu}


def cpu_adjoint_baseline(scene, cam, cfg, seconds):
    """The oracle's host gradient (forward score accumulation) plus its primal
    frame, all host cores, on passes of the 512x512 frame at 1 sample per
    pixel until `seconds` are used; scaled to gradient evaluations of 64 spp."""
    import numpy as np
    from oracle import oracle as O
    from rustracer_amd import Renderer
    cores = O.threads()
    orc = O.OracleScene(scene)
    P = cfg["params"]
    orc.set_params(P, cfg["theta_start"])
    W, H, spp = cfg["width"], cfg["height"], cfg["spp"]
    dl = np.full((H, W, 3), 1.0 / (W * H * 3), np.float32)
    done = 0
    t0 = time.perf_counter()
    while True:
        sp = Renderer.params(W, H, sample_idx=done, max_depth=cfg["max_depth"], seed=7)
        orc.render_sample(cam, sp, mode=1, nthreads=cores)
        orc.backward(cam, sp, 1, P, dl, mode=1, nthreads=cores)
        done += 1
        if time.perf_counter() - t0 >= seconds or done >= 64:
            break
    el = time.perf_counter() - t0
    return {"value": done / el / spp, "unit": "gradient evaluations/s", "cores": cores, "kind": "port",
            "sample": f"{done} passes of {W}x{H} at 1 sample/pixel (primal + forward-score gradient) of the {spp} "
                      f"per evaluation, {el:.1f}s"}


def measure_rmse(args, with_cpu=True):
    """"At fixed RMSE" on the headline gem (Sellmeier, k = 64 within 0.15) at
    640 x 360, depth 24.  Target: a 4096-spp image of the oracle (reference
    row streams, its own 2M-photon map).  The device renders with photons it
    traced itself (rt_photon_build).  RMSE is linear HDR over mean radiance.
    Two comparisons of 64-spp frames:
      * paired: frame k of the wavefront renderer against the oracle's frame
        on the same per-pixel streams (mode 1, same seed).  Both estimators
        draw the same numbers, so the frame-to-frame spread of the caustics
        cancels and the ratio of mean squared errors is pinned to a fraction
        of a percent;
      * independent: the oracle's own row streams at disjoint sample indices,
        the reference's estimator as the reference draws it.
    Equal RMSE at equal spp makes samples/s the figure of merit.  Every
    render (target included) draws four hero wavelengths over [380, 780) nm,
    the mode the gem workload is timed in."""
    import numpy as np
    import torch
    from oracle import oracle as O
    from rustracer_amd import Renderer, scenes
    from rustracer_amd import _ffi as F

    W, H, spp, ref_spp, K = 640, 360, 64, 4096, 8
    sc, cam, cfg = scenes.gem_caustics(W, H, **scenes.GEM_BENCH)
    D = cfg["max_depth"]
    cores = O.threads()
    orc = O.OracleScene(sc)
    lq = cfg["light_quad"]
    pp = F.rt_photon_params()
    pp.emitter = F.EMIT_QUAD
    for k in range(3):
        pp.origin[k], pp.u[k], pp.v[k], pp.color[k] = lq["q"][k], lq["u"][k], lq["v"][k], cfg["light_color"][k]
    pp.num_photons, pp.gather_radius, pp.knn = cfg["num_photons"], cfg["gather_radius"], cfg["knn"]
    WL = 4
    pp.wavelengths = WL
    t0 = time.time()
    nph = orc.photon_build(pp, nthreads=cores)
    p = Renderer.params(W, H, sample_idx=0, max_depth=D, use_photons=True, wavelengths=WL)
    ref = orc.render_accumulate(cam, p, ref_spp, mode=0, nthreads=cores).astype(np.float64) / ref_spp
    ref_s = time.time() - t0
    log(f"[rmse] reference: {ref_spp} spp at {W}x{H} on {cores} host threads in {ref_s:.1f}s, {nph} photons")

    def mse(img):
        return float(((np.asarray(img, np.float64) / spp - ref) ** 2).mean())

    r = Renderer(0).upload(sc)
    nd = r.photon_build(pp)                          # the device traces its own map
    a = torch.zeros((H, W, 3), dtype=torch.float32, device="cuda")
    m_wf, m_pair, m_ind = [], [], []
    for k in range(K):
        r.render_wavefront(a, cam, W, H, spp, max_depth=D, photon_map=True, seed=11 + k, wavelengths=WL)
        m_wf.append(mse(a.cpu().numpy()))
    r.render_accumulate(a, cam, W, H, spp, max_depth=D, photon_map=True, seed=11, wavelengths=WL)
    m_plain = mse(a.cpu().numpy())
    r.close()
    for k in range(K):
        pc = Renderer.params(W, H, max_depth=D, use_photons=True, seed=11 + k, wavelengths=WL)
        m_pair.append(mse(orc.render_accumulate(cam, pc, spp, mode=1, nthreads=cores)))
        pc = Renderer.params(W, H, sample_idx=100000 + 1000 * k, max_depth=D, use_photons=True, wavelengths=WL)
        m_ind.append(mse(orc.render_accumulate(cam, pc, spp, mode=0, nthreads=cores)))
    mw, mp, mi = np.array(m_wf), np.array(m_pair), np.array(m_ind)
    e_wf, e_pair, e_ind = (float(np.sqrt(x.mean())) for x in (mw, mp, mi))
    ratio = e_wf / e_pair
    # paired: delta method on sqrt(mean(mw) / mean(mp)) with the per-frame differences
    d = mw / mp.mean() - mp / mp.mean()
    se = 0.5 * float(np.sqrt(d.var(ddof=1) / K))
    ratio_ind = e_wf / e_ind
    se_ind = 0.5 * ratio_ind * float(np.sqrt(mw.var(ddof=1) / K / mw.mean() ** 2 + mi.var(ddof=1) / K / mi.mean() ** 2))
    log(f"[rmse] {spp} spp vs the {ref_spp}-spp target, {K} frames each: wavefront {e_wf:.5f}, oracle on the same "
        f"streams {e_pair:.5f} (ratio {ratio:.4f} +- {se:.4f}), oracle row streams {e_ind:.5f} (ratio {ratio_ind:.4f} "
        f"+- {se_ind:.4f}); device photons {nd}, oracle photons {nph}")
    return {"value": ratio, "unit": "RMSE ratio (wavefront / oracle, same streams) at equal spp", "ms_per_step": None,
            "config": {"workload": "gem_caustics_640x360_rmse", "width": W, "height": H, "spp": spp,
                       "reference_spp": ref_spp, "frames": K, "photons": {"device": nd, "oracle": nph},
                       "wavelengths": "four hero wavelengths over [380, 780) nm",
                       "rmse": {"wavefront": e_wf, "oracle_same_streams": e_pair, "oracle_row_streams": e_ind,
                                "plain_first_frame": float(np.sqrt(m_plain))},
                       "rmse_per_frame": {"wavefront": [float(np.sqrt(x)) for x in mw],
                                          "oracle_same_streams": [float(np.sqrt(x)) for x in mp],
                                          "oracle_row_streams": [float(np.sqrt(x)) for x in mi]},
                       "ratio_se": se, "within_2pct": bool(abs(ratio - 1.0) <= 0.02 and se <= 0.007),
                       "ratio_row_streams": ratio_ind, "ratio_row_streams_se": se_ind,
                       "row_streams_consistent_2se": bool(abs(ratio_ind - 1.0) <= 2 * se_ind),
                       "reference_seconds": round(ref_s, 1)},
            "roofline": {"achieved": None, "frac": None, "kernel": None, "avg_launch_ms": None},
            "cpu_baseline": None}


def measured_traffic(workload, opts, bytes_per_launch):
    """HBM bytes per k_wf_extend dispatch from the committed rocprofv3 PMC
    summary (FETCH_SIZE + WRITE_SIZE, copy-calibrated) of the same kernel
    variant (occupancy target, block size, material sort) on this workload.
    A dispatch's DRAM bytes scale with the rays it traces, so when the
    record was measured at another queue capacity its measured ratio of DRAM
    bytes to algorithmic bytes is applied to this launch's algorithmic bytes
    (the note says so).  Otherwise None and the reason."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f).get(workload)
    except (OSError, ValueError):
        return None, "no PMC summary"
    if not rec:
        return None, "no PMC summary for this workload"
    # graph replay and chunk streams change neither what a dispatch reads nor writes
    same = [r for r in (rec if isinstance(rec, list) else [rec])
            if all(opts.get(k, d) == r["wavefront_opts"].get(k, d) for k, d in (("occupancy", 0), ("block", 128), ("sort", 1)))]
    for r in same:
        if all(opts.get(k) == r["wavefront_opts"].get(k) for k in ("capacity", "grid")):
            return r["traffic_bytes_per_dispatch"], r["source"]
    if same:
        r = same[0]
        o = r["wavefront_opts"]
        ratio = r["traffic_bytes_per_dispatch"] / r["algorithmic_bytes_per_dispatch"]
        return ratio * bytes_per_launch, (f"{r['source']} (measured at capacity {o.get('capacity')}, grid "
                                          f"{o.get('grid')}: DRAM / algorithmic bytes {ratio:.3f}, applied to "
                                          f"this launch's algorithmic bytes)")
    return None, f"no PMC summary measured for the kernel variant of {opts}"


def measure_output(args, with_cpu=True):
    """The render loop's per-pixel work and the saved image on the device, at
    1280x720 with every buffer in HBM: accumulate_sample (12 B sample + 29 B
    state read, 28 B written per pixel), mark_converged (29 B read + 1 B
    written), the 8-bit resolve with optics (4 x 12 B + 4 x 4 B read through
    the bilinear taps, 3 B written), the AOV pass (one closest-hit ray per
    pixel on the 1.31M-triangle mesh) and the render loop on the mesh with the
    firefly clamp on: fused per pixel (rt_render_progressive, 16 passes) and
    through the wavefront (rt_render_progressive_wavefront, 64 passes).  Kernel time from HIP events on
    the library stream; beside it the checker's accumulate_sample on the host."""
    import numpy as np
    import torch
    from rustracer_amd import Renderer, scenes
    W, H = 1280, 720
    n = W * H
    r = Renderer(0)
    g = torch.Generator(device="cuda").manual_seed(1)
    # four different passes in turn: the per-pixel variance stays well above
    # the convergence threshold, so every call does every pixel's work
    sc_bufs = [torch.rand((n, 3), generator=g, device="cuda") * 2.0 for _ in range(4)]
    turn = [0]

    def next_pass():
        turn[0] += 1
        return sc_bufs[turn[0] % 4]
    st = dict(accum=torch.zeros((n, 3), device="cuda"), samples=torch.zeros(n, dtype=torch.int32, device="cuda"),
              var_m2=torch.zeros((n, 3), device="cuda"), converged=torch.zeros(n, dtype=torch.uint8, device="cuda"))
    o = Renderer.accum_opts(adaptive=True, threshold=0.05)
    rgb8 = torch.zeros((H, W, 3), dtype=torch.uint8, device="cuda")

    def timed(fn, reps):
        """Kernel ms per call: the library's HIP events on its own stream."""
        fn()
        torch.cuda.synchronize()
        r.set_timing(True)
        for _ in range(reps):
            fn()
        ms, _ = r.timing()
        r.set_timing(False)
        return ms / reps

    t_acc = timed(lambda: r.accumulate_sample(next_pass(), st, o), 20)
    t_mark = timed(lambda: r.mark_converged(st, o), 20)
    t_res = timed(lambda: r.resolve_rgb8(st["accum"], W, H, samples=st["samples"], tonemapper="agx", vignette=0.2,
                                         distortion=0.05, out=rgb8), 20)
    converged = float(st["converged"].float().mean())
    assert converged < 0.01, f"timing state converged ({converged}): the bookkeeping kernels would skip work"
    scene, cam, cfg = scenes.mesh_envmap(W, H, level=8)
    r.upload(scene)
    alb = torch.zeros((H, W, 3), device="cuda")
    nrm = torch.zeros((H, W, 3), device="cuda")
    t_aux = timed(lambda: r.render_aux(cam, W, H, alb, nrm), 5)
    st2 = dict(accum=torch.zeros((n, 3), device="cuda"), samples=torch.zeros(n, dtype=torch.int32, device="cuda"),
               var_m2=torch.zeros((n, 3), device="cuda"), converged=torch.zeros(n, dtype=torch.uint8, device="cuda"))
    passes = 16
    t_prog = timed(lambda: r.render_progressive(st2, cam, W, H, passes, Renderer.accum_opts(), max_depth=cfg["max_depth"]), 1)
    wopts = dict(capacity=1 << 23, grid=8192, sort=1, occupancy=4, streams=3)
    wpasses = 64
    t_progw = timed(lambda: r.render_progressive(st2, cam, W, H, wpasses, Renderer.accum_opts(), max_depth=cfg["max_depth"],
                                                 wavefront=True, **wopts), 1)
    gbs = lambda b, ms: b / (ms * 1e-3) / 1e9
    b_acc, b_mark, b_res = n * (12 + 29 + 28), n * (29 + 1), n * (4 * 12 + 4 * 4 + 3)
    log(f"[output] accumulate {t_acc:.3f} ms ({gbs(b_acc, t_acc):.0f} GB/s), mark {t_mark:.3f} ms "
        f"({gbs(b_mark, t_mark):.0f} GB/s), resolve {t_res:.3f} ms ({gbs(b_res, t_res):.0f} GB/s), aux {t_aux:.2f} ms, "
        f"progressive {n * passes / (t_prog * 1e-3) / 1e6:.1f} M samples/s (megakernel), "
        f"{n * wpasses / (t_progw * 1e-3) / 1e6:.1f} M samples/s (wavefront)")
    cpu = None
    if with_cpu:
        from oracle import oracle as O
        hs = {k: (v.cpu().numpy().view(np.uint32) if k == "samples" else v.cpu().numpy()) for k, v in st.items()}
        x = sc_bufs[1].cpu().numpy()
        reps, t0 = 0, time.perf_counter()
        while time.perf_counter() - t0 < 3.0:
            O.accumulate_sample(x, hs)
            reps += 1
        el = time.perf_counter() - t0
        cpu = {"value": n * reps / el, "unit": "pixels/s", "cores": 1, "kind": "port",
               "sample": f"{reps} accumulate_sample passes over 1280x720 on one thread, {el:.1f}s"}
    r.close()
    return {"value": n / (t_acc * 1e-3), "unit": "pixels/s (accumulate_sample)", "ms_per_step": t_acc,
            "config": {"workload": "render_loop_output_1280x720", "width": W, "height": H,
                       "accumulate_ms": t_acc, "mark_converged_ms": t_mark, "resolve_rgb8_ms": t_res,
                       "aux_pass_ms": t_aux, "aux_rays_per_s": n / (t_aux * 1e-3),
                       "progressive_samples_per_s": n * passes / (t_prog * 1e-3), "progressive_passes": passes,
                       "progressive_wavefront_samples_per_s": n * wpasses / (t_progw * 1e-3),
                       "progressive_wavefront_passes": wpasses, "progressive_wavefront_opts": wopts,
                       "bytes_per_pixel": {"accumulate": 69, "mark_converged": 30, "resolve_rgb8": 67}},
            "roofline": {"bound": "hbm", "achieved": gbs(b_acc, t_acc), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": gbs(b_acc, t_acc) / HBM_PEAK_GBS, "kernel": "k_accumulate",
                         "avg_launch_ms": t_acc, "bytes_per_launch": b_acc,
                         "others": {"k_mark_converged": {"achieved": gbs(b_mark, t_mark)},
                                    "k_resolve": {"achieved": gbs(b_res, t_res)}}},
            "cpu_baseline": cpu}


def cpu_baseline(scene, cam, cfg, seconds, photons=False):
    """The oracle's render_tiles with the reference's row streams, all host
    cores, passes of the same frame until `seconds` of wall time are used
    (scene build and photon tracing excluded, as in the reference's --bench)."""
    from oracle import oracle as O
    from rustracer_amd import Renderer
    from rustracer_amd import _ffi as F
    cores = O.threads()
    orc = O.OracleScene(scene)
    if photons and "caustics" in cfg:
        orc.photon_build(Renderer.caustic_params(cfg["caustics"]), nthreads=cores)
    elif photons:
        lq = cfg["light_quad"]
        pp = F.rt_photon_params()
        pp.emitter = F.EMIT_QUAD
        for k in range(3):
            pp.origin[k], pp.u[k], pp.v[k], pp.color[k] = lq["q"][k], lq["u"][k], lq["v"][k], cfg["light_color"][k]
        pp.num_photons, pp.gather_radius, pp.knn = cfg["num_photons"], cfg["gather_radius"], cfg.get("knn", 0)
        pp.wavelengths = cfg.get("wavelengths", 0)
        orc.photon_build(pp, nthreads=cores)
    W, H, spp = cfg["width"], cfg["height"], cfg["spp"]
    strata = int(spp ** 0.5)

    def params(k):
        return Renderer.params(W, H, sample_idx=k, strata=strata, max_depth=cfg["max_depth"], use_photons=photons,
                               wavelengths=cfg.get("wavelengths", 0))
    orc.render_sample(cam, params(spp - 1), mode=0, nthreads=cores)      # warm-up pass (untimed)
    done = 0
    t0 = time.perf_counter()
    while True:
        orc.render_sample(cam, params(done), mode=0, nthreads=cores)
        done += 1
        if time.perf_counter() - t0 >= seconds or done >= 100000:
            break
    el = time.perf_counter() - t0
    # one thread, one pass: the scalar rate beside the all-core one
    t1 = time.perf_counter()
    orc.render_sample(cam, params(0), mode=0, nthreads=1)
    one = W * H / (time.perf_counter() - t1)
    return {"value": W * H * done / el, "unit": "samples/s", "cores": cores, "kind": "port",
            "sample": f"{done} full passes of {W}x{H} at 1 sample/pixel (of {spp}), {el:.1f}s, after one warm-up pass",
            "one_thread": one, "host": host_cpu()}


def host_cpu():
    """The host CPU model and logical CPU count, for the CPU baseline."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"model": model, "logical_cpus": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="mesh", choices=sorted(workloads()) + ["batch", "adjoint", "rmse", "output"])
    ap.add_argument("--also", default="gem,batch,adjoint,rmse,output,bench_toml",
                    help="comma-separated extra workloads ('' for none): " + ", ".join(sorted(workloads())) +
                         ", batch, adjoint, rmse, output")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--wf-opts", default="",
                    help='fixed wavefront options as JSON (e.g. \'{"occupancy": 4, "capacity": 4194304}\'), no autotune')
    ap.add_argument("--renderer", default="both", choices=["both", "wavefront", "plain"])
    ap.add_argument("--calib-copy-mb", type=int, default=0,
                    help="first run a device-to-device copy of this many MiB (counter calibration for --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (every box here has one MI355X)",
              file=sys.stderr)
        sys.exit(2)
    if args.calib_copy_mb:
        import torch
        n = args.calib_copy_mb * (1 << 20) // 4
        a = torch.ones(n, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        b.copy_(a)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calib_copy_mb} MiB read + {args.calib_copy_mb} MiB written")
    if args.config == "batch":
        main_res = measure_batch(args, not args.no_cpu)
    elif args.config == "adjoint":
        main_res = measure_adjoint(args, not args.no_cpu)
    elif args.config == "rmse":
        main_res = measure_rmse(args)
    elif args.config == "output":
        main_res = measure_output(args, not args.no_cpu)
    else:
        main_res = measure(args.config, args, with_cpu=not args.no_cpu)
    also = {}
    for name in [a for a in args.also.split(",") if a and a != args.config]:
        if name == "batch":
            m = measure_batch(args, not args.no_cpu)
        elif name == "adjoint":
            m = measure_adjoint(args, not args.no_cpu)
        elif name == "rmse":
            m = measure_rmse(args)
        elif name == "output":
            m = measure_output(args, not args.no_cpu)
        else:
            m = measure(name, args, with_cpu=not args.no_cpu)
        also[m["config"]["workload"]] = {"value": m["value"], "unit": m.get("unit", "samples/s"), "config": m["config"],
                                          "roofline": {k: m["roofline"][k] for k in ("achieved", "frac", "kernel", "avg_launch_ms",
                                                                                     "stages") if k in m["roofline"]},
                                          "cpu_baseline": m["cpu_baseline"]}
    line = {
        "metric": "GPU samples/s (wavefront vs plain) at fixed RMSE; adjoint grad evals/s",
        "value": main_res["value"],
        "unit": main_res.get("unit", "samples/s"),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": main_res["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": main_res["config"],
        "roofline": main_res["roofline"],
        "cpu_baseline": main_res["cpu_baseline"],
        "also": also,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
