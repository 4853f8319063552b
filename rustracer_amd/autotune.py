"""Autotuner for the wavefront renderer's launch parameters.

Picks the register budget (occupancy target) and block size of the traversal
stages, the path-queue capacity, the persistent grid of the stage kernels, the
material sort, hipGraph replay, one to four chunk streams and, with k-NN
photon maps, the dense-query count up to which the nearest-photon search runs
one query per wave, by coordinate descent on a short frame of
the actual workload (same scene, camera, depth and photon setting), timing
each candidate with HIP events on the library's stream.  The search is cheap
(each probe renders a few passes) and the result is returned as the kwargs of
`Renderer.render_wavefront`.
"""
import time

# up to 32 M paths per chunk (about 13 GB of queues per chunk stream): on
# 288 GB the larger chunks cost nothing and leave fewer latency-bound tails
CAPACITIES = [1 << 20, 1 << 21, 1 << 22, 1 << 23, 1 << 24, 1 << 25]
GRIDS = [1024, 2048, 4096, 8192, 16384]
BLOCKS = [64, 128, 256]
KNN_WAVE_MAX = [16384, 65536, 262144, 1 << 20]


def _probe(r, run, opts, repeats=2):
    r.set_timing(True)
    run(opts, 99)                       # warm / allocate
    warm_ms, _ = r.timing()
    if warm_ms > 500.0:                 # long probe frames: one timed run is enough
        repeats = 1
    r.set_timing(True)                  # (re-enabling resets the totals)
    for k in range(repeats):
        run(opts, 100 + k)
    ms, _ = r.timing()
    r.set_timing(False)
    return ms / repeats


def tune_wavefront(r, out, cam, W, H, spp, depth, photons=False, probe_spp=None, log=None, knn=False, wavelengths=0):
    """Tunes `Renderer.render_wavefront` on one frame.  Returns (best_opts,
    info) where info["table"] lists every probe as (opts, ms per probe frame)."""
    npix = W * H
    if probe_spp is None:
        # enough paths for two chunks of the largest capacity (eight of 8 M, so
        # one to four chunk streams differ in the probe), at most the real spp
        probe_spp = max(1, min(spp, (2 * CAPACITIES[-1] + npix - 1) // npix))

    def run(o, seed):
        r.render_wavefront(out, cam, W, H, probe_spp, max_depth=depth, photon_map=photons, seed=seed,
                           wavelengths=wavelengths, **o)
    best, info = tune(r, run, log, knn=photons and knn)
    info["probe_spp"] = probe_spp
    return best, info


def tune_batch(r, views, probe_spp=4, log=None, knn=True):
    """Tunes `Renderer.render_batch` on copies of the views with at most
    `probe_spp` samples per pixel each."""
    probe = [dict(v, spp=min(v["spp"], probe_spp)) for v in views]

    def run(o, seed):
        for v in probe:
            v["seed"] = seed
        r.render_batch(probe, **o)
    best, info = tune(r, run, log, knn=knn and any(v.get("photon_map") for v in views))
    info["probe_spp"] = probe_spp
    return best, info


def tune(r, run, log=None, knn=False):
    """Coordinate descent over the launch options; `run(opts, seed)` renders
    one probe frame on the renderer's stream."""
    best = dict(capacity=1 << 22, grid=2048, sort=1, use_graph=True, occupancy=0, streams=1, block=128)
    if knn:
        best["knn_wave_max"] = 65536
    table = []

    def measure(o):
        ms = _probe(r, run, o)
        table.append((dict(o), ms))
        if log:
            log(f"  autotune {o} -> {ms:.2f} ms")
        return ms

    t0 = time.time()
    best_ms = measure(best)
    axes = [("occupancy", [0, 4, 5]), ("block", BLOCKS), ("capacity", CAPACITIES), ("grid", GRIDS), ("sort", [0, 1, 2]),
            ("use_graph", [True, False]), ("streams", [1, 2, 3, 4])]
    if knn:
        axes.append(("knn_wave_max", KNN_WAVE_MAX))
    for key, values in axes:
        for v in values:
            if v == best[key]:
                continue
            o = dict(best)
            o[key] = v
            ms = measure(o)
            if ms < best_ms:
                best, best_ms = o, ms
    return best, {"probes": len(table), "seconds": round(time.time() - t0, 2),
                  "best_ms": best_ms, "table": [(o, round(ms, 3)) for o, ms in table]}
