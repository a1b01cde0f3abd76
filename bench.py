"""Benchmark of the MI355X render path on one GPU.

One step = one full frame of the instanced-icosphere workload: a 1,003,520-
triangle displaced icosphere instanced 64 times, 1920x1080 at 64 spp, max 8
bounces, mix(diffuse, transparent) materials, constant world background.
Inputs (scene tables, BVH, film) are resident in HBM before the timed region.

The timed path is the batched wavefront integrator (trace and shade kernels
over path state in HBM, all 64 samples of every pixel in one dispatch).  The plain
megakernel (one lane runs whole paths) is timed on one frame beside it.

`--workload many_lights` times the many-light scene instead (the instanced
scene plus 10,000 emissive triangles and 321 lamps, light-tree NEE,
1920x1080 at 32 spp, max 6 bounces), `--workload cornell` the Cornell box
(640x400 at 16 spp, max 4 bounces); the default line is the instanced one.
Shading runs the scene's surface programs, generated per closure-leaf layout
and compiled with hipRTC on the first (untimed) dispatch; `--interpreter`
shades with the built-in closure interpreter instead (the same film).
Before the timed region the session's wavefront launch shape is autotuned
(probe dispatches of 4 samples; the film does not depend on the shape).
Trace launches are timed with HIP events, which this runtime cannot time
inside a graph, so the timed dispatches launch from the host;
`--untimed-trace` replays each dispatch as a graph instead (no roofline).

Prints one JSON line with samples/s (value), Mrays/s, the HBM roofline of the
dominant kernel (the wavefront trace kernel, timed with HIP events around
each of its launches on the session stream) and the CPU oracle timed on a
bounded sample of the same workload on the host cores.  The default line
also carries `serving`, the metric's latency half: the first 60 requests of
the serving stream (Poisson arrivals at 200 req/s, seed 4, 30 % parameter
edits over the resident Cornell, instanced and many-light scenes) through the
render server and through the plain per-request path, with p50 / p99 request
latency, viewport p99 and the batched-vs-plain ratio.
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

METRIC = "Mrays/s & samples/s (batched vs plain) at 1 GPU; p50/p99 request latency"

FILM_BYTES_PER_PIXEL_DISPATCH = 104.0
HBM_PEAK_GBS = 8000.0
SAMPLES_PER_DISPATCH = 64


def bytes_per_ray(nodes: float, tris: float) -> float:
    """Algorithmic bytes per ray, B_ray = 48 + 32 * N_node + 36 * N_tri: ray in
    and hit out, one binary node, one triangle."""
    return 48.0 + 32.0 * nodes + 36.0 * tris


# Per workload: the scene, its settings, and the mean binary-SAH node visits
# and triangle tests per ray measured by the CPU oracle's BVH (16 bins,
# leaves <= 4) over one full-frame sample of that workload.  `traffic` is the
# HBM bytes the trace kernel moves per ray: rocprofv3 FETCH_SIZE x 2.0 +
# WRITE_SIZE x 1.0 (scale factors from a 1 GiB copy) divided by the rays of
# the profiled dispatch (profiles/wf_kernels_pmc.md), or None where no
# counter profile was taken.
WORKLOADS = {
    "cornell": {
        "scene": "cornell_box", "settings": "cornell_settings", "spp": 16,
        "nodes": 11.17, "tris": 7.26, "traffic": None,
        "describe": "Cornell box: 34 triangles, emissive ceiling quad, diffuse walls and boxes; "
                    "640x400, {spp} spp, max 4 bounces, seed 0",
    },
    "instanced": {
        "scene": "instanced_icospheres", "settings": "instanced_settings", "spp": 64,
        "nodes": 46.83, "tris": 8.59, "traffic": 1111.3,
        "describe": "instanced icospheres: 1,003,520-tri displaced icosphere x 64 instances, "
                    "mix(diffuse, transparent) materials, ground quad, constant background; "
                    "1920x1080, {spp} spp, max 8 bounces, seed 0",
    },
    "many_lights": {
        "scene": "many_lights", "settings": "many_light_settings", "spp": 32,
        "nodes": 94.81, "tris": 15.06, "traffic": 1847.9,
        "describe": "many lights: the instanced icospheres plus 10,000 emissive add(emission, diffuse) "
                    "triangles, 256 point, 64 spot and 1 distant lamp, light-tree NEE; "
                    "1920x1080, {spp} spp, max 6 bounces, seed 0",
    },
}


def cpu_baseline(snap, settings, total, seconds_target=15.0):
    """Time the CPU oracle on the full frame of the same workload: a 1 spp
    calibration pass (which also yields the oracle BVH's node visits and
    triangle tests per ray behind B_ray), then as many consecutive samples as
    fill the target duration, on the host cores."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_binding as ob

    threads = min(os.cpu_count() or 1, 16)
    scene = ob.OracleScene(snap)
    t0 = time.perf_counter()
    _, st0 = scene.render(settings, 0, 1, total, threads=threads)
    dt = max(time.perf_counter() - t0, 1e-3)
    spp = int(max(1, min(total - 1, round(seconds_target / dt))))
    t0 = time.perf_counter()
    _, st = scene.render(settings, 1, spp, total, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": st.samples / dt,
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full {settings.width}x{settings.height} frame, samples 1-{spp} of {total} "
                  f"({st.samples} samples) in {dt:.1f} s, {st.rays / max(dt, 1e-9) / 1e6:.2f} Mrays/s",
        "oracle_nodes_per_ray": st0.node_visits / max(st0.rays, 1),
        "oracle_tris_per_ray": st0.triangle_tests / max(st0.rays, 1),
    }


def calibration_copy(nbytes: int = 1 << 30) -> None:
    """One device-to-device copy of `nbytes` through the HIP runtime (its copy
    kernel reads and writes exactly nbytes): the known-size reference that
    scales FETCH_SIZE / WRITE_SIZE counters to bytes."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(dst), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        hip.hipMemset(src, 1, C.c_size_t(nbytes))
        hip.hipDeviceSynchronize()
        if hip.hipMemcpy(dst, src, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    print(f"bench: calibration copy of {nbytes} bytes done", file=sys.stderr, flush=True)


def serving_leg(requests: int, with_plain: bool, log) -> dict:
    """The request stream over the resident Cornell, instanced and many-light
    scenes: the render server (continuous batching) and, optionally, the
    plain path (one megakernel session per request, in arrival order)."""
    from psycles_amd import abi, scenes, serving
    from psycles_amd.render import RendererBackend
    from psycles_amd.server import RenderServer

    log("building the Cornell, instanced and many-light scenes")
    snaps = {"cornell": scenes.cornell_box(), "instanced": scenes.instanced_icospheres(),
             "many_lights": scenes.many_lights()}
    stream = serving.make_stream(snaps, requests)
    batched_backend = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT)
    compiled = {n: batched_backend.compile_scene(sn).scene for n, sn in snaps.items()}
    server = RenderServer(batched_backend)
    # warm the kernels and allocator with one small request per scene
    warm = []
    for n in snaps:
        st = serving.scene_settings(n, 64, 64)
        warm.append((st, server.submit(compiled[n], st, 0, 1, 1)))
    for _, h in warm:
        h.wait(-1.0)
        h.release()
    log(f"stream of {len(stream)} requests over {stream[-1].arrival:.2f} s; running the server")
    batched = serving.run_batched(server, compiled, stream, log)
    server.close()
    del compiled, batched_backend
    log(f"server: {batched['samples_per_s'] / 1e6:.1f} M samples/s, viewport p99 "
        f"{batched['viewport_latency_ms']['p99']} ms")
    out = {"workload": f"serving: {len(stream)} requests, Poisson arrivals at 200 req/s (seed 4) over resident "
                       "Cornell, instanced and many-light scenes; resolutions 256x256 to 1920x1080, 1-64 spp, "
                       "30% parameter-only edits",
           "batched": batched}
    if with_plain:
        plain_backend = RendererBackend(0, integrator=abi.INTEGRATOR_MEGAKERNEL)
        plain_compiled = {n: plain_backend.compile_scene(sn).scene for n, sn in snaps.items()}
        log("running the plain per-request path")
        out["plain"] = serving.run_plain(plain_backend, plain_compiled, stream, log)
        out["batched_vs_plain"] = batched["samples_per_s"] / out["plain"]["samples_per_s"]
    return out


def serving_summary(leg: dict) -> dict:
    """The serving block of the default line: throughput, latency
    percentiles and the batched-vs-plain ratio."""
    b = leg["batched"]
    out = {"workload": leg["workload"], "samples_per_s": b["samples_per_s"], "makespan_s": b["makespan_s"],
           "latency_p50_ms": b["latency_ms"]["p50"], "latency_p99_ms": b["latency_ms"]["p99"],
           "viewport_p50_ms": b["viewport_latency_ms"]["p50"], "viewport_p99_ms": b["viewport_latency_ms"]["p99"],
           "viewport_requests": b["viewport_latency_ms"]["count"], "requests": b["requests"], "edits": b["edits"]}
    if "plain" in leg:
        p = leg["plain"]
        out.update({"plain_samples_per_s": p["samples_per_s"], "plain_latency_p50_ms": p["latency_ms"]["p50"],
                    "plain_viewport_p99_ms": p["viewport_latency_ms"]["p99"],
                    "batched_vs_plain": leg["batched_vs_plain"]})
    return out


def serving_bench(args) -> int:
    """`--workload serving`: the stream alone on the line."""
    def log(msg):
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    leg = serving_leg(args.requests, not args.no_plain, log)
    batched = leg["batched"]
    result = {
        "metric": METRIC,
        "value": batched["samples_per_s"],
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": batched["makespan_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": leg["workload"],
                   "integrator": "render server (continuous batching of wavefront batches)"},
        "batched": batched,
    }
    if "plain" in leg:
        result["plain"] = leg["plain"]
        result["batched_vs_plain"] = leg["batched_vs_plain"]
    print(json.dumps(result))
    return 0


def uneven_bench(args) -> int:
    """The first `--requests` requests of the uneven-path batch (stacked
    transparent slabs, every request its own scene, side 128-1024, 1-256 spp,
    seed 5) submitted together: the render server against one megakernel
    session per request."""
    from psycles_amd import abi, scenes, serving
    from psycles_amd.render import RendererBackend
    from psycles_amd.server import RenderServer

    def log(msg):
        print(f"bench: {msg}", file=sys.stderr, flush=True)

    batch = scenes.uneven_batch()[: args.requests]
    log(f"compiling {len(batch)} uneven-batch scenes")
    backend = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT)
    compiled = [backend.compile_scene(r.snapshot).scene for r in batch]
    server = RenderServer(backend, path_capacity=args.server_paths, streams=args.server_streams,
                          target_batch_ms=args.server_batch_ms)
    warm_settings = batch[0].settings(64)
    warm = server.submit(compiled[0], warm_settings, 0, 1, 1)
    warm.wait(-1.0)
    warm.release()
    batched = serving.run_uneven_batched(server, compiled, batch, log)
    server.close()
    result = {
        "metric": METRIC,
        "value": batched["samples_per_s"],
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": batched["makespan_s"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"uneven-path batch: the first {len(batch)} of 512 requests (seed 5), each its own scene "
                        "of 1-16 stacked transparent slabs (transparency U[0.5, 0.95]) over diffuse geometry, "
                        "square side U{128..1024}, U{1..256} spp, max 8 bounces with RR, 32 transparent bounces; "
                        "the optional homogeneous box is not built",
            "integrator": "render server (wavefront batches)",
            "server_streams": args.server_streams or 6,
            "server_paths": args.server_paths or 8388608,
            "server_batch_ms": args.server_batch_ms or 12,
        },
        "batched": batched,
    }
    if not args.no_plain:
        plain_backend = RendererBackend(0, integrator=abi.INTEGRATOR_MEGAKERNEL)
        plain_compiled = [plain_backend.compile_scene(r.snapshot).scene for r in batch]
        log("running the plain per-request path")
        plain = serving.run_uneven_plain(plain_backend, plain_compiled, batch, log)
        result["plain"] = plain
        result["batched_vs_plain"] = batched["samples_per_s"] / plain["samples_per_s"]
    print(json.dumps(result))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", choices=sorted(WORKLOADS) + ["serving", "uneven"], default="instanced")
    ap.add_argument("--requests", type=int, default=None,
                    help="serving: requests in the stream (120); uneven: leading requests of the batch (16)")
    ap.add_argument("--spp", type=int, default=None, help="samples per pixel (default: the workload's)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true", help="skip the megakernel frame")
    ap.add_argument("--serving-requests", type=int, default=60,
                    help="requests of the serving leg on the default line (0 skips it)")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default wavefront launch shape instead of timing candidates first")
    ap.add_argument("--no-graphs", action="store_true",
                    help="launch every kernel of a dispatch from the host instead of replaying a graph")
    ap.add_argument("--server-streams", type=int, default=0,
                    help="uneven: render-server streams (2-8; 0: the server's default of 6)")
    ap.add_argument("--server-paths", type=int, default=0,
                    help="uneven: render-server paths per batch (0: the server's default of 8,388,608)")
    ap.add_argument("--server-batch-ms", type=int, default=0,
                    help="uneven: render-server target batch duration (0: the server's default of 12 ms)")
    ap.add_argument("--tuning", default=None,
                    help="explicit wavefront launch shape instead of the autotuner, e.g. tile_width=16,sort_shading=1")
    ap.add_argument("--interpreter", action="store_true",
                    help="shade with the closure interpreter instead of the scene's hipRTC-generated surface "
                         "programs")
    ap.add_argument("--untimed-trace", action="store_true",
                    help="do not time each trace launch with HIP events (no roofline; dispatches may then "
                         "replay as graphs)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="finish with a 1 GiB device copy (counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(every box here has one MI355X)", file=sys.stderr)
        return 2

    from psycles_amd import abi, scenes
    from psycles_amd.render import RendererBackend

    if args.workload == "serving":
        args.requests = args.requests or 120
        return serving_bench(args)
    if args.workload == "uneven":
        args.requests = args.requests or 16
        return uneven_bench(args)
    w = WORKLOADS[args.workload]
    print(f"bench: building the {args.workload} scene", file=sys.stderr, flush=True)
    snap = getattr(scenes, w["scene"])()
    settings = getattr(scenes, w["settings"])()
    b_ray = bytes_per_ray(w["nodes"], w["tris"])
    bvh = None
    if not args.no_autotune and not args.tuning:
        # scene set-up: the BVH branching factor, by compiling the scene at
        # each width and timing the same 4-sample probe (films are the same)
        from psycles_amd.tuning import tune_bvh_width
        t0 = time.perf_counter()
        bvh = tune_bvh_width(snap, settings, probe_samples=4)
        bvh["seconds"] = time.perf_counter() - t0
        print(f"bench: BVH width {bvh['chosen']} (probe ms by width {bvh['ms']}) in {bvh['seconds']:.1f} s",
              file=sys.stderr, flush=True)
    backend = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT,
                              max_samples_per_dispatch=SAMPLES_PER_DISPATCH, graphs=not args.no_graphs,
                              time_trace=not args.untimed_trace, generated_programs=not args.interpreter,
                              bvh_width=bvh["chosen"] if bvh else 4)
    t0 = time.perf_counter()
    comp = backend.compile_scene(snap)
    if not comp.ok():
        print(json.dumps({"error": [d.message for d in comp.diagnostics]}))
        return 1
    compile_s = time.perf_counter() - t0
    print(f"bench: scene compiled in {compile_s:.2f} s", file=sys.stderr, flush=True)
    spp = args.spp or w["spp"]
    pixels = settings.width * settings.height

    session = backend.create_session(comp.scene, settings)
    tuning = None
    if args.tuning:
        # an explicit launch shape, e.g. "tile_width=16,sort_shading=1"
        session.set_tuning({k: int(v) for k, v in (kv.split("=") for kv in args.tuning.split(","))})
    elif not args.no_autotune:
        # session set-up, like the compile: time candidate launch shapes on a
        # 4-sample probe and keep the fastest (the film does not depend on it)
        t0 = time.perf_counter()
        rep = session.autotune(4)
        tuning = {"chosen": rep.chosen.as_dict(), "probe_ms_default": rep.default_ms,
                  "probe_ms_chosen": rep.chosen_ms, "candidates": rep.candidates,
                  "probe_samples": rep.probe_samples, "seconds": time.perf_counter() - t0}
        print(f"bench: autotune chose {tuning['chosen']} ({rep.chosen_ms:.2f} ms vs default "
              f"{rep.default_ms:.2f} ms per probe) in {tuning['seconds']:.2f} s", file=sys.stderr, flush=True)

    def step(sess):
        # a fresh frame: reset zeroes the resident film, then the whole
        # sample range is dispatched and waited for
        sess.reset(settings)
        sess.render_device(0, spp, spp)
        return sess.stats()

    for i in range(args.warmup):
        st = step(session)
        print(f"bench: warmup {i} {st.kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    rays = 0
    trace_ms = 0.0
    trace_launches = 0
    graph_launches = 0
    for i in range(args.steps):
        st = step(session)
        rays += st.rays
        trace_ms += st.trace_ms
        trace_launches += st.trace_launches
        graph_launches += st.graph_launches
        print(f"bench: step {i} {st.kernel_ms:.1f} ms (trace {st.trace_ms:.1f} ms), {st.rays / 1e6:.1f} Mrays",
              file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    samples = pixels * spp * args.steps
    value = samples / wall
    mrays = rays / wall / 1e6
    trace_launch_ms = trace_ms / max(trace_launches, 1)
    rays_per_launch = rays / max(trace_launches, 1)
    timed = trace_launches > 0 and trace_launch_ms > 0.0
    achieved = rays_per_launch * b_ray / (trace_launch_ms * 1e-3) / 1e9 if timed else 0.0
    traffic_gbs = (rays_per_launch * w["traffic"] / (trace_launch_ms * 1e-3) / 1e9
                   if timed and w["traffic"] else None)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": w["describe"].format(spp=spp),
            "integrator": "wavefront (batched)",
            "samples_per_dispatch": SAMPLES_PER_DISPATCH,
            "launch": "graph replay per dispatch" if graph_launches else "host launches",
            "shading": "closure interpreter" if args.interpreter else "generated surface programs (hipRTC)",
            "tuning": tuning if tuning else session.tuning.as_dict(),
            "bvh_width": bvh if bvh else 4,
        },
        "mrays_per_s": mrays,
        "rays_per_sample": rays / samples,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": rays_per_launch * w["traffic"] if w["traffic"] else None,
            # `achieved` prices each ray at the algorithmic binary-BVH bytes
            # B_ray; the measured HBM bytes of the kernel (FETCH_SIZE and
            # WRITE_SIZE per ray) over the same launch time give the
            # bandwidth the kernel really draws
            "basis": "achieved: algorithmic B_ray = 48 + 32 N_node + 36 N_tri bytes per ray (binary SAH BVH); "
                     "traffic: measured HBM bytes per launch",
            "traffic_gbs": traffic_gbs,
            "traffic_frac": traffic_gbs / HBM_PEAK_GBS if traffic_gbs else None,
            "kernel": "wf_trace_kernel",
            "launch_ms": trace_launch_ms,
            "launches": trace_launches,
            "bytes_per_ray": b_ray,
        },
    }
    if not timed:
        result["roofline"] = None
    if not args.no_plain:
        # the plain per-request path on one frame of the same workload
        plain = RendererBackend(0, integrator=abi.INTEGRATOR_MEGAKERNEL, max_samples_per_dispatch=2)
        pcomp = plain.compile_scene(snap)
        psess = plain.create_session(pcomp.scene, settings)
        step(psess)
        t1 = time.perf_counter()
        pst = step(psess)
        pwall = time.perf_counter() - t1
        film_bytes = pixels * FILM_BYTES_PER_PIXEL_DISPATCH
        pl_ms = pst.kernel_ms / max(pst.dispatches, 1)
        pl_achieved = (pst.rays / max(pst.dispatches, 1) * b_ray + film_bytes) / (pl_ms * 1e-3) / 1e9
        result["plain"] = {
            "integrator": "megakernel",
            "samples_per_s": pixels * spp / pwall,
            "mrays_per_s": pst.rays / pwall / 1e6,
            "ms_per_step": pwall * 1e3,
            "launch_ms": pl_ms,
            "roofline_frac": pl_achieved / HBM_PEAK_GBS,
        }
        result["batched_vs_plain"] = value / result["plain"]["samples_per_s"]
        print(f"bench: plain megakernel frame {pwall * 1e3:.1f} ms", file=sys.stderr, flush=True)
        del psess, pcomp, plain
    if not args.no_cpu_baseline:
        print("bench: timing the CPU oracle baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(snap, settings, spp)
    if args.serving_requests > 0 and args.workload == "instanced":
        # the metric's latency half: the serving stream, batched and plain
        del session, comp, backend
        result["serving"] = serving_summary(serving_leg(
            args.serving_requests, not args.no_plain,
            lambda msg: print(f"bench: {msg}", file=sys.stderr, flush=True)))
    print(json.dumps(result))
    if args.calibrate_copy:
        calibration_copy()
    return 0


if __name__ == "__main__":
    sys.exit(main())
