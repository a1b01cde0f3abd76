"""Benchmark of the MI355X render path on one GPU.

One step = one full frame of the instanced-icosphere workload: a 1,003,520-
triangle displaced icosphere instanced 64 times, 1920x1080 at 64 spp, max 8
bounces, mix(diffuse, transparent) materials, constant world background.
Inputs (scene tables, BVH, film) are resident in HBM before the timed region.

Prints one JSON line with samples/s (value), Mrays/s, the HBM roofline of the
render kernel and the CPU oracle timed on a bounded sample of the same
workload on the host cores.
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

# Algorithmic bytes per ray, B_ray = 48 + 32 * N_node + 36 * N_tri, with the
# mean binary-SAH node visits and triangle tests per ray measured by the CPU
# oracle's BVH (16 bins, leaves <= 4) over the full frame of this workload.
ORACLE_NODES_PER_RAY = 46.83
ORACLE_TRIS_PER_RAY = 8.59
BYTES_PER_RAY = 48.0 + 32.0 * ORACLE_NODES_PER_RAY + 36.0 * ORACLE_TRIS_PER_RAY
FILM_BYTES_PER_PIXEL_DISPATCH = 104.0
HBM_PEAK_GBS = 8000.0


def cpu_baseline(snap, settings, seconds_target=15.0):
    """Time the CPU oracle on the full frame of the same workload: a 1 spp
    calibration pass, then as many consecutive samples as fill the target
    duration, on the host cores."""
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_binding as ob

    threads = min(os.cpu_count() or 1, 16)
    scene = ob.OracleScene(snap)
    t0 = time.perf_counter()
    _, st = scene.render(settings, 0, 1, 64, threads=threads)
    dt = max(time.perf_counter() - t0, 1e-3)
    spp = int(max(1, min(63, round(seconds_target / dt))))
    t0 = time.perf_counter()
    _, st = scene.render(settings, 1, spp, 64, threads=threads)
    dt = time.perf_counter() - t0
    return {
        "value": st.samples / dt,
        "unit": "samples/s",
        "cores": threads,
        "kind": "port",
        "sample": f"full 1920x1080 frame, samples 1-{spp} of 64 ({st.samples} samples) in "
                  f"{dt:.1f} s, {st.rays / max(dt, 1e-9) / 1e6:.2f} Mrays/s",
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--spp", type=int, default=64)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported "
              "(every box here has one MI355X)", file=sys.stderr)
        return 2

    from psycles_amd import scenes
    from psycles_amd.render import RendererBackend

    print("bench: building instanced scene (1,003,520 tris x 64 instances)", file=sys.stderr, flush=True)
    snap = scenes.instanced_icospheres()
    settings = scenes.instanced_settings()
    backend = RendererBackend(0)
    t0 = time.perf_counter()
    comp = backend.compile_scene(snap)
    if not comp.ok():
        print(json.dumps({"error": [d.message for d in comp.diagnostics]}))
        return 1
    compile_s = time.perf_counter() - t0
    print(f"bench: scene compiled in {compile_s:.2f} s", file=sys.stderr, flush=True)
    spp = args.spp
    pixels = settings.width * settings.height

    session = backend.create_session(comp.scene, settings)

    def step():
        # a fresh frame: reset zeroes the resident film, then the whole
        # sample range is dispatched and waited for
        session.reset(settings)
        session.render_device(0, spp, spp)
        return session.stats()

    for i in range(args.warmup):
        st = step()
        print(f"bench: warmup {i} {st.kernel_ms:.1f} ms", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    rays = 0
    kernel_ms = 0.0
    dispatches = 0
    for i in range(args.steps):
        st = step()
        rays += st.rays
        kernel_ms += st.kernel_ms
        dispatches += st.dispatches
        print(f"bench: step {i} {st.kernel_ms:.1f} ms, {st.rays / 1e6:.1f} Mrays", file=sys.stderr,
              flush=True)
    wall = time.perf_counter() - t0
    samples = pixels * spp * args.steps
    value = samples / wall
    mrays = rays / wall / 1e6
    launch_ms = kernel_ms / max(dispatches, 1)
    rays_per_launch = rays / max(dispatches, 1)
    samples_per_launch = spp * args.steps / max(dispatches, 1)
    bytes_per_launch = rays_per_launch * BYTES_PER_RAY + pixels * FILM_BYTES_PER_PIXEL_DISPATCH
    achieved = bytes_per_launch / (launch_ms * 1e-3) / 1e9
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
            "workload": "instanced icospheres: 1,003,520-tri displaced icosphere x 64 instances, "
                        "mix(diffuse, transparent) materials, ground quad, constant background; "
                        f"1920x1080, {spp} spp, max 8 bounces, seed 0",
            "integrator": "megakernel",
            "samples_per_dispatch": samples_per_launch,
        },
        "mrays_per_s": mrays,
        "rays_per_sample": rays / samples,
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
            "kernel": "render_kernel",
            "launch_ms": launch_ms,
            "bytes_per_ray": BYTES_PER_RAY,
        },
    }
    if not args.no_cpu_baseline:
        print("bench: timing the CPU oracle baseline", file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(snap, settings)
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
