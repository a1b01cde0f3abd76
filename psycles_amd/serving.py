"""Serving workload: a Poisson request stream over resident scenes.

The stream: exponential inter-arrival times at `rate` requests/s
(SplitMix64 seed 4).  Each request is, with probability 0.3, a
parameter-only edit of a random material of a random scene (the diffuse
albedo, or the emission strength of an emitting material, scaled by
U[0.8, 1.2], albedo clamped to 1), and otherwise a render of a random scene
at a resolution drawn from {256x256, 640x400, 1280x720, 1920x1080} and a
sample count from {1, 4, 16, 64}, with seed = request id.  Viewport requests
are the renders at 256x256 or 640x400 with at most 4 samples.

`run_batched` replays the stream against a RenderServer (continuous
batching); `run_plain` replays it one request at a time through megakernel
sessions, the plain per-request GPU path.  Both report per-request latency
(completion minus arrival) and whole-stream samples/s.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import abi, scenes

RESOLUTIONS = ((256, 256), (640, 400), (1280, 720), (1920, 1080))
SAMPLE_COUNTS = (1, 4, 16, 64)
VIEWPORT_PIXELS = 640 * 400
VIEWPORT_SAMPLES = 4


@dataclass
class Request:
    index: int
    arrival: float              # seconds from the start of the stream
    scene: str
    edit: bool = False
    material: int = 0
    parameters: np.ndarray | None = None
    width: int = 0
    height: int = 0
    spp: int = 0

    @property
    def samples(self) -> int:
        return 0 if self.edit else self.width * self.height * self.spp

    @property
    def viewport(self) -> bool:
        return not self.edit and self.width * self.height <= VIEWPORT_PIXELS and self.spp <= VIEWPORT_SAMPLES


@dataclass
class Workload:
    snapshots: dict
    settings: dict                      # scene -> settings factory(width, height)
    requests: list = field(default_factory=list)


def scene_settings(name: str, width: int, height: int) -> abi.RenderSettings:
    if name == "cornell":
        return scenes.cornell_settings(width, height)
    if name == "instanced":
        return scenes.instanced_settings(width, height)
    return scenes.many_light_settings(width, height)


def edited_parameters(instructions, parameters: np.ndarray, factor: float) -> np.ndarray:
    """Emission strength (emitting materials) or diffuse albedo scaled by
    `factor`; albedo stays within [0, 1]."""
    p = np.array(parameters, dtype=np.float32, copy=True)
    emission = [i for i in instructions if i[0] == abi.CLOSURE_EMISSION]
    if emission:
        for ins in emission:
            p[ins[5]] *= factor
        return p
    for ins in instructions:
        if ins[0] == abi.CLOSURE_DIFFUSE:
            c = ins[4]
            p[c:c + 3] = np.minimum(p[c:c + 3] * factor, 1.0)
    return p


def make_stream(snapshots: dict, count: int, rate: float = 200.0, seed: int = 4,
                edit_fraction: float = 0.3) -> list:
    """`count` requests of the serving stream over `snapshots` (name ->
    SceneSnapshot, in a fixed order).  Each edit scales the values left by
    the edits before it; the snapshots themselves are not modified."""
    rng = scenes.SplitMix64(seed)
    names = list(snapshots)
    current = {n: [m.parameters for m in snapshots[n].materials] for n in names}
    t = 0.0
    out = []
    for i in range(count):
        t += -math.log(1.0 - rng.uniform()) / rate
        name = names[min(int(rng.uniform() * len(names)), len(names) - 1)]
        if rng.uniform() < edit_fraction:
            snap = snapshots[name]
            m = min(int(rng.uniform() * len(snap.materials)), len(snap.materials) - 1)
            factor = 0.8 + 0.4 * rng.uniform()
            params = edited_parameters(snap.materials[m].instructions, current[name][m], factor)
            current[name][m] = params
            out.append(Request(i, t, name, edit=True, material=m, parameters=params))
        else:
            w, h = RESOLUTIONS[min(int(rng.uniform() * 4), 3)]
            spp = SAMPLE_COUNTS[min(int(rng.uniform() * 4), 3)]
            out.append(Request(i, t, name, width=w, height=h, spp=spp))
    return out


def _percentiles(values):
    if not values:
        return {"p50": None, "p99": None, "count": 0}
    v = np.asarray(values, dtype=np.float64)
    return {"p50": float(np.percentile(v, 50)), "p99": float(np.percentile(v, 99)), "count": int(v.size)}


def summarize(stream, latencies_ms: dict, makespan_s: float, rays: int | None = None) -> dict:
    renders = [r for r in stream if not r.edit]
    samples = sum(r.samples for r in renders)
    view = [latencies_ms[r.index] for r in renders if r.viewport]
    final = [latencies_ms[r.index] for r in renders if not r.viewport]
    out = {
        "samples_per_s": samples / makespan_s,
        "makespan_s": makespan_s,
        "samples": samples,
        "requests": len(renders),
        "edits": len(stream) - len(renders),
        "latency_ms": _percentiles([latencies_ms[r.index] for r in renders]),
        "viewport_latency_ms": _percentiles(view),
        "final_latency_ms": _percentiles(final),
    }
    if rays is not None:
        out["mrays_per_s"] = rays / makespan_s / 1e6
    return out


def run_batched(server, compiled: dict, stream, log=None) -> dict:
    """Submits each request at its arrival time (edits in order with them),
    waits for all, and measures latency on the server clock."""
    handles = {}
    keep = []
    start = time.perf_counter()
    for r in stream:
        delay = r.arrival - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        if r.edit:
            server.update_parameters(compiled[r.scene], r.material, r.parameters)
            continue
        s = scene_settings(r.scene, r.width, r.height)
        s.seed = r.index
        keep.append(s)
        handles[r.index] = server.submit(compiled[r.scene], s, 0, r.spp, r.spp)
    if log:
        log(f"batched: all {len(handles)} renders submitted after {time.perf_counter() - start:.2f} s")
    latencies, first_submit, last_done = {}, None, 0.0
    for index, h in handles.items():
        if not h.wait(-1.0):
            raise RuntimeError(f"request {index} did not complete")
        st = h.status()
        latencies[index] = st.done_ms - st.submit_ms
        first_submit = st.submit_ms if first_submit is None else min(first_submit, st.submit_ms)
        last_done = max(last_done, st.done_ms)
        h.release()
    makespan = (last_done - first_submit) * 1e-3
    stats = server.stats()
    if log:
        log(f"batched: {stats.batches} batches, {stats.paths / 1e6:.1f} M paths, busy {stats.busy_ms:.0f} ms, "
            f"peak HBM {stats.peak_hbm / 2**30:.2f} GiB")
    out = summarize(stream, latencies, makespan, stats.rays)
    out["batches"] = int(stats.batches)
    out["peak_hbm_gib"] = stats.peak_hbm / 2**30
    return out


def run_plain(backend, compiled: dict, stream, log=None) -> dict:
    """One request at a time in arrival order: a megakernel session per
    render (created, rendered, film read back), edits applied between them."""
    latencies = {}
    rays = 0
    start = time.perf_counter()
    first, last = None, 0.0
    for r in stream:
        now = time.perf_counter() - start
        if r.arrival > now:
            time.sleep(r.arrival - now)
        if r.edit:
            compiled[r.scene].update_parameters(r.material, r.parameters)
            continue
        s = scene_settings(r.scene, r.width, r.height)
        s.seed = r.index
        sess = backend.create_session(compiled[r.scene], s)
        sess.render_device(0, r.spp, r.spp)
        sess.read_film()
        rays += sess.stats().rays
        sess.close()
        done = time.perf_counter() - start
        latencies[r.index] = (done - r.arrival) * 1e3
        first = r.arrival if first is None else first
        last = done
        if log and len(latencies) % 20 == 0:
            log(f"plain: {len(latencies)} requests done at {done:.1f} s")
    out = summarize(stream, latencies, last - first, rays)
    if log:
        log(f"plain: makespan {out['makespan_s']:.2f} s")
    return out


def run_uneven_batched(server, compiled: list, batch, log=None) -> dict:
    """The uneven-path batch submitted at once (every request its own scene)
    and rendered to completion by the server; samples/s over the makespan
    on the server clock, per-request latency from submission."""
    keep, handles = [], []
    for req, comp in zip(batch, compiled):
        s = req.settings()
        keep.append(s)
        handles.append(server.submit(comp, s, 0, req.spp, req.spp))
    latencies, first_submit, last_done = [], None, 0.0
    for n, (req, h) in enumerate(zip(batch, handles)):
        if not h.wait(-1.0):
            raise RuntimeError(f"request {req.index} did not complete")
        if log and (n + 1) % 32 == 0:
            log(f"uneven batched: {n + 1} of {len(batch)} requests done")
        st = h.status()
        latencies.append(st.done_ms - st.submit_ms)
        first_submit = st.submit_ms if first_submit is None else min(first_submit, st.submit_ms)
        last_done = max(last_done, st.done_ms)
        h.release()
    makespan = (last_done - first_submit) * 1e-3
    stats = server.stats()
    samples = sum(r.side * r.side * r.spp for r in batch)
    if log:
        log(f"uneven batched: {stats.batches} batches, {stats.rays / 1e9:.2f} G rays, {makespan:.2f} s")
    return {"samples_per_s": samples / makespan, "mrays_per_s": stats.rays / makespan / 1e6,
            "makespan_s": makespan, "samples": samples, "requests": len(batch), "batches": int(stats.batches),
            "latency_ms": _percentiles(latencies), "peak_hbm_gib": stats.peak_hbm / 2**30}


def run_uneven_plain(backend, compiled: list, batch, log=None) -> dict:
    """The same batch one request at a time through megakernel sessions."""
    rays = 0
    start = time.perf_counter()
    latencies = []
    for n, (req, comp) in enumerate(zip(batch, compiled)):
        s = req.settings()
        sess = backend.create_session(comp, s)
        sess.render_device(0, req.spp, req.spp)
        sess.read_film()
        rays += sess.stats().rays
        sess.close()
        latencies.append((time.perf_counter() - start) * 1e3)
        if log and (n + 1) % 32 == 0:
            log(f"uneven plain: {n + 1} of {len(batch)} requests done, {time.perf_counter() - start:.1f} s")
    makespan = time.perf_counter() - start
    samples = sum(r.side * r.side * r.spp for r in batch)
    if log:
        log(f"uneven plain: {makespan:.2f} s")
    return {"samples_per_s": samples / makespan, "mrays_per_s": rays / makespan / 1e6, "makespan_s": makespan,
            "samples": samples, "latency_ms": _percentiles(latencies)}
