"""Scene-level tuning: the BVH branching factor.

`RenderSession.autotune` (psx_session_autotune) picks a session's wavefront
launch shape.  The BVH's branching factor is fixed when a scene compiles
(PSX_BACKEND_BVH_WIDTH_*), so it is tuned here by compiling the scene at
each width and timing the same probe dispatches: device time of
`probe_samples` samples over the full window, the fastest of `repeats`
after a warm-up.  A width replaces the default of four only when it is at
least 1 % faster, the session autotuner's rule.  Every width renders the
same film; only traversal time differs.
"""
from __future__ import annotations

from . import abi
from .render import RendererBackend

DEFAULT_WIDTH = 4


def probe_ms(snapshot, settings, width: int, probe_samples: int = 4, repeats: int = 3, device: int = 0,
             **backend_options) -> float:
    """Device milliseconds of one `probe_samples` dispatch on a scene
    compiled with `width` children per BVH node."""
    backend = RendererBackend(device, integrator=abi.INTEGRATOR_WAVEFRONT,
                              max_samples_per_dispatch=probe_samples, bvh_width=width, **backend_options)
    try:
        comp = backend.compile_scene(snapshot)
        if not comp.ok():
            raise ValueError("; ".join(d.message for d in comp.diagnostics))
        session = backend.create_session(comp.scene, settings)
        session.render_device(0, probe_samples, probe_samples)
        best = float("inf")
        for _ in range(repeats):
            session.render_device(0, probe_samples, probe_samples)
            best = min(best, session.stats().kernel_ms)
        session.close()
        comp.scene.close()
        return best
    finally:
        backend.close()


def tune_bvh_width(snapshot, settings, probe_samples: int = 4, widths=(4, 3, 2), repeats: int = 3,
                   device: int = 0, **backend_options) -> dict:
    """{"chosen": width, "ms": {width: probe ms}} over `widths`."""
    ms = {w: probe_ms(snapshot, settings, w, probe_samples, repeats, device, **backend_options) for w in widths}
    chosen = DEFAULT_WIDTH if DEFAULT_WIDTH in ms else min(ms, key=ms.get)
    for w, t in sorted(ms.items(), key=lambda kv: kv[1]):
        if w != chosen and t < 0.99 * ms[chosen]:
            chosen = w
        break
    return {"chosen": chosen, "ms": ms}
