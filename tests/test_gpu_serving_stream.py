"""The serving stream over the full-size resident scenes: the first 30
requests of the Poisson stream (200 req/s, seed 4, 30 % parameter-only
edits) over the Cornell box, the instanced icospheres (1,003,520 triangles x
64 instances) and the many-light scene, submitted to the render server at
their arrival times.

Every render's film is bit-identical to a standalone session of the same
range on a second compilation of the scene that receives the same edits in
stream order (each result is the same alone and in the mix), and every
viewport render (256x256 or 640x400, at most 4 spp) is within the film bar
of the oracle rendering the same request with the same edits."""
import sys
import time

import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import abi, scenes, serving
from psycles_amd.render import RendererBackend
from psycles_amd.server import RenderServer

pytestmark = pytest.mark.gpu

REQUESTS = 30
REL_TOL = 1.0e-4


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relative_ok(a, b, tol=REL_TOL):
    diff = np.abs(a - b).max(axis=-1)
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)).max(axis=-1), 1e-3)
    return diff <= tol * scale


def test_serving_stream_full_scenes():
    snaps = {"cornell": scenes.cornell_box(), "instanced": scenes.instanced_icospheres(),
             "many_lights": scenes.many_lights()}
    stream = serving.make_stream(snaps, REQUESTS)
    renders = [r for r in stream if not r.edit]
    log(f"case: serving stream, {len(renders)} renders and {len(stream) - len(renders)} edits over "
        f"{stream[-1].arrival * 1e3:.0f} ms")
    backend = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT)
    served = {n: backend.compile_scene(s).scene for n, s in snaps.items()}
    server = RenderServer(backend)
    handles, keep = {}, []
    start = time.perf_counter()
    for r in stream:
        delay = r.arrival - (time.perf_counter() - start)
        if delay > 0:
            time.sleep(delay)
        if r.edit:
            server.update_parameters(served[r.scene], r.material, r.parameters)
            continue
        s = serving.scene_settings(r.scene, r.width, r.height)
        s.seed = r.index
        keep.append(s)
        handles[r.index] = server.submit(served[r.scene], s, 0, r.spp, r.spp)
    films = {}
    for index, h in handles.items():
        assert h.wait(-1.0), index
        films[index] = h.read_film()
        h.release()
    log(f"  server done in {time.perf_counter() - start:.2f} s")
    server.close()

    # standalone sessions and the oracle, with the edits replayed in order
    alone_backend = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT)
    alone = {n: alone_backend.compile_scene(s).scene for n, s in snaps.items()}
    oracles = {}
    checked_oracle = 0
    for r in stream:
        if r.edit:
            assert alone[r.scene].update_parameters(r.material, r.parameters)
            if r.scene in oracles:
                assert oracles[r.scene].update_parameters(r.material, r.parameters) == 0
            else:
                oracles.setdefault("pending", []).append(r)
            continue
        s = serving.scene_settings(r.scene, r.width, r.height)
        s.seed = r.index
        sess = alone_backend.create_session(alone[r.scene], s)
        sess.render_device(0, r.spp, r.spp)
        ref = sess.read_film()
        sess.close()
        got = films[r.index]
        for name, a, b in zip(("combined", "normal", "albedo", "count"), got, ref):
            assert np.array_equal(a.view(np.uint32), b.view(np.uint32)), (r.index, r.scene, name)
        if r.viewport:
            if r.scene not in oracles:
                osc = ob.OracleScene(snaps[r.scene])
                for e in oracles.get("pending", []):
                    if e.scene == r.scene:
                        assert osc.update_parameters(e.material, e.parameters) == 0
                oracles[r.scene] = osc
            film, _ = oracles[r.scene].render(s, 0, r.spp, r.spp)
            assert np.array_equal(got[3], film.sample_count)
            for name, g, c in (("combined", got[0], film.combined), ("normal", got[1], film.normal),
                               ("albedo", got[2], film.albedo)):
                frac = float(np.mean(relative_ok(g, c)))
                log(f"  request {r.index} ({r.scene} {r.width}x{r.height}x{r.spp}) {name} vs oracle {frac:.5f}")
                assert frac >= 0.999, (r.index, name)
            checked_oracle += 1
    log(f"  {len(renders)} films bit-identical to standalone sessions; {checked_oracle} viewport films vs oracle")
    assert checked_oracle >= 3
