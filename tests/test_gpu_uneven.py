"""The uneven-path batch on the GPU: stacked transparent slabs (1-16 per
request, transparency U[0.5, 0.95]) over diffuse geometry, every request its
own scene, resolution and sample count (SplitMix64 seed 5).

The first requests of the batch, at reduced resolution and sample count,
submitted together to the render server: each film bit-identical to the same
request rendered alone in a wavefront session, and within 1e-4 relative of
the CPU oracle on >= 99.9 % of pixels (Combined / Normal / Albedo, sample
counts identical).  The first 64 at full resolution (at most 16 samples):
bit-identical alone and in the mix, the smallest against the oracle.  The optional homogeneous box of a request is not built
(no volume transport on this path).
"""
import sys

import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import abi, scenes
from psycles_amd.render import RendererBackend
from psycles_amd.server import RenderServer

pytestmark = pytest.mark.gpu

REL_TOL = 1.0e-4
PIXEL_FRACTION = 0.999
REQUESTS = 10


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def relative_ok(a, b, tol=REL_TOL):
    diff = np.abs(a - b).max(axis=-1)
    scale = np.maximum(np.maximum(np.abs(a), np.abs(b)).max(axis=-1), 1e-3)
    return diff <= tol * scale


def reduced(req):
    """Side / 8 (16..128 pixels) and at most 8 samples."""
    return req.settings(max(req.side // 8, 16)), min(req.spp, 8)


def test_uneven_batch_through_server():
    batch = scenes.uneven_batch(REQUESTS)
    backend = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT, max_samples_per_dispatch=8)
    comps = [backend.compile_scene(r.snapshot).scene for r in batch]
    server = RenderServer(backend, path_capacity=1 << 17)
    submitted = []
    for req, comp in zip(batch, comps):
        s, spp = reduced(req)
        log(f"case: uneven request {req.index}: {req.slabs} slabs, {s.width}^2, {spp} spp")
        submitted.append((req, comp, s, spp, server.submit(comp, s, 0, spp, spp)))
    for req, comp, s, spp, r in submitted:
        assert r.wait(120000), req.index
    st = server.stats()
    log(f"  batches {st.batches}, paths {st.paths}, rays {st.rays}")
    assert st.completed == REQUESTS
    for req, comp, s, spp, r in submitted:
        film = r.read_film()
        sess = backend.create_session(comp, s)
        sess.render_device(0, spp, spp)
        alone = sess.read_film()
        for x, y in zip(film, alone):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), req.index
        ref, _ = ob.OracleScene(req.snapshot).render(s, 0, spp, spp)
        assert np.array_equal(film[3], ref.sample_count)
        for name, a, b in (("combined", film[0], ref.combined), ("normal", film[1], ref.normal),
                           ("albedo", film[2], ref.albedo)):
            frac = relative_ok(a, b).mean()
            log(f"  request {req.index} {name}: {frac:.6f}")
            assert frac >= PIXEL_FRACTION, (req.index, name)
    server.close()


def test_uneven_megakernel_matches_oracle():
    """The deepest stacks of the first requests through the megakernel."""
    batch = sorted(scenes.uneven_batch(REQUESTS), key=lambda r: -r.slabs)[:2]
    backend = RendererBackend(0)
    for req in batch:
        s, spp = reduced(req)
        log(f"case: uneven request {req.index} (megakernel): {req.slabs} slabs, {s.width}^2, {spp} spp")
        comp = backend.compile_scene(req.snapshot).scene
        sess = backend.create_session(comp, s)
        sess.render_device(0, spp, spp)
        g = sess.read_film()
        ref, _ = ob.OracleScene(req.snapshot).render(s, 0, spp, spp)
        assert np.array_equal(g[3], ref.sample_count)
        for name, a, b in (("combined", g[0], ref.combined), ("normal", g[1], ref.normal),
                           ("albedo", g[2], ref.albedo)):
            frac = relative_ok(a, b).mean()
            log(f"  {name}: {frac:.6f}")
            assert frac >= PIXEL_FRACTION, name


def test_uneven_full_size_batch_identity():
    """The first 64 requests at full resolution (128-1024 pixels square),
    at most 16 samples each, through a server with default options (arenas
    of 8 M paths, batches filling them): every film bit-identical to the
    request rendered alone, every pixel's sample count the request's, and
    the three smallest requests within the film bar of the oracle."""
    batch = scenes.uneven_batch(64)
    backend = RendererBackend(0, integrator=abi.INTEGRATOR_WAVEFRONT)
    comps = [backend.compile_scene(r.snapshot).scene for r in batch]
    server = RenderServer(backend)
    submitted = []
    for req, comp in zip(batch, comps):
        s, spp = req.settings(), min(req.spp, 16)
        submitted.append((req, comp, s, spp, server.submit(comp, s, 0, spp, spp)))
    log(f"case: first 64 uneven requests at full size, {sum(s.width * s.height * n for _, _, s, n, _ in submitted)} "
        "samples")
    for req, comp, s, spp, r in submitted:
        assert r.wait(300000), req.index
    st = server.stats()
    log(f"  batches {st.batches}, paths {st.paths}, rays {st.rays}")
    smallest = sorted(submitted, key=lambda e: e[2].width)[:3]
    for req, comp, s, spp, r in submitted:
        film = r.read_film()
        assert np.all(film[3] == spp), req.index
        sess = backend.create_session(comp, s)
        sess.render_device(0, spp, spp)
        alone = sess.read_film()
        sess.close()
        for x, y in zip(film, alone):
            assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), req.index
        if any(req is e[0] for e in smallest):
            ref, _ = ob.OracleScene(req.snapshot).render(s, 0, spp, spp)
            for name, a, b in (("combined", film[0], ref.combined), ("normal", film[1], ref.normal),
                               ("albedo", film[2], ref.albedo)):
                frac = relative_ok(a, b).mean()
                log(f"  request {req.index} ({s.width}^2, {spp} spp) {name}: {frac:.6f}")
                assert frac >= PIXEL_FRACTION, (req.index, name)
        r.release()
    server.close()
