"""The uneven-path batch generator and its CPU oracle render.

The stream is deterministic (SplitMix64 seed 5) and draws what the batch
specifies: 1-16 slabs with transparency in [0.5, 0.95], a square side in
128..1024, 1..256 samples, a homogeneous box on about half the requests.
On the oracle, a stack of black slabs over nothing but the world transmits
world x the product of the transparencies in expectation (the closed form
of the transparent mix's stochastic lobe choice).
"""
import numpy as np

import oracle_binding as ob
from psycles_amd import abi, scenes


def test_stream_is_deterministic_and_in_range():
    a = scenes.uneven_batch(64)
    b = scenes.uneven_batch(64)
    assert [(r.slabs, r.side, r.spp, r.volume) for r in a] == [(r.slabs, r.side, r.spp, r.volume) for r in b]
    full = scenes.uneven_batch()
    assert len(full) == 512
    assert min(r.slabs for r in full) == 1 and max(r.slabs for r in full) == 16
    assert 128 <= min(r.side for r in full) and max(r.side for r in full) <= 1024
    assert 1 <= min(r.spp for r in full) and max(r.spp for r in full) <= 256
    assert 0.4 < np.mean([r.volume for r in full]) < 0.6
    assert all(0.1 <= r.sigma_t <= 2.0 for r in full)
    for r in full[:16]:
        slabs = [m for m in r.snapshot.materials if m.name.startswith("slab")]
        assert len(slabs) == r.slabs
        for m in slabs:
            assert 0.5 <= float(m.parameters[-1]) <= 0.95


def test_black_slab_stack_transmits_the_product():
    ts = [0.55, 0.8, 0.93, 0.6, 0.75]
    snap = scenes.uneven_slabs(ts, [(0.0, 0.0, 0.0)] * (3 + len(ts)), [(0.0, -2.0, 0.0)] * 3)
    snap.instances = [i for i in snap.instances if i.name.startswith("slab")]
    snap.lights = []
    s = scenes.uneven_settings(24, 0)
    film, _ = ob.OracleScene(snap).render(s, 0, 64, 64)
    c = film.combined[..., :3] / film.sample_count[..., None]
    expect = 0.3 * np.prod(ts)
    # The lobe choice is stochastic (transparent with probability t, weight
    # 1), so the product holds in expectation: the upper 15 rows cross all
    # five slabs; 360 pixels x 64 samples put the mean within 5 %.
    region = c[:15]
    assert abs(region.mean() / expect - 1.0) < 0.05
    # a black slab absorbs whatever takes its diffuse lobe: no pixel exceeds
    # the world seen through the nearest slab alone
    assert c.max() <= 0.3 + 1e-6
