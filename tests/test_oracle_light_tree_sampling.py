"""Light-tree selection in the CPU oracle (src/luisa/path_tracer_light_tree.cpp).

Invariants that pin the restatement beyond the construction test: every
selection's pdf equals the reverse pdf of the emitter it chose (the descent
and the climb multiply the same branch probabilities), selection frequencies
follow the reverse pdfs, the probability of selecting nothing (a descent that
ends where no emitter is important) is one minus their sum, and renders with
the tree agree with renders using the flat CDF.
"""
import math

import numpy as np

import oracle_binding as ob
from psycles_amd import abi, scenes


def tree_scene():
    snap = scenes.cornell_lamps()
    # more lamps so that the tree has several levels
    rng = np.random.default_rng(7)
    for i in range(24):
        pos = rng.uniform(-0.8, 0.8, 3)
        snap.lights.append(scenes.lamp(abi.LIGHT_POINT, scenes.translation(*pos), tuple(rng.uniform(0.2, 1.0, 3)),
                                       float(rng.uniform(0.5, 5.0)), size=0.02))
    return snap


def test_pdfs_sum_to_one_and_selection_matches():
    osc = ob.OracleScene(tree_scene())
    n = osc.tree_emitters()
    assert n > 20
    rng = np.random.default_rng(3)
    for _ in range(20):
        point = rng.uniform(-0.9, 0.9, 3)
        normal = rng.normal(size=3)
        normal /= np.linalg.norm(normal)
        pdfs = np.array([osc.tree_pdf(i, point, normal, False) for i in range(n)])
        assert pdfs.sum() <= 1.0 + 2e-5
        counts = np.zeros(n + 1)
        m = 4000
        for k in range(m):
            eid, ps, pr = osc.tree_query(point, normal, False, (k + 0.5) / m)
            if eid == 0xFFFFFFFF:
                counts[n] += 1
                continue
            assert eid < n
            assert abs(ps - pr) <= 2e-6 * max(ps, pr), (ps, pr)
            assert abs(ps - pdfs[eid]) <= 2e-6 * pdfs[eid]
            counts[eid] += 1
        expected = np.append(pdfs, 1.0 - pdfs.sum())
        # stratified u: frequencies follow the pdfs
        assert np.abs(counts / m - expected).max() < 0.01


def test_tree_and_flat_renders_agree():
    snap = tree_scene()
    means = []
    for use_tree in (0, 1):
        s = scenes.cornell_settings(32, 20)
        s.integrator.use_light_tree = use_tree
        film, _ = ob.OracleScene(snap).render(s, 0, 256, 256)
        means.append(float((film.combined[..., :3] / film.sample_count[..., None]).mean()))
    assert abs(means[0] - means[1]) < 0.02 * means[0], means


def reference_contract_tree():
    """The four emitters of tests/test_luisa_light_tree.cpp:165-219: three
    local boxes at x = -1.5, 0, 1.5 (z = 2, energies 1, 2, 4, cones about
    -z with theta_o = pi, theta_e = pi/2) and one distant emitter about +z
    (theta_o 0, theta_e 0.1, energy 0.75)."""
    out = []
    for i, (x, energy) in enumerate(((-1.5, 1.0), (0.0, 2.0), (1.5, 4.0))):
        e = ob.TreeEmitter()
        e.bounds_min[:] = (x - 0.2, -0.2, 1.8)
        e.bounds_max[:] = (x + 0.2, 0.2, 2.2)
        e.bounds_empty = 0
        e.axis[:] = (0.0, 0.0, -1.0)
        e.theta_o, e.theta_e, e.orientation_empty = math.pi, math.pi / 2, 0
        e.energy = energy
        e.centroid[:] = (x, 0.0, 2.0)
        e.emitter_id, e.distant = i, 0
        out.append(e)
    d = ob.TreeEmitter()
    d.bounds_empty = 1
    d.axis[:] = (0.0, 0.0, 1.0)
    d.theta_o, d.theta_e, d.orientation_empty = 0.0, 0.1, 0
    d.energy = 0.75
    d.centroid[:] = (0.0, 0.0, 1.0)
    d.emitter_id, d.distant = 3, 1
    out.append(d)
    return out


def test_reference_device_tree_contract():
    """tests/test_luisa_light_tree.cpp:272-447 on the oracle: 8192 stratified
    descents from the origin facing +z always select an emitter whose
    selection pdf equals its reverse pdf (2e-5), frequencies follow the
    reverse pdfs (0.015), the pdfs sum to one (2e-5); from (0, 0, 3) facing
    +z every emitter is behind the surface unless it transmits."""
    tree = ob.BuiltTree(reference_contract_tree())
    origin, up = (0.0, 0.0, 0.0), (0.0, 0.0, 1.0)
    n = 8192
    freq = np.zeros(4)
    for i in range(n):
        eid, ps, pr = tree.sample(origin, up, False, (i + 0.5) / n)
        assert eid < 4 and ps > 0.0 and abs(ps - pr) <= 2.0e-5, (i, eid, ps, pr)
        freq[eid] += 1
    pdfs = np.array([tree.pdf(i, origin, up, False) for i in range(4)])
    assert abs(pdfs.sum() - 1.0) <= 2.0e-5
    assert np.all(np.abs(freq / n - pdfs) <= 0.015), (freq / n, pdfs)
    behind, _, _ = tree.sample((0.0, 0.0, 3.0), up, False, 0.25)
    assert behind == 0xFFFFFFFF
    transmitted, _, _ = tree.sample((0.0, 0.0, 3.0), up, True, 0.25)
    assert transmitted < 4
