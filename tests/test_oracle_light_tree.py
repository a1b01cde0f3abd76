"""Light tree construction in the CPU oracle.

Pinned by the reference's own construction test (tests/test_light_tree.cpp:
cone union, distant branch, root energy, leaf capacity, identity checks) and,
where the reference sources compile here, bit-identical against
build_cycles_light_tree (src/sampling/light_tree.cpp) on random populations:
same nodes, same emitter order inside every leaf, same reverse mappings.
"""
import math

import numpy as np
import pytest

import oracle_binding as ob


def emitter(identity, position, energy, distant=False, axis=(0.0, 0.0, 1.0), theta_o=None, theta_e=None):
    e = ob.TreeEmitter()
    for i in range(3):
        e.bounds_min[i] = e.bounds_max[i] = position[i]
        e.centroid[i] = position[i]
        e.axis[i] = axis[i]
    e.bounds_empty = 1 if distant else 0
    e.theta_o = (0.0 if distant else math.pi) if theta_o is None else theta_o
    e.theta_e = (0.1 if distant else math.pi / 2) if theta_e is None else theta_e
    e.orientation_empty = 0
    e.energy = energy
    e.emitter_id = identity
    e.distant = 1 if distant else 0
    return e


def node_view(tree):
    n = tree["nodes"].reshape(-1, 20)
    return {"measure": n[:, :12].view(np.float32), "flags": n[:, 12:14], "parent": n[:, 14], "left": n[:, 15],
            "right": n[:, 16], "first": n[:, 17], "count": n[:, 18], "kind": n[:, 19]}


def test_reference_construction_case():
    """tests/test_light_tree.cpp:65-118."""
    inputs = [emitter(i, (float(i), float((i * 7) % 3), 0.0), float(i + 1)) for i in range(10)]
    inputs.append(emitter(10, (0.0, 0.0, 1.0), 4.0, distant=True))
    inputs.append(emitter(11, (0.0, 0.0, -1.0), 2.0, distant=True))
    tree = ob.light_tree(inputs, 3)
    nodes = node_view(tree)
    assert tree["root"] == 0
    assert nodes["kind"][0] == 0
    right = nodes["right"][0]
    assert nodes["kind"][right] == 2 and nodes["count"][right] == 2
    assert abs(float(nodes["measure"][0][11]) - 61.0) < 1e-6
    ids = tree["emitters"].reshape(-1, 19)[:, 17]
    assert sorted(ids.tolist()) == list(range(12))
    for tree_index, identity in enumerate(ids):
        assert tree["to_tree"][identity] == tree_index
        leaf = tree["to_leaf"][identity]
        assert nodes["kind"][leaf] == 2 or nodes["count"][leaf] <= 3
        assert nodes["first"][leaf] <= tree_index < nodes["first"][leaf] + nodes["count"][leaf]


def test_reference_cone_union():
    """tests/test_light_tree.cpp:55-63: cones about x (0.25) and y (0.5)."""
    a = emitter(0, (0.0, 0.0, 0.0), 1.0, axis=(1.0, 0.0, 0.0), theta_o=0.0, theta_e=0.25)
    b = emitter(1, (1.0, 0.0, 0.0), 1.0, axis=(0.0, 1.0, 0.0), theta_o=0.0, theta_e=0.5)
    tree = ob.light_tree([a, b], 8)
    m = node_view(tree)["measure"][0]
    assert abs(m[9] - 0.78539816339) < 1e-6
    assert abs(m[6] - 0.70710677) < 1e-6 and abs(m[7] - 0.70710677) < 1e-6
    assert m[10] == 0.5


def test_identity_checks():
    assert ob.light_tree([emitter(7, (0, 0, 0), 1.0)], 8) is None
    zero = ob.light_tree([emitter(0, (0, 0, 0), 0.0)], 8)
    assert zero is not None
    assert float(node_view(zero)["measure"][zero["root"]][11]) == 0.0


def random_population(rng, n, distant_fraction=0.05):
    out = []
    for i in range(n):
        distant = rng.random() < distant_fraction
        pos = rng.normal(0.0, 10.0, 3)
        size = rng.exponential(0.5, 3)
        e = ob.TreeEmitter()
        for k in range(3):
            e.bounds_min[k] = pos[k] - size[k]
            e.bounds_max[k] = pos[k] + size[k]
            e.centroid[k] = pos[k]
        axis = rng.normal(0.0, 1.0, 3)
        if rng.random() < 0.1:
            axis = np.zeros(3)
        for k in range(3):
            e.axis[k] = axis[k]
        e.bounds_empty = 1 if distant else 0
        e.theta_o = float(rng.choice([0.0, math.pi / 2, math.pi, rng.random() * math.pi]))
        e.theta_e = float(rng.choice([math.pi / 2, rng.random() * math.pi]))
        e.orientation_empty = 0
        e.energy = float(rng.choice([0.0, rng.exponential(5.0), rng.exponential(50.0)]))
        e.emitter_id = i
        e.distant = 1 if distant else 0
        out.append(e)
    perm = rng.permutation(n)
    return [out[j] for j in perm]


@pytest.mark.parametrize("n,leaf", [(1, 8), (2, 1), (37, 3), (500, 8), (2000, 8), (3000, 1)])
def test_matches_reference_build(n, leaf):
    ref = ob.ref_library()
    if ref is None or not hasattr(ref, "ref_light_tree_build"):
        pytest.skip("reference light tree sources not compiled")
    rng = np.random.default_rng(n * 31 + leaf)
    pop = random_population(rng, n)
    mine = ob.light_tree(pop, leaf)
    theirs = ob.light_tree(pop, leaf, library=ref)
    assert mine is not None and theirs is not None
    assert mine["root"] == theirs["root"]
    for key in ("nodes", "emitters", "to_tree", "to_leaf"):
        assert np.array_equal(mine[key], theirs[key]), key


def test_scene_upload_contract():
    """tests/test_luisa_light_tree.cpp:49-161: one front-sampled emissive
    triangle (0,0,3) (2,0,3) (0,2,3) whose instance moves it +10 in x and
    overrides its material with emission (0.25, 0.5, 1) x 4, plus one point
    lamp.  The tree is usable, the flat distribution has two entries of
    selection pdf 0.5, and the triangle's tree record carries world bounds
    x in [10, 12], z = 3 and energy area x mean emission = 2 x 7/3."""
    from psycles_amd import abi, scenes
    b = scenes.ProgramBuilder()
    base = b.build(b.emission((1.0, 1.0, 1.0), 0.25), emission_sampling=abi.EMISSION_SAMPLING_FRONT)
    b = scenes.ProgramBuilder()
    override = b.build(b.emission((0.25, 0.5, 1.0), 4.0), emission_sampling=abi.EMISSION_SAMPLING_FRONT)
    mesh = scenes.MeshData(np.array([[0, 0, 3], [2, 0, 3], [0, 2, 3]], np.float32), np.array([[0, 1, 2]], np.uint32),
                           [0])
    inst = scenes.InstanceData(0, scenes.translation(10.0, 0.0, 0.0), material_overrides=[1])
    cam = scenes.CameraData(scenes.translation(0.0, 0.0, 10.0))
    snap = scenes.SceneSnapshot([base, override], [mesh], [inst], cam,
                                lights=[scenes.lamp(abi.LIGHT_POINT, scenes.translation(20.0, 0.0, 2.0))])
    osc = ob.OracleScene(snap)
    _, pdf, kind, _ = osc.distribution()
    assert list(kind) == [0, 1]
    assert np.allclose(pdf, [0.5, 0.5], atol=1e-6)
    tree = osc.light_tree()
    assert tree is not None and osc.tree_emitters() == 2
    records = tree["emitters"].reshape(-1, 19)
    tri = records[records[:, 17] == 0][0]
    m = tri[:12].view(np.float32)
    assert abs(m[0] - 10.0) <= 1e-6 and abs(m[2] - 3.0) <= 1e-6
    assert abs(m[3] - 12.0) <= 1e-6 and abs(m[5] - 3.0) <= 1e-6
    assert abs(m[11] - 2.0 * (1.0 + 2.0 + 4.0) / 3.0) <= 1e-5
