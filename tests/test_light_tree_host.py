"""The library's host light-tree builder (psycles_amd/csrc/psx_light_tree.cpp)
against the oracle's build (oracle/psx_oracle_tree.cpp, itself bit-identical
to build_cycles_light_tree, src/sampling/light_tree.cpp:90-408).

Populations above 4096 emitters make the builder evaluate split axes and
build subtrees on several threads; the tables must still be word-for-word
the serial pre-order build: same topology, same emitter order in every
leaf, same measures and the same reverse mappings.
"""
import ctypes as C
import math
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import oracle_binding as ob

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "psycles_amd" / "csrc"
ROCM = Path("/opt/rocm")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    exe = tmp_path_factory.mktemp("tree") / "light_tree_driver"
    subprocess.run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", f"-I{ROCM / 'include'}",
                    "-D__HIP_PLATFORM_AMD__", f"-I{CSRC}", str(ROOT / "tests" / "host" / "light_tree_driver.cpp"),
                    str(CSRC / "psx_light_tree.cpp"), "-pthread", "-o", str(exe)], check=True)
    return exe


def population(rng, n, distant_fraction=0.03):
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
        for k in range(3):
            e.axis[k] = axis[k]
        e.bounds_empty = 1 if distant else 0
        e.theta_o = float(rng.choice([0.0, math.pi / 2, rng.random() * math.pi]))
        e.theta_e = float(rng.choice([math.pi / 2, rng.random() * math.pi]))
        e.orientation_empty = 0
        e.energy = float(rng.choice([0.0, rng.exponential(5.0), rng.exponential(50.0)]))
        e.emitter_id = i
        e.distant = 1 if distant else 0
        out.append(e)
    perm = rng.permutation(n)
    return [out[j] for j in perm]


def library_tree(driver, tmp_path, emitters, max_leaf):
    src, dst = tmp_path / "in.bin", tmp_path / "out.bin"
    arr = (ob.TreeEmitter * len(emitters))(*emitters)
    with open(src, "wb") as f:
        f.write(np.array([len(emitters), max_leaf], np.uint32).tobytes())
        f.write(bytes(arr))
    subprocess.run([str(driver), str(src), str(dst)], check=True)
    words = np.fromfile(dst, np.uint32)
    ok, root, count = (int(v) for v in words[:3])
    at = 3
    nodes = words[at:at + 24 * count].reshape(count, 24)
    at += 24 * count
    n = len(emitters)
    records = words[at:at + 20 * n].reshape(n, 20)
    at += 20 * n
    mapping = words[at:at + 2 * n].reshape(n, 2)
    return ok, root, nodes, records, mapping


@pytest.mark.parametrize("n,leaf", [(300, 8), (5000, 8), (12000, 8), (9000, 1)])
def test_library_tree_matches_oracle(driver, tmp_path, n, leaf):
    rng = np.random.default_rng(n * 7 + leaf)
    pop = population(rng, n)
    ok, root, nodes, records, mapping = library_tree(driver, tmp_path, pop, leaf)
    ref = ob.light_tree(pop, leaf)
    assert ok == 1 and ref is not None
    assert root == ref["root"]
    rn = ref["nodes"].reshape(-1, 20)
    assert nodes.shape[0] == rn.shape[0]
    # topology: parent, left, right, kind; leaf ranges
    assert np.array_equal(nodes[:, 16:19], rn[:, 14:17])
    assert np.array_equal(nodes[:, 19], rn[:, 19])
    assert np.array_equal(nodes[:, 20:22], rn[:, 17:19])
    # measures: energy always, bounds and cone where the oracle has them
    m = rn[:, :12].view(np.float32)
    f = nodes[:, :16].view(np.float32)
    assert np.array_equal(f[:, 3].view(np.uint32), m[:, 11].view(np.uint32))
    boxed = rn[:, 12] == 0
    assert np.array_equal(f[boxed][:, 0:3].view(np.uint32), m[boxed][:, 0:3].view(np.uint32))
    assert np.array_equal(f[boxed][:, 4:7].view(np.uint32), m[boxed][:, 3:6].view(np.uint32))
    coned = rn[:, 13] == 0
    assert np.array_equal(f[coned][:, 8:11].view(np.uint32), m[coned][:, 6:9].view(np.uint32))
    assert np.array_equal(f[coned][:, 7].view(np.uint32), m[coned][:, 9].view(np.uint32))
    assert np.array_equal(f[coned][:, 11].view(np.uint32), m[coned][:, 10].view(np.uint32))
    # emitters in tree order and the reverse mappings
    re = ref["emitters"].reshape(-1, 19)
    assert np.array_equal(records[:, 16], re[:, 17])
    assert np.array_equal(mapping[:, 0], ref["to_tree"])
    assert np.array_equal(mapping[:, 1], ref["to_leaf"])
