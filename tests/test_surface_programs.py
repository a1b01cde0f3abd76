"""Generated surface programs without a device: the source a scene's
materials generate and its hipRTC build for gfx950.

Materials whose flattened closure leaves have the same layout share one
program, whatever their parameter values; every program appears once in the
switch; the source compiles with the flags of the compiled-in kernels.
"""
import ctypes as C

import numpy as np
import pytest

from psycles_amd import abi, programs, scenes


def test_cornell_programs_share_layouts():
    snap = scenes.cornell_box()
    src, n = programs.program_source(snap.materials, 0)
    # five diffuse materials of different colours share one program; the
    # emissive ceiling has its own
    assert len(snap.materials) > 2 and n == 2
    assert src.count("__device__ __forceinline__ void surface_program_") == n
    for i in range(n):
        assert f"case {i}u: surface_program_{i}(p, cs); return;" in src
    assert "psx::shade_launch<false, false>" in src
    assert "leaf 0: emission" in src and "leaf 0: diffuse" in src


def test_layout_not_values_decides_the_program():
    a = scenes.diffuse_material((0.1, 0.2, 0.3))
    b = scenes.diffuse_material((0.9, 0.8, 0.7), roughness=0.5)
    _, n = programs.program_source([a, b], 0)
    assert n == 1
    # two mix(diffuse, transparent) materials and a diffuse ground: two
    # programs, and one more plain diffuse material adds none
    c = scenes.instanced_icospheres(frequency=4, instances=2).materials
    assert programs.program_source(c, 0)[1] == 2
    assert programs.program_source(c + [a, b], 0)[1] == 2


def test_variants_name_their_shade_kernel():
    mats = scenes.cornell_lamps().materials
    for v, lp, lights in ((0, "false", "false"), (1, "true", "false"), (2, "false", "true"), (3, "true", "true")):
        src, _ = programs.program_source(mats, v)
        assert f"psx::shade_launch<{lp}, {lights}>" in src


def test_mix_leaves_are_unrolled():
    snap = scenes.instanced_icospheres(frequency=4, instances=2)
    src, n = programs.program_source(snap.materials, 0)
    assert "leaf 1: transparent" in src
    assert "w = w * (1.0f - f);" in src and "w = w * f;" in src


def test_invalid_requests():
    lib = abi.load_library()
    keep = []
    mats = scenes.material_structs(scenes.cornell_box().materials, keep)
    length = C.c_size_t()
    assert lib.psx_surface_program_source(mats, 2, 4, None, 0, C.byref(length), None) == abi.PSX_ERROR_INVALID_ARGUMENT
    assert lib.psx_surface_program_source(mats, 2, 0, None, 0, None, None) == abi.PSX_ERROR_INVALID_ARGUMENT
    small = C.create_string_buffer(16)
    assert lib.psx_surface_program_source(mats, 2, 0, small, 16, C.byref(length), None) == \
        abi.PSX_ERROR_INVALID_ARGUMENT
    assert length.value > 16
    bad = scenes.material_structs(scenes.cornell_box().materials[:1], keep)
    bad[0].instructions = None
    assert lib.psx_surface_program_source(bad, 1, 0, None, 0, C.byref(length), None) == \
        abi.PSX_ERROR_INVALID_ARGUMENT


@pytest.mark.parametrize("scene,variant", [("cornell", abi.PROGRAM_LIGHTS),
                                           ("instanced", 0)])
def test_generated_source_builds_for_gfx950(scene, variant):
    snap = scenes.cornell_box() if scene == "cornell" else scenes.instanced_icospheres(frequency=4, instances=2)
    ok, code, log = programs.program_build(snap.materials, variant)
    assert ok, log
    assert code > 10000
    assert "error" not in log.lower()


def test_build_reports_a_bad_architecture():
    ok, code, log = programs.program_build(scenes.cornell_box().materials, 0, arch="gfx000")
    assert not ok and code == 0 and log
