"""Generated surface programs through the C ABI, without a device.

Materials whose flattened closure leaves have the same layout share one
straight-line HIP program; a scene's programs, a switch over them and a
wavefront shade kernel form one translation unit that the library compiles
with hipRTC (the reference's per-structure-signature SurfaceProgram JIT,
docs/architecture.md:63-110).  These helpers return that source and build it,
as psx_surface_program_source / psx_surface_program_build do.
"""
from __future__ import annotations

import ctypes as C

from psycles_amd import abi
from psycles_amd.scenes import material_structs


def program_source(materials, variant: int = 0) -> tuple[str, int]:
    """(generated source, number of distinct programs) for `materials`."""
    lib = abi.load_library()
    keep: list = []
    mats = material_structs(materials, keep)
    length, programs = C.c_size_t(), C.c_uint32()
    abi.check("psx_surface_program_source",
              lib.psx_surface_program_source(mats, len(materials), variant, None, 0, C.byref(length),
                                             C.byref(programs)))
    buf = C.create_string_buffer(length.value + 1)
    abi.check("psx_surface_program_source",
              lib.psx_surface_program_source(mats, len(materials), variant, buf, length.value + 1,
                                             C.byref(length), None))
    return buf.value.decode(), programs.value


def program_build(materials, variant: int = 0, arch: str = "gfx950") -> tuple[bool, int, str]:
    """hipRTC build of the generated source: (compiled, code-object bytes, log)."""
    lib = abi.load_library()
    keep: list = []
    mats = material_structs(materials, keep)
    log = C.create_string_buffer(1 << 16)
    code = C.c_uint64()
    rc = lib.psx_surface_program_build(mats, len(materials), variant, arch.encode(), log, len(log),
                                       C.byref(code))
    if rc not in (abi.PSX_OK, abi.PSX_ERROR_UNSUPPORTED):
        abi.check("psx_surface_program_build", rc)
    return rc == abi.PSX_OK, code.value, log.value.decode(errors="replace")
