"""The C-ABI library loads, exports every symbol include/psx.h declares and
its host-only entry points agree with the oracle (no GPU needed)."""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

import oracle_binding as ob
from psycles_amd import abi

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    text = (ROOT / "include" / "psx.h").read_text()
    return sorted(set(re.findall(r"\b(psx_[a-z_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not abi.LIBRARY_PATH.exists():
        pytest.skip("libpsx.so not built")
    return abi.load_library()


def test_exports_every_declared_symbol(lib):
    names = declared_symbols()
    assert len(names) >= 18
    assert sorted(abi.EXPORTED_SYMBOLS) == names
    for name in names:
        assert hasattr(lib, name), name


def test_version(lib):
    assert b"psx" in lib.psx_version()


def test_settings_defaults_match_contract(lib):
    s = abi.RenderSettings()
    lib.psx_render_settings_default(C.byref(s))
    d = abi.default_settings(0, 0)
    for field, _ in abi.IntegratorSettings._fields_:
        assert getattr(s.integrator, field) == pytest.approx(getattr(d.integrator, field)), field
    assert s.filter_width == 1.0 and s.pass_alpha_threshold == 0.5


def test_host_sobol_table_matches_oracle(lib):
    for total in (1, 16, 300, 4096):
        count = C.c_uint32()
        assert lib.psx_sobol_table(total, None, C.byref(count)) == 0
        out = np.zeros(count.value * 4, dtype=np.float32)
        assert lib.psx_sobol_table(total, out.ctypes.data_as(C.POINTER(C.c_float)), None) == 0
        size = ob.load().oracle_sequence_size_for_samples(total)
        assert count.value == 256 * size
        assert np.array_equal(out.view(np.uint32), ob.sobol_table(size).reshape(-1).view(np.uint32))


def test_invalid_arguments_fail_cleanly(lib):
    out = C.c_void_p()
    assert lib.psx_scene_compile(None, None, abi.DIAG_FN(0), None, C.byref(out)) == \
        abi.PSX_ERROR_INVALID_ARGUMENT
    assert lib.psx_sobol_table(0, None, None) == abi.PSX_ERROR_INVALID_ARGUMENT
    assert lib.psx_session_render(None, 0, 1, 0, 1, None) == abi.PSX_ERROR_INVALID_ARGUMENT
