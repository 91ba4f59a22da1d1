"""The C-ABI library loads and exports every symbol include/kiyosi_amd.h
declares (no compute calls: runs without a GPU)."""

import ctypes
import re

from conftest import ROOT


def declared_symbols():
    header = (ROOT / "include" / "kiyosi_amd.h").read_text()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    return sorted(set(re.findall(r"\b(kg_[a-z0-9_]+)\s*\(", header)))


def test_header_declares_the_engine_entry_points():
    names = declared_symbols()
    for required in ("kg_context_create", "kg_vanilla_price", "kg_autocallable_price",
                     "kg_pcg32_outputs", "kg_sobol_points"):
        assert required in names


def test_library_exports_every_declared_symbol():
    from kiyosi_amd import _native
    lib = ctypes.CDLL(str(_native.LIB_PATH))
    missing = [name for name in declared_symbols() if not hasattr(lib, name)]
    assert not missing, missing


def test_python_binding_covers_every_declared_symbol():
    from kiyosi_amd import _native
    assert set(declared_symbols()) <= set(_native.SIGNATURES)
    _native.lib()  # declares every signature without calling into the GPU


def test_abi_struct_sizes_match_the_header():
    from kiyosi_amd import _abi
    assert ctypes.sizeof(_abi.Settings) == 40
    assert ctypes.sizeof(_abi.Moments) == 32
    assert ctypes.sizeof(_abi.Estimate) == 80
    assert ctypes.sizeof(_abi.BsmMarket) == 40
    assert ctypes.sizeof(_abi.EuropeanTerms) == 32
    assert ctypes.sizeof(_abi.AutocallEvent) == 40


def test_header_limits_match_the_python_binding():
    from kiyosi_amd import _abi
    header = (ROOT / "include" / "kiyosi_amd.h").read_text()
    limits = dict(re.findall(r"#define KG_(MAX_[A-Z_]+) (\d+)", header))
    assert limits, "no KG_MAX_* limits found"
    for name, value in limits.items():
        assert getattr(_abi, name) == int(value), name
