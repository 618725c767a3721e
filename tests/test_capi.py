"""The C-ABI library loads and exports every entry point declared in
include/pearl_gpu.h (no GPU calls)."""
import os
import re
import ctypes

from conftest import ROOT, load_pkg


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "pearl_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pearl_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_expected_api():
    syms = declared_symbols()
    for s in ["pearl_create", "pearl_advance", "pearl_maybe_shift", "pearl_set_state", "pearl_get_state",
              "pearl_last_error", "pearl_destroy"]:
        assert s in syms


def test_library_exports_every_declared_symbol():
    pkg = load_pkg()
    lib = ctypes.CDLL(pkg.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert sorted(pkg.EXPORTS) == declared_symbols()


def test_error_messages_match_reference_text():
    pkg = load_pkg()
    assert pkg.WHY == ["phase increment", "linear/closure", "composition/phase bounds", "mass conservation",
                       "free-energy increase"]
