"""The C ABI library loads without a GPU and exports every entry point that
include/htr_amd.h declares (no compute calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "htr_amd.h")
LIB = os.path.join(ROOT, "htr-solver_amd", "libhtr_amd.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(htr_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared()
    for must in ("htr_euler_flux", "htr_diffusion_flux", "htr_chem_update", "htr_full_rhs",
                 "htr_max_spectral_radius", "htr_update_primitive", "htr_shock_sensor"):
        assert must in names


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_python_binding_covers_the_header():
    import sys
    sys.path.insert(0, os.path.join(ROOT, "htr-solver_amd"))
    import htr_amd
    assert sorted(htr_amd.EXPORTS) == declared()
    assert htr_amd.lib().htr_version().startswith(b"htr-amd")
