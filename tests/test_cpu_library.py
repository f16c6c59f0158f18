"""Host-side checks that need no GPU: the C-ABI library loads and exports every
function include/cantera_gpu.h declares, and the ctypes table covers them."""
import os
import re

import pytest

from conftest import ROOT


def header_functions():
    text = open(os.path.join(ROOT, "include", "cantera_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(cg_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    import cantera_amd
    L = cantera_amd.lib()
    names = header_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(L, n), f"{n} declared in include/cantera_gpu.h but not exported"


def test_ctypes_table_matches_header():
    from cantera_amd._lib import SIGNATURES
    assert sorted(n for n, _, _ in SIGNATURES) == header_functions()


def test_library_reports_errors_without_gpu():
    """Calls on invalid handles return CG_ERR with a message (no exception
    crosses the C ABI)."""
    import cantera_amd
    from cantera_amd._lib import last_error
    L = cantera_amd.lib()
    assert L.cg_mech_nspecies(12345) == -999
    assert "invalid mechanism handle" in last_error()
    assert L.cg_mech_load(b"/nonexistent.mech") == -999
    assert "cannot open" in last_error()
