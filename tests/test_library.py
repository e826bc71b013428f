"""CPU checks of the C ABI library: it is built, it loads, and it exports every function that
include/bte_hip.h declares (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

from helpers import ROOT

HEADER = os.path.join(ROOT, "include", "bte_hip.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bte_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    assert "bte_g1_msm" in names and "bte_g2_msm_indexed" in names
    assert len(names) >= 10


def test_library_exports_every_declared_symbol():
    import weighted_bte_amd

    lib = weighted_bte_amd.lib()
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert b"gfx950" in lib.bte_version()


def test_python_binding_covers_header():
    import weighted_bte_amd

    missing = [n for n in declared_functions() if n not in weighted_bte_amd._SIGNATURES]
    assert not missing, missing


def test_error_path_without_gpu_work():
    """Argument validation returns a status instead of aborting (reference panics)."""
    import weighted_bte_amd

    lib = weighted_bte_amd.lib()
    rc = lib.bte_msm_set_window(7, 16, 3)
    assert rc == 1
    assert b"bad arguments" in lib.bte_last_error()
