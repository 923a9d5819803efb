"""CPU checks of the C-ABI library: it is built, loads, and exports every symbol include/lss_amd.h declares."""
import os
import re
import subprocess

import pytest

import linear_system_solvers_amd as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = set()
    for h in ("lss_amd.h", "lss_heat1d.h"):
        src = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"^\s*(?:int|int32_t|const char \*)\s*\**\s*(lss_\w+)\s*\(", src, re.M))
    return names


def test_header_and_binding_agree():
    from linear_system_solvers_amd import heat1d
    assert _declared() == set(L.EXPORTS) | set(heat1d.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not L.available():
        pytest.skip("liblss_amd.so not built")
    out = subprocess.run(["nm", "-D", "--defined-only", L.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r"\sT\s(lss_\w+)$", out, re.M))
    missing = _declared() - exported
    assert not missing, missing
    lib = L.lib()
    for name in _declared():
        assert getattr(lib, name) is not None


def test_invalid_arguments_are_rejected_without_a_gpu():
    if not L.available():
        pytest.skip("liblss_amd.so not built")
    import ctypes as C
    lib = L.lib()
    h = C.c_void_p()
    assert lib.lss_adi_plan_create(None, 1, C.byref(h)) == -1
    assert b"null" in lib.lss_last_error()
    assert lib.lss_tridiag_batched(7, 0, 1, 8, None, None, None, None, None, None) == -1
    assert lib.lss_model_plan_create(None, 0, 0, C.byref(h)) == -1
    bad = L.Model(nx=3)
    with pytest.raises(L.LssError):
        L.Plan.from_models([bad])
