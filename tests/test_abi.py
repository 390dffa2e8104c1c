"""The C-ABI library loads without a GPU and exports every entry point the header declares."""
import ctypes
import os
import re

from spiffe_amd import abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "spiffe_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(spf_[a-z_0-9]+)\s*\(", text)))


def test_header_declares_the_bound_entry_points():
    fns = header_functions()
    assert len(fns) >= 25
    assert sorted(abi.SIGNATURES) == fns


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(abi.LIB_PATH)
    missing = [f for f in header_functions() if not hasattr(lib, f)]
    assert not missing, missing


def test_struct_layouts_match_the_header():
    # sizes follow from the C declarations (natural alignment on x86-64 / amdgcn)
    assert ctypes.sizeof(abi.ScreenRow) == 12 * 8 + 8 + 8 + 4 + 2 + 2
    assert ctypes.sizeof(abi.LostRow) == 8 + 7 * 4 + 4 + 8 + 8 + 2 + 2 + 4
    assert ctypes.sizeof(abi.PoissonParams) == 6 * 8


def test_context_creation_fails_cleanly_without_a_gpu():
    lib = abi.load()
    h = ctypes.c_void_p()
    rc = lib.spf_ctx_create(0, ctypes.byref(h))
    if rc != 0:
        assert rc in (abi.SPF_EHIP, abi.SPF_EINVAL)
        assert not h.value
    else:
        lib.spf_ctx_destroy(h)
