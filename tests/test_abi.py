"""The C-ABI library loads and exports every function include/hci_gpu.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

from risq_amd import hci

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "hci_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hci_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_binding_table():
    names = declared_functions()
    assert set(names) == set(hci.SIGNATURES), set(names) ^ set(hci.SIGNATURES)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(hci.LIB_PATH):
        import __graft_entry__
        __graft_entry__.build()
    lib = ctypes.CDLL(hci.LIB_PATH)
    for name in declared_functions():
        assert hasattr(lib, name), name
    hci.load_library()


def test_no_gpu_context_fails_cleanly():
    """Without a device hci_create returns an error code instead of aborting."""
    lib = hci.load_library()
    if os.environ.get("HIP_VISIBLE_DEVICES", None) is None and os.path.exists("/dev/kfd"):
        return  # a GPU may be present; creation could succeed
    h = ctypes.c_void_p()
    rc = lib.hci_create(0, ctypes.byref(h))
    assert rc != 0 and not h.value
