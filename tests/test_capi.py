"""The C-ABI library loads and exports every entry point include/dpmi.h
declares (no GPU calls)."""

import ctypes
import re
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "dpmi.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dpmi_\w+)\s*\(", text)))


def test_library_exports_header_symbols():
    from dpmi import _lib

    if not _lib.LIB_PATH.exists():
        pytest.skip("libdpmi.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = _declared()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, f"declared but not exported: {missing}"
    # every declared entry point is bound with a signature on the Python side
    unbound = [n for n in names if n not in _lib.exported_symbols()]
    assert not unbound, f"declared but not bound in dpmi._lib: {unbound}"


def test_abi_version_and_error_text():
    from dpmi import _lib

    if not _lib.LIB_PATH.exists():
        pytest.skip("libdpmi.so not built")
    lib = _lib.load()
    assert lib.dpmi_abi_version() == 1
    # argument errors are reported without touching the device
    out = ctypes.c_int64(0)
    rc = lib.dpmi_nlist_workspace_bytes(ctypes.byref(out), None, 10, ctypes.c_float(-1.0))
    assert rc == 1
    assert b"rcut" in lib.dpmi_last_error()


def test_workspace_size_host_only():
    import numpy as np

    from dpmi import _lib

    if not _lib.LIB_PATH.exists():
        pytest.skip("libdpmi.so not built")
    lib = _lib.load()
    box = np.diag([99.36] * 3).ravel()
    out = ctypes.c_int64(0)
    assert lib.dpmi_nlist_workspace_bytes(ctypes.byref(out), box.ctypes.data, 110592, ctypes.c_float(6.0)) == 0
    # per-atom arrays and (16+2)^3 extended cells fit in a few MB
    assert 110592 * 12 < out.value < 8 << 20


def test_ops_refuse_cpu_tensors():
    """No CPU fallback: the product path raises on host tensors."""
    import torch

    from dpmi import _lib, ops

    if not _lib.LIB_PATH.exists():
        pytest.skip("libdpmi.so not built")
    with pytest.raises(RuntimeError, match="GPU path only"):
        ops.tabulate_forward(torch.zeros((2, 48)), torch.tensor([0, 1, 2, 0.1, 0.1, -1.0]), torch.zeros((4, 1)),
                             torch.zeros((1, 4, 4)), 8)
