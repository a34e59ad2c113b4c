"""The C-ABI library loads on a machine without a GPU and exports every
entry point include/tci.h declares; device calls fail loudly there."""
import ctypes as C
import os
import re

import pytest

from conftest import load_pkg

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    txt = open(os.path.join(ROOT, "include", "tci.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:tci_status|const char \*|void)\s*\*?\s*(tci_\w+)\s*\(", txt, re.M)))


def test_header_declares_entry_points():
    names = _declared()
    assert "tci_rrlu_f64" in names and "tci_solve_submit" in names and "tci_pi_eval_f64" in names
    assert len(names) >= 20


def test_library_exports_every_declared_symbol():
    pkg = load_pkg()
    lib = C.CDLL(pkg._abi.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(pkg._abi.EXPORTED_SYMBOLS) == _declared()


def test_no_device_is_a_loud_error():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    pkg = load_pkg()
    with pytest.raises(pkg.TCIError) as e:
        pkg.Context(0)
    assert e.value.status == 12        # TCI_ERR_DEVICE: no CPU fallback


def test_ctx_stream_rejects_null_arguments():
    pkg = load_pkg()
    lib = pkg._abi.lib()
    sp = C.c_void_p()
    assert lib.tci_ctx_stream(None, C.byref(sp)) == 3     # TCI_ERR_ARG
    assert sp.value is None
