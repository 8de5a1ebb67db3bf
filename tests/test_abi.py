"""CPU checks of the drop-in boundary: both shared libraries load without a
GPU and export every entry point their headers declare; the product refuses
to run without its native libraries (no fallback).  No compute calls."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header, pattern):
    text = open(os.path.join(ROOT, "include", header)).read()
    return sorted(set(re.findall(pattern, text, re.M)))


@pytest.fixture(scope="module")
def built():
    from slepc_amd import build
    if not (os.path.exists(build.KS_LIB) and os.path.exists(build.HOST_LIB)):
        build.build_all(oracle=False)
    return build


def test_ks_hip_exports_every_declared_symbol(built):
    names = _declared("ks_hip.h", r"^\s*(?:int|void \*|const char \*)\s*(ks_\w+)\s*\(")
    assert len(names) >= 30
    lib = ctypes.CDLL(built.KS_LIB, mode=ctypes.RTLD_GLOBAL)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_slepc_amd_exports_every_declared_symbol(built):
    names = _declared("slepc_amd.h", r"^(?:PetscErrorCode|const char\s*\*|void \*)\s*(\w+)\s*\(")
    assert "EPSSolve" in names and "BVCreate_HIPKS" in names and "MatCreate_HIPCSR" in names
    ctypes.CDLL(built.KS_LIB, mode=ctypes.RTLD_GLOBAL)
    lib = ctypes.CDLL(built.HOST_LIB)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_no_fallback_when_library_missing(tmp_path, monkeypatch):
    import importlib
    from slepc_amd import ks
    monkeypatch.setattr(ks, "LIB_PATH", str(tmp_path / "absent.so"))
    monkeypatch.setattr(ks, "_lib", None)
    with pytest.raises(ks.KSError):
        ks.lib()
    importlib.reload(ks)


def test_context_refuses_without_device(built):
    # on a machine without a gfx950 device the runtime reports it instead of computing
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    from slepc_amd import ks
    lib = ks.lib()
    assert lib.ks_init(0) == -4   # KS_ERR_NODEV
