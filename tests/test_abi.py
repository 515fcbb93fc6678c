"""The native libraries load without a GPU and export every symbol the
headers in include/ declare (no compute calls)."""
import ctypes as C
import re
from pathlib import Path

import pytest

import lpmcxx_amd as L

ROOT = Path(__file__).resolve().parent.parent


def _declared(header: str) -> list[str]:
    text = (ROOT / "include" / header).read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lpm_[a-z0-9_]+)\s*\(", text)))


@pytest.mark.parametrize("header,lib", [("lpm_gpu.h", "liblpm_gpu.so"), ("lpm_mesh.h", "liblpm.so")])
def test_exports(header, lib):
    names = _declared(header)
    assert names, header
    so = L.gpu_lib() if lib == "liblpm_gpu.so" else L.host_lib()
    missing = [n for n in names if not hasattr(so, n)]
    assert not missing, missing


def test_status_strings():
    lib = L.gpu_lib()
    assert lib.lpm_status_string(0) == b"ok"
    assert lib.lpm_status_string(-7) == b"invalid kernel"
    assert lib.lpm_status_string(-4) == b"device allocation failed"


def test_mesh_capi_errors():
    lib = L.host_lib()
    h = C.c_void_p()
    assert lib.lpm_mesh_create(C.byref(h), 7, 2, 1.0) == -1      # unknown seed
    assert lib.lpm_mesh_create(C.byref(h), 1, -1, 1.0) == -1     # bad level
    assert lib.lpm_mesh_create(C.byref(h), 1, 2, -1.0) == -1     # bad radius
    assert lib.lpm_mesh_create(None, 1, 2, 1.0) == -2
