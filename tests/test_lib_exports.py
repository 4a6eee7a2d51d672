"""The C-ABI library loads and exports every entry point the header declares
(no device calls: these run without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cardiotrust_hip.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ct_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def libpath():
    from cardiotrust_amd import build, gpu
    if not os.path.exists(gpu.LIB_PATH):
        build.build()
    return gpu.LIB_PATH


def test_header_declares_python_exports():
    from cardiotrust_amd import gpu
    assert declared() == sorted(gpu.EXPORTS)


def test_library_exports_every_declared_symbol(libpath):
    out = subprocess.run(["nm", "-D", "--defined-only", libpath], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r" T (ct_[a-z0-9_]+)$", out, flags=re.M))
    missing = [s for s in declared() if s not in exported]
    assert not missing, missing


def test_library_loads_and_binds(libpath):
    from cardiotrust_amd import gpu
    L = gpu.lib()
    for name in gpu.EXPORTS:
        assert getattr(L, name) is not None
    assert isinstance(L.ct_last_error(), bytes)


def test_library_has_no_oracle_dependency(libpath):
    out = subprocess.run(["nm", "-D", libpath], capture_output=True, text=True, check=True).stdout
    assert "ct_oracle" not in out
    deps = subprocess.run(["ldd", libpath], capture_output=True, text=True).stdout
    assert "oracle" not in deps


def test_device_code_is_gfx950(libpath):
    # the offload bundle names its target
    data = open(libpath, "rb").read()
    assert b"gfx950" in data
