"""The C-ABI library loads and exports every function include/pmgpu.h declares."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "pmgpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pm_[a-z_]+)\s*\(", text)))


def test_header_matches_binding():
    from protomol_amd import gpu
    assert declared() == sorted(gpu.EXPORTS)


def test_library_exports_all_symbols():
    from protomol_amd import gpu
    if not os.path.exists(gpu.LIB_PATH):
        pytest.skip("libpmgpu.so not built")
    lib = ctypes.CDLL(gpu.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name
    assert lib.pm_abi_version() == 1


def declared_front():
    text = open(os.path.join(ROOT, "include", "pmfront.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(pmf_[a-z_]+)\s*\(", text)))


def test_front_library_exports_all_symbols():
    from protomol_amd import front
    if not os.path.exists(front.LIB_PATH):
        pytest.skip("libpmfront.so not built")
    lib = ctypes.CDLL(front.LIB_PATH)
    names = declared_front()
    assert len(names) >= 17
    for name in names:
        assert hasattr(lib, name), name
