"""The C ABI: libtd3_hip.so loads and exports every entry point include/td3.h declares, and
the Python binding declares the same set.  No kernel is launched (runs without a GPU)."""

import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    text = open(os.path.join(ROOT, "include", "td3.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(td3_\w+)\s*\(", text)))


def test_library_exports_every_declared_symbol(td3):
    lib_path = os.path.join(ROOT, "tree-d-3d_amd", "libtd3_hip.so")
    assert os.path.exists(lib_path), "build libtd3_hip.so first (tree-d-3d_amd/build.py)"
    lib = ctypes.CDLL(lib_path)
    names = declared()
    assert len(names) >= 10
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    lib.td3_abi_version.restype = ctypes.c_int
    assert lib.td3_abi_version() == 1


def test_binding_matches_header(td3):
    from tree_d_3d_amd import _abi

    assert sorted(_abi.SIGNATURES) == declared()


def test_no_cpu_fallback_without_gpu(td3):
    import torch

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(RuntimeError, match="GPU"):
        td3.direct_field([[0.0], [0.0]], [[1.0], [1.0]], [1.0])


def test_moments_made_on_first_use_and_again_after_a_failure(td3):
    """A tree's array dict makes the deferred cell moments on the first read of one of
    them, once; a maker that fails stays deferred for the next read; other names stay
    missing."""
    import sys

    arrays = sys.modules["tree_d_3d_amd.flat_tree"]._Arrays(pos=1)
    calls = []

    def make():
        calls.append(1)
        if len(calls) == 1:
            raise RuntimeError("first attempt")
        arrays.update(moment_re=5, moment_im=6, node_qabs=7)

    arrays.defer(make)
    assert "moment_re" not in arrays
    with pytest.raises(RuntimeError):
        arrays["moment_re"]
    assert arrays["moment_im"] == 6 and arrays["moment_re"] == 5 and len(calls) == 2
    with pytest.raises(KeyError):
        arrays["node_center"]
