"""The C-ABI library loads and exports every entry point include/trt_capi.h declares (CPU only)."""
import ctypes
import os
import re

from conftest import ROOT


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "trt_capi.h")).read()
    return sorted(set(re.findall(r"\b(trt_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol(trt):
    lib = trt.native.load()
    names = declared_symbols()
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(trt.native.EXPORTS) == set(names)


def test_create_without_gpu_reports_error(trt):
    """Without a GPU the context cannot be created; the error comes back through trt_last_error."""
    import pytest
    try:
        import torch  # noqa: F401
    except Exception:
        pass
    prob = trt.problem.small()
    if os.path.exists("/dev/kfd"):
        pytest.skip("a GPU is visible")
    with pytest.raises(trt.native.TrtError):
        trt.native.Context(prob)
