"""CPU checks of the C-ABI library: it loads, exports every function include/mdx.h declares,
and its host-side constants agree with the Python unit module. No device work is done."""
import os
import re

import pytest

from conftest import ROOT


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mdx.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mdx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from mdpy_amd import _lib as L
    lib = L.load()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding
    assert set(names) <= set(L.SIGNATURES), set(names) - set(L.SIGNATURES)


def test_host_constants_agree():
    from mdpy_amd import _lib as L
    from mdpy_amd.unit import ONE_OVER_4PI_EPS0
    lib = L.load()
    assert lib.mdx_coulomb_constant() == pytest.approx(ONE_OVER_4PI_EPS0, rel=1e-15)
    assert lib.mdx_version() >= 100


def test_context_creation_fails_loudly_without_device():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    from mdpy_amd import _lib as L
    from mdpy_amd.error import DeviceError
    import ctypes
    p = ctypes.c_void_p()
    with pytest.raises(DeviceError):
        L.call("mdx_ctx_create", 0, None, ctypes.byref(p))
