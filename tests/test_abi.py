"""The C-ABI library loads and exports every entry point include/jscc_hip.h declares
(no compute calls: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "jscc_hip.h")
LIB = os.path.join(ROOT, "jscc_recon_spect_cylinder_amd", "libjscc_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(jscc_\w+)\s*\(", text)))


@pytest.mark.skipif(not os.path.exists(LIB), reason="libjscc_hip.so not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(LIB)
    names = declared_functions()
    assert len(names) >= 20
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_python_signatures_cover_the_header():
    from jscc_recon_spect_cylinder_amd import _lib
    assert sorted(_lib.SIGNATURES) == declared_functions()


@pytest.mark.skipif(not os.path.exists(LIB), reason="libjscc_hip.so not built")
def test_abi_version_and_error_string():
    from jscc_recon_spect_cylinder_amd import _lib
    lib = _lib.load()
    assert lib.jscc_abi_version() == 1
    # a rejected call reports a message without touching the GPU
    rc = lib.jscc_em_update(None, None, 1.0, None, 0.0, None, 5, 9, 1e-12, None, None)
    assert rc != 0
    assert b"bad arguments" in lib.jscc_last_error()


def test_product_package_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "jscc_recon_spect_cylinder_amd")
    for fn in os.listdir(pkg):
        if fn.endswith(".py"):
            src = open(os.path.join(pkg, fn)).read()
            assert "import oracle" not in src and "from oracle" not in src, fn


def test_source_cache_hits_only_the_same_tensor_at_the_same_version():
    """Derived-data caches (rotation tables, voxel arrays, aligned matrices) must not
    return a result for a different tensor that happens to reuse the same memory."""
    import gc
    import weakref
    import numpy as np
    import torch
    from jscc_recon_spect_cylinder_amd.ops import SourceCache
    c = SourceCache(4)
    a = torch.from_numpy(np.arange(12, dtype=np.int64).reshape(6, 2))
    c.put(a, ("d",), "A")
    assert c.get(a, ("d",)) == "A"
    assert c.get(a, ("other",)) is None
    assert c.get(a.view(6, 2), ("d",)) is None           # same memory, another tensor
    assert c.get(torch.from_numpy(a.numpy()), ("d",)) is None
    a.add_(1)
    assert c.get(a, ("d",)) is None                       # modified in place
    ref = weakref.ref(a)
    c.put(a, ("d",), "A2")
    del a
    gc.collect()
    assert ref() is not None                              # the entry keeps its source alive
    for k in range(4):
        c.put(torch.zeros(1), (k,), k)                    # capacity reached: entries dropped
    gc.collect()
    assert ref() is None


@pytest.mark.skipif(not os.path.exists(LIB), reason="libjscc_hip.so not built")
def test_projection_column_limit_covers_registers_and_lds():
    """jscc_proj_max_cols is the smaller of the register and the LDS limit (32768 columns for every storage
    type); a wider matrix is rejected by the host checks before any launch."""
    from jscc_recon_spect_cylinder_amd import _lib
    lib = _lib.load()
    for dt in (0, 1, 2, 3):
        assert lib.jscc_proj_max_cols(dt) == 32768
    err = ctypes.c_int32(0)
    rc = lib.jscc_proj(0, None, 1, 40960, 40960, None, 8, 0, None, None, None, 0, 40960, None, 0, None, 1e-12, 1,
                       None, None, 0, 1.0, 0, 0, None, 40960, None, 0, ctypes.addressof(err), None)
    assert rc != 0 and b"too many columns" in lib.jscc_last_error()
