"""The C-ABI library loads without a GPU and exports every symbol include/pdx_amd.h declares."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    hdr = open(os.path.join(ROOT, "include", "pdx_amd.h")).read()
    return sorted(set(re.findall(r"(?:pdxg_status|size_t|int|uint32_t|const char\s*\*)\s*(pdxg_\w+)\s*\(", hdr)))


def test_header_declarations_are_listed_in_binding():
    from pdx_amd import _capi

    assert sorted(_capi.EXPORTS) == _declared()


def test_library_loads_and_exports_every_declared_symbol():
    lib = ctypes.CDLL(os.path.join(ROOT, "pdx_amd", "libpdx_amd.so"))
    for name in _declared():
        assert hasattr(lib, name), name


def test_errors_without_gpu_are_reported_not_raised():
    from pdx_amd import _capi

    L = _capi.lib()
    h = ctypes.c_void_p()
    st = L.pdxg_index_create(0, None, ctypes.byref(h))
    assert st == 1  # PDXG_ERR_INVALID for a null description
    assert b"null" in L.pdxg_last_error()


def test_python_api_module_imports_and_validates_config():
    import pytest

    import pdx_amd

    with pytest.raises(ValueError):
        pdx_amd.IndexPDXIVF(num_dimensions=0)
    with pytest.raises(ValueError):
        pdx_amd.IndexPDXIVFTree(num_dimensions=8, num_clusters=4, num_meso_clusters=4)
    with pytest.raises(ValueError):
        pdx_amd.IndexPDXIVF(num_dimensions=8, kmeans_iters=100)
    with pytest.raises(RuntimeError):
        pdx_amd.load_index("/nonexistent/file.pdx")
