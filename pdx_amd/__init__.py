"""MI355X-native PDX search path: HIP kernels behind a C-ABI (include/pdx_amd.h), the reference's
index API in C++ (pdx_amd/csrc/host) and Python (pdx_amd.index_factory)."""
__all__ = ["IndexPDXIVF", "IndexPDXIVFSQ8", "IndexPDXIVFTree", "IndexPDXIVFTreeSQ8", "load_index", "Server"]


def __getattr__(name):  # lazy: importing the package does not load the native module
    if name in __all__:
        from pdx_amd import index_factory

        return getattr(index_factory, name)
    raise AttributeError(name)
