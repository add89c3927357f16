"""MI355X-native PDX search path: HIP kernels behind a C-ABI (include/pdx_amd.h)."""
