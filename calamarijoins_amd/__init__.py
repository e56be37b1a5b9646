"""MI355X-native join engine for the contest API (Plan / ColumnarTable /
8 KiB pages).  The engine is libcalamarijoins_amd.so (HIP kernels for gfx950
plus the C++ executor); this package holds its ctypes bindings."""
from .abi import load_library  # noqa: F401
