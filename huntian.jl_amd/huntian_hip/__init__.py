"""huntian_hip -- MI355X (gfx950) implementation of Huntian.jl's halo-finding hot path.

Host-side mirror of the reference plugin API for this path (the CUDA.jl extension methods
looked up by Huntian's drivers), calling hand-written HIP kernels in libhuntian_hip.so
through its C ABI (include/huntian_abi.h).
"""
from .abi import HuntianError, NativeLibraryMissing  # noqa: F401
from .homeltab import read_homeltab, write_homeltab, make_ctab, WEBSKY_TABLE  # noqa: F401
