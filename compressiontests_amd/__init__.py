"""MI355X (gfx950) MG-PCG / SpMV solver backend for the FE stiffness systems of
nickreserved/CompressionTests.  The compute lives in libmgx.so (HIP kernels behind the C ABI
of include/mgx.h); this package is its Python binding and the mirror of the reference's
solver and model interfaces."""
from . import _lib
from ._lib import (MgxError, MgxInvalidArgument, MgxInvalidOperation, MgxNotImplemented, MgxNotSPD,
                   device_count)

__all__ = ["_lib", "MgxError", "MgxInvalidArgument", "MgxInvalidOperation", "MgxNotImplemented", "MgxNotSPD",
           "device_count"]
