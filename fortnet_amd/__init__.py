"""MI355X (gfx950) implementation of Fortnet's BPNN hot path.

ACSF descriptors over periodic neighbour lists, per-element networks, analytic forces and
weight gradients run in libfnet_hip.so (hand-written HIP kernels); this package is the host
side: model/batch management mirroring the reference's TAcsf/TBpnn/TNetwork data model.
"""
from ._lib import (FNET_BATCHED, FNET_BF16, FNET_BF16_FAST, FNET_FP32,  # noqa: F401
                   FNET_FP32_NET64, FNET_FP64, FNET_PER_STRUCT, FnetError)
from .core import (AA__BOHR, ACTIVATIONS, BOHR__AA, Batch, Context, GFunctions,  # noqa: F401
                   Geometry, Netstat, nparams)
