"""ReconNet projector / reconstruction path for AMD Instinct MI355X (gfx950).

The compute path is libReconNet4D_no_tf.so (hand-written HIP kernels behind a
C-ABI that matches the reference library's exports).  ``recon.ReconNet`` is
the host-array front end with the reference Python class's interface;
``torch_ops`` and ``sqs`` drive the same kernels on device tensors.
"""

from ._lib import LIB_PATH, lib, last_error  # noqa: F401
from .geometry import FanGeometry, WORKLOADS, fan_angles, interleaved_order, HU_WATER  # noqa: F401

__all__ = ["lib", "last_error", "LIB_PATH", "FanGeometry", "WORKLOADS", "fan_angles", "interleaved_order",
           "HU_WATER"]
