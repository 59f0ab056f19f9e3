"""verona_amd: MI355X-native hot path of the Verona SRHD solver.

Compute runs in libverona_hip.so (hand-written gfx950 HIP kernels behind the
C ABI of include/verona_hip.h); this package is the host-side mirror of
Verona's Julia API (Verona.EosTypes, Verona2D, Verona3D).
"""
from . import eos as EosTypes
from . import verona1d as Verona1D
from . import verona2d as Verona2D
from . import verona3d as Verona3D
from .bc import Boundary, Cart
from .c2p import Con2PrimBatch
from .device import DeviceGrid, host_array

__all__ = ["EosTypes", "Verona1D", "Verona2D", "Verona3D", "Boundary", "Cart", "Con2PrimBatch", "DeviceGrid",
           "host_array"]
