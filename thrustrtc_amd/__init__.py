"""ThrustRTC primitive core on MI355X — the Python API of the reference's
`ThrustRTC` package (python/ThrustRTC/__init__.py), backed by libPyThrustRTC.so.

    import thrustrtc_amd as trtc
    d = trtc.device_vector_from_list([1, 4, 2, 8, 5, 7], 'int32_t')
    trtc.Sort(d)

Like the reference, importing the package initialises the GPU and raises
ImportError when no usable device (or no built library) is present.
"""
from .native import native, check_i

check_i(native.n_trtc_try_init())

from .Context import *  # noqa: E402,F401,F403
from .DeviceViewable import *  # noqa: E402,F401,F403
from .DVVector import (device_vector, device_vector_from_numpy, device_vector_from_list,  # noqa: E402,F401
                       device_vector_from_dvs, DVVectorLike, DVVector, DVRange, DVVectorAdaptor, DVTorchVector)
from .Functor import *  # noqa: E402,F401,F403
from .FakeVectors import *  # noqa: E402,F401,F403
from .Algorithms import *  # noqa: E402,F401,F403

__version__ = "0.1.0"
