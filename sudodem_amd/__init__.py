"""sudodem_amd -- SudoDEM3D's superellipsoid step loop on AMD MI355X.

The step (bounds, sort-and-sweep broadphase, Levenberg-Marquardt / Nelder-Mead narrowphase,
Superquadrics laws, Newton integration) runs in libsdm_hip.so, hand-written HIP for gfx950,
behind the C ABI of include/sdm.h.  This package is the Python face of that ABI with the
reference's class names (see scene.py).
"""
from ._native import (SDM_FP64, SDM_LAW_HERTZ, SDM_LAW_LINEAR, SDM_MIXED, Context, SdmError, SdmLibrary,  # noqa: F401
                      load_hip)
from .scene import *  # noqa: F401,F403
