"""vamtoolbox_amd: MI355X-native light-propagation path of VAMToolbox.

Drop-in for the reference's projector / optimiser path: ``geometry``,
``projectorconstructor``, ``optimize`` (CAL, OSMO, BCLP), ``response``,
``medium``, ``metrics``, ``util.data`` and ``imagesequence``, with the propagators running as
hand-written gfx950 HIP kernels (libvamx_hip.so, C ABI in include/vamx.h).
"""

from . import geometry, metrics, response  # noqa: F401
from . import util  # noqa: F401
from . import medium  # noqa: F401
from . import projectorconstructor  # noqa: F401
from . import optimize  # noqa: F401
from . import imagesequence  # noqa: F401

__version__ = "0.1.0"
