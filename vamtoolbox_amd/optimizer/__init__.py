"""CAL, OSMO and BCLP sinogram optimisers running on the GPU."""
from . import BCLP, CAL, OSMO  # noqa: F401
