"""pyseistr_amd: MI355X (gfx950) drop-ins for pyseistr's structural hot path.

Drop-in names (same signatures and output shapes as pyseistr):
dip2dc, dip3dc (and dip2d, dip3d), pwspray2d, pwspray3d, somean2dc,
somean3dc, somf2dc, somf3dc, soint2dc, soint3dc (and soint2d, soint3d),
sint2dc, sint3dc, smoothc, bandpassc, patch2d, patch3d, patch2d_inv,
patch3d_inv, fkdip.  All computation runs in libseistr_hip.so; there is no CPU
fallback.
"""
from ._lib import SeistrError, check_finite  # noqa: F401
from .bandpass import bandpassc  # noqa: F401
from .dip import dip2d, dip2dc, dip3d, dip3dc  # noqa: F401
from .fk import fkdip  # noqa: F401
from .filters import pwspray2d, pwspray3d, somean2dc, somean3dc, somf2dc, somf3dc  # noqa: F401
from .interp import sint2dc, sint3dc, soint2d, soint2dc, soint3d, soint3dc  # noqa: F401
from .patch import patch2d, patch2d_inv, patch3d, patch3d_inv  # noqa: F401
from .smooth import smoothc  # noqa: F401

__version__ = "0.1.0"
