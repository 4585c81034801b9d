"""pyseistr_amd: MI355X (gfx950) drop-ins for pyseistr's structural hot path.

Drop-in names (same signatures and output shapes as pyseistr):
dip2dc, dip3dc (and dip2d, dip3d), pwspray2d, pwspray3d, somean2dc,
somean3dc, somf2dc, somf3dc, soint2dc, soint3dc (and soint2d, soint3d),
sint2dc, sint3dc, smoothc, bandpassc.  All computation runs in
libseistr_hip.so; there is no CPU fallback.
"""
from ._lib import SeistrError
from .dip import dip2d, dip2dc, dip3d, dip3dc

__version__ = "0.1.0"

try:  # modules added as the library grows
    from .filters import pwspray2d, pwspray3d, somean2dc, somean3dc, somf2dc, somf3dc  # noqa: F401
except ImportError:  # pragma: no cover
    pass
try:
    from .interp import sint2dc, sint3dc, soint2d, soint2dc, soint3d, soint3dc  # noqa: F401
except ImportError:  # pragma: no cover
    pass
from .bandpass import bandpassc  # noqa: E402,F401
from .smooth import smoothc  # noqa: E402,F401
