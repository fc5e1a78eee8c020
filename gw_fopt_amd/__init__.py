"""MI355X (gfx950) generalized bulk-flow gravitational-wave spectrum.

A drop-in for the many-bubble spectrum path of hoquoctrung98/gw_fopt
(``bubble_gw::many_bubbles::GeneralizedBulkFlow``): the C ABI of
``libbubble_gw_c.so`` (declared in ``include/bubble_gw.h``) and the Python
surface of ``gw_fopt.bubble_gw.many_bubbles`` (``gw_fopt_amd.many_bubbles``).
"""
from . import many_bubbles  # noqa: F401

__version__ = "0.1.0"
