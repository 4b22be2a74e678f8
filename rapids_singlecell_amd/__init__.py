"""MI355X-native Harmony integration with rapids_singlecell's interface.

``rapids_singlecell_amd.pp.harmony_integrate`` is a drop-in for
``rapids_singlecell.pp.harmony_integrate``; the hot path runs in hand-written
gfx950 HIP kernels (``libharmony_hip.so``) called through a C ABI.
"""

from __future__ import annotations

from . import pp

__all__ = ["pp"]
__version__ = "0.1.0"
