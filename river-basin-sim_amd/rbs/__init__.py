"""rbs -- river-basin-sim's solver time step on MI355X (gfx950).

The host side keeps the reference's contracts (the ``compute_dt``/``step`` scheme
seam, ``State``, ``MassLedger``, the multi-rate scheduler and the run loop); every
kernel runs in ``librbs.so`` (hand-written HIP) through a ctypes C-ABI
(``include/rbs.h``).  There is no CPU fallback.
"""

__version__ = "0.1.0"
