"""pot_amd -- MI355X (gfx950) drop-in for POT's entropic Sinkhorn hot path.

Public surface (same names and arguments as POT 0.9.7):

* ``sinkhorn``, ``sinkhorn2``                     (ot.sinkhorn, ot.sinkhorn2)
* ``bregman.sinkhorn_knopp``, ``bregman.sinkhorn_log``
* ``batch.solve_batch``, ``batch.solve_sample_batch``,
  ``batch.bregman_log_projection_batch``, ``batch.bregman_projection_batch``,
  ``batch.dist_batch`` and the batch losses
* ``dist``, ``unif``, ``utils.OTResult``

All iterations run in ``libpot_amd.so`` (hand-written HIP for gfx950, C ABI
in include/pot_amd.h); nothing falls back to a CPU implementation.
"""

from . import batch, bregman, utils
from .batch import dist_batch, solve_batch, solve_sample_batch
from .bregman import sinkhorn, sinkhorn2
from .utils import dist, unif

__version__ = "0.1.0"

__all__ = [
    "batch", "bregman", "utils", "sinkhorn", "sinkhorn2", "dist", "unif",
    "solve_batch", "solve_sample_batch", "dist_batch",
]
