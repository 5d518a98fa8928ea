"""pot_amd -- MI355X (gfx950) drop-in for POT's entropic Sinkhorn hot path.

Public surface (same names and arguments as POT 0.9.7):

* ``sinkhorn``, ``sinkhorn2``                     (ot.sinkhorn, ot.sinkhorn2)
* ``bregman.sinkhorn_knopp``, ``bregman.sinkhorn_log``
* ``batch.solve_batch``, ``batch.solve_sample_batch``,
  ``batch.bregman_log_projection_batch``, ``batch.bregman_projection_batch``,
  ``batch.dist_batch`` and the batch losses
* ``solve``, ``solve_sample``                      (ot.solve, ot.solve_sample; the
  balanced entropic route)
* ``dist``, ``unif``, ``utils.OTResult``
* ``empirical_sinkhorn``, ``empirical_sinkhorn2`` (ot.bregman; ``isLazy=True``
  runs on the online-cost solver)
* ``online.sinkhorn_online`` -- sinkhorn_log on the squared Euclidean cost of
  two point clouds without storing the cost (MFMA x . y^T in fp32, bf16,
  compensated bf16, fp8, or fp64)

All iterations run in ``libpot_amd.so`` (hand-written HIP for gfx950, C ABI
in include/pot_amd.h); nothing falls back to a CPU implementation.
"""

from . import batch, bregman, empirical, online, solvers, utils
from .empirical import empirical_sinkhorn, empirical_sinkhorn2
from .batch import dist_batch, solve_batch, solve_sample_batch
from .bregman import sinkhorn, sinkhorn2
from .solvers import solve, solve_sample
from .utils import dist, unif

__version__ = "0.1.0"

__all__ = [
    "batch", "bregman", "online", "solvers", "utils", "sinkhorn", "sinkhorn2", "dist", "unif",
    "solve", "solve_sample", "solve_batch", "solve_sample_batch", "dist_batch",
    "empirical", "empirical_sinkhorn", "empirical_sinkhorn2",
]
