"""Learned warm start of sinkhorn_log (no counterpart in the reference).

A DeepSets encoder over the target cloud and an MLP over each source point
predict the dual potential f; the solver starts from u0 = f / reg and then
runs the ordinary iterations to the same stopThr, so the method itself
decides convergence.
"""

import os

# A warm-started problem whose first checked marginal error exceeds this
# restarts from the cold start.
FALLBACK_THR = 1e-1

WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "weights.npz")


def load_default(device):
    """The trained network, or None when no weights have been trained yet."""
    if not os.path.exists(WEIGHTS):
        return None
    from .net import WarmStartNet
    return WarmStartNet.load(WEIGHTS, device)


def predict_u0(net, R, X, Y, regs):
    """Initial log potentials (sum n,) float64 for the ragged batch R."""
    return net.predict_u0(R, X, Y, regs)
