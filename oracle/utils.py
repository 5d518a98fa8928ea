"""Float64 NumPy restatement of the POT helpers used on the hot path.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

* ``unif``                 ot/utils.py:239-259
* ``euclidean_distances``  ot/utils.py:270-307 (clamped at 0, zero diagonal if X is Y)
* ``dist``                 ot/utils.py:400-527 (sqeuclidean / euclidean branches)
* ``dist0``                ot/utils.py:529-550
* ``make_1D_gauss``        ot/datasets.py:14-33

Pinning: ``dist`` against a brute-force double loop (the check of
test/test_utils.py:207-227, atol 1e-14) in tests/test_oracle.py.
"""

import numpy as np


def unif(n):
    """Uniform histogram ``1/n`` of length n."""
    return np.ones((n,)) / n


def euclidean_distances(X, Y, squared=False):
    """Pairwise (squared) Euclidean distances via ``|x|^2 + |y|^2 - 2 x.y``."""
    c = -2.0 * (X @ Y.T)
    c += np.einsum("ij,ij->i", X, X)[:, None]
    c += np.einsum("ij,ij->i", Y, Y)[None, :]
    c = np.maximum(c, 0)
    if not squared:
        c = np.sqrt(c)
    if X is Y:
        c = c * (1 - np.eye(X.shape[0]))
    return c


def dist(x1, x2=None, metric="sqeuclidean"):
    """``ot.dist`` for the two metrics the hot path uses."""
    if x2 is None:
        x2 = x1
    if metric == "sqeuclidean":
        return euclidean_distances(x1, x2, squared=True)
    if metric == "euclidean":
        return euclidean_distances(x1, x2, squared=False)
    raise NotImplementedError(metric)


def dist0(n):
    """``lin_square`` cost of size (n, n): ``(i - j)^2``."""
    x = np.arange(n, dtype=np.float64).reshape((n, 1))
    return dist(x, x)


def make_1D_gauss(n, m, s):
    """Normalised 1-D Gaussian histogram on ``n`` bins."""
    x = np.arange(n, dtype=np.float64)
    h = np.exp(-((x - m) ** 2) / (2 * s ** 2))
    return h / h.sum()
