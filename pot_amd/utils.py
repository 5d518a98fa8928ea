"""Helpers of the drop-in API: ``unif``, ``list_to_array``, ``dist``,
``euclidean_distances`` and the ``OTResult`` container.

* ``list_to_array``        ot/utils.py:59-94 (lists become NumPy float64 arrays)
* ``unif``                 ot/utils.py:239-259
* ``euclidean_distances``  ot/utils.py:270-307 (GPU kernel, clamped at 0,
                           zero diagonal when ``X is Y``)
* ``dist``                 ot/utils.py:400-527 ('sqeuclidean' / 'euclidean' on the GPU)
* ``OTResult``             ot/utils.py:1127-1400 (the fields the batch solvers fill)
* ``LazyTensor``           ot/utils.py:1425-1492 (a tensor computed from slices on demand)
"""

import numpy as np
import torch

from . import _solve
from ._device import Ingress


def list_to_array(*lst):
    """Convert Python lists to NumPy float64 arrays; other inputs pass through."""
    out = [np.array(a, dtype=np.float64) if isinstance(a, list) else a for a in lst]
    return out if len(lst) > 1 else out[0]


def unif(n, type_as=None):
    """Uniform histogram of length n (values 1/n) of the type of ``type_as``."""
    if type_as is None:
        return np.ones((n,)) / n
    if isinstance(type_as, torch.Tensor):
        return torch.ones((n,), dtype=type_as.dtype, device=type_as.device) / n
    return np.ones((n,), dtype=np.asarray(type_as).dtype) / n


def euclidean_distances(X, Y, squared=False):
    """Pairwise (squared) Euclidean distances ``|x|^2 + |y|^2 - 2 x.y``, clamped at 0."""
    same = X is Y
    ing = Ingress(X, Y)
    Xt = ing.dev(X)
    Yt = Xt if same else ing.dev(Y)
    M = _solve.sqeuclidean(Xt[None], Yt[None], squared=squared, clamp=True,
                           zero_diag=same)
    return ing.out(M[0])


def dist(x1, x2=None, metric="sqeuclidean", p=2, w=None, backend="auto", nx=None,
         use_tensor=False):
    """Cost matrix between samples; 'sqeuclidean' and 'euclidean' run on the GPU."""
    if x2 is None:
        x2 = x1
    if metric == "sqeuclidean":
        return euclidean_distances(x1, x2, squared=True)
    if metric == "euclidean":
        return euclidean_distances(x1, x2, squared=False)
    raise NotImplementedError(
        "pot_amd.dist implements the 'sqeuclidean' and 'euclidean' metrics")


class OTResult:
    """Result container with the fields of ot.utils.OTResult used by the batch API.

    ``plan``, ``potentials``, ``potential_a``, ``potential_b``, ``value``,
    ``value_linear``, ``log``, ``marginals``, ``marginal_a``, ``marginal_b``
    (marginals are the plan's row / column sums, _utils.py:1336-1360).
    """

    def __init__(self, potentials=None, value=None, value_linear=None, value_quad=None,
                 plan=None, log=None, backend=None, sparse_plan=None, lazy_plan=None,
                 status=None, batch_size=100):
        self._potentials = potentials
        self._value = value
        self._value_linear = value_linear
        self._value_quad = value_quad
        self._plan = plan
        self._log = log
        self._status = status

    def __repr__(self):
        s = "OTResult("
        if self._value is not None:
            s += "value={},".format(self._value)
        if self._value_linear is not None:
            s += "value_linear={},".format(self._value_linear)
        if self._plan is not None:
            s += "plan={}(shape={}),".format(self._plan.__class__.__name__, self._plan.shape)
        return (s[:-1] + ")") if s[-1] != "(" else s + ")"

    @property
    def potentials(self):
        return self._potentials

    @property
    def potential_a(self):
        return None if self._potentials is None else self._potentials[0]

    @property
    def potential_b(self):
        return None if self._potentials is None else self._potentials[1]

    @property
    def plan(self):
        return self._plan

    @property
    def value(self):
        return self._value

    @property
    def value_linear(self):
        return self._value_linear

    @property
    def value_quad(self):
        return self._value_quad

    @property
    def log(self):
        return self._log

    @property
    def status(self):
        return self._status

    @property
    def marginals(self):
        return None if self._plan is None else (self.marginal_a, self.marginal_b)

    @property
    def marginal_a(self):
        if self._plan is None:
            return None
        return self._plan.sum(-1)

    @property
    def marginal_b(self):
        if self._plan is None:
            return None
        return self._plan.sum(-2)

    @property
    def citation(self):
        return ""


class LazyTensor:
    """A tensor that is never stored: ``T[key]`` calls ``getitem(i, j, ...,
    **kwargs)`` on the indices (ot.utils.LazyTensor, ot/utils.py:1425-1492).

    An integer or a slice indexes the first axis (the others are taken
    whole), a tuple indexes the leading axes (the rest taken whole); any
    other key raises NotImplementedError.  The keyword arguments are kept as
    attributes of the same names.
    """

    def __init__(self, shape, getitem, **kwargs):
        self._getitem = getitem
        self.shape = tuple(shape)
        self.ndim = len(self.shape)
        self.kwargs = kwargs
        for key, value in kwargs.items():
            setattr(self, key, value)

    def __getitem__(self, key):
        if isinstance(key, (int, np.integer, slice)):
            k = [key]
        elif isinstance(key, tuple):
            k = list(key)
        else:
            raise NotImplementedError("Only integer, slice, and tuple indexing is supported")
        k += [slice(None)] * (self.ndim - len(k))
        return self._getitem(*k, **self.kwargs)

    def __repr__(self):
        return "LazyTensor(shape={},attributes=({}))".format(self.shape,
                                                              ",".join(self.kwargs.keys()))
