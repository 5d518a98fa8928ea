"""CPU checks of the C-ABI library: it is built, current, exports every symbol
include/pot_amd.h declares, and rejects bad arguments before touching a GPU."""

import ctypes
import glob
import os
import re

import pytest

from pot_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = sorted(glob.glob(os.path.join(ROOT, "include", "*.h")))


def _declared_functions():
    names = set()
    for h in HEADERS:
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(potamd_\w+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_library_exists_and_is_current():
    assert os.path.exists(_lib.LIB_PATH), "run python -m pot_amd.build"
    lib_mtime = os.path.getmtime(_lib.LIB_PATH)
    srcs = glob.glob(os.path.join(ROOT, "pot_amd", "csrc", "*")) + HEADERS
    stale = [s for s in srcs if os.path.getmtime(s) > lib_mtime + 1e-6]
    assert not stale, f"libpot_amd.so is older than {stale}; rebuild"


def test_every_declared_symbol_is_exported_and_bound():
    declared = _declared_functions()
    assert declared, "no declarations found"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == set(_lib.SIGNATURES), declared ^ set(_lib.SIGNATURES)
    _lib.load()


def test_status_strings_and_version():
    lib = _lib.load()
    assert lib.potamd_version() >= 100
    for code in (0, 1, 2, 3, -1, -2, -3, -4):
        assert lib.potamd_status_string(code)


def test_dense_argument_validation():
    lib = _lib.load()
    P = ctypes.c_void_p(0x1000)
    info = (ctypes.c_int * 4)()

    def call(**kw):
        args = dict(dtype=0, domain=0, order=0, kg=0, M=P, ldm=8, bs=0, B=1, n=8, m=8,
                    reg=0.1, max_iter=10, err_norm=0, ws=None, wsb=0)
        args.update(kw)
        return lib.potamd_sinkhorn_dense(
            args["dtype"], args["domain"], args["order"], args["kg"], args["M"],
            args["ldm"], args["bs"], args["B"], args["n"], args["m"], args["reg"],
            P, P, P, P, args["max_iter"], 1e-9, args["err_norm"], P, 4, 0, 0, 0, 0,
            args["ws"], args["wsb"], None, None, info)

    assert call() == -4                      # valid, but no workspace
    assert call(dtype=5) == -1
    assert call(ldm=4) == -1                 # ldm < m
    assert call(reg=0.0) == -1               # log domain needs reg > 0
    assert call(domain=1, reg=0.0) == -4     # scaling domain accepts reg = 0 (nan guard)
    assert call(max_iter=0) == -1
    assert call(err_norm=3) == -1
    assert call(m=1 << 20, ldm=1 << 20) == -3  # beyond the compiled tile set


def test_plan_and_dist_argument_validation():
    lib = _lib.load()
    P = ctypes.c_void_p(0x1000)
    rc = lib.potamd_plan_dense(0, 0, 0, P, 8, 0, 1, 8, 8, 0.1, P, P, None, None, P, None,
                               64, None, None, 0, None)
    assert rc == -4
    rc = lib.potamd_plan_dense(0, 0, 0, P, 4, 0, 1, 8, 8, 0.1, P, P, None, None, P, None,
                               64, None, None, 0, None)
    assert rc == -1
    rc = lib.potamd_sqeuclidean(0, P, P, 1, 4, 4, 0, 16, 16, P, 4, 16, 1, 1, 0, None)
    assert rc == -1


def test_pairwise_and_gemm_argument_validation():
    # every rejection happens before any launch, so no GPU is touched
    lib = _lib.load()
    P = ctypes.c_void_p(0x1000)
    assert lib.potamd_pairwise_batch(2, 0, P, P, 1, 4, 4, 3, 2.0, 1.0, 0.0, P, None) == -1
    assert lib.potamd_pairwise_batch(0, 3, P, P, 1, 4, 4, 3, 2.0, 1.0, 0.0, P, None) == -1
    assert lib.potamd_pairwise_batch(0, 0, P, P, 1, 4, 4, 3, 0.0, 1.0, 0.0, P, None) == -1
    assert lib.potamd_pairwise_batch(0, 1, P, P, 1, 4, 4, 3, 0.0, 0.0, -1.0, P, None) == -1
    assert lib.potamd_pairwise_batch(0, 0, P, P, 1, 4, 0, 3, 2.0, 1.0, 0.0, P, None) == -1
    assert lib.potamd_gemm_batched(0, 1, 4, 4, 4, 1.0, P, 16, 4, 1, P, 16, 4, 1, 0.0, P, 16, 3,
                                   None) == -1   # ldc < N
    assert lib.potamd_gemm_batched(0, 1, 4, 4, 0, 1.0, P, 16, 4, 1, P, 16, 4, 1, 0.0, P, 16, 4,
                                   None) == -1   # K = 0
    assert lib.potamd_gemm_batched(0, 1, 4, 4, 4, 1.0, None, 16, 4, 1, P, 16, 4, 1, 0.0, P, 16,
                                   4, None) == -1


def test_no_gpu_means_loud_failure():
    torch = pytest.importorskip("torch")
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    import numpy as np
    import pot_amd
    with pytest.raises(RuntimeError, match="ROCm GPU"):
        pot_amd.sinkhorn([0.5, 0.5], [0.5, 0.5], np.eye(2), 1.0)


def test_batch_tuner_key_and_lookup(monkeypatch):
    import numpy as np

    from pot_amd import autotune, ragged
    ns = np.full(8, 1000)
    ms = np.full(8, 700)
    key = autotune._batch_key("points", ns, ms, 16)
    assert key.endswith("|batch-points|1024x1024|mean1024x512|d16")
    monkeypatch.setattr(autotune, "_mem", {key: {"config": [16, 0, 0]}})
    assert autotune.lookup_batch("points", ns, ms, 16) == (16, 0, 0)
    assert ragged._tuned("points", ns, ms, 16, 0) == (16, 0, 0)
    # an explicit configuration wins; an untuned class gets the library default
    assert ragged._tuned("points", ns, ms, 16, 8) == (8, 0, 0)
    assert ragged._tuned("dense", ns, ms, 0, 0) == (0, 0, 0)
    assert autotune.lookup_batch("points", [], [], 16) is None


def test_split_plan_tiers_and_budget():
    import numpy as np

    from pot_amd import ragged
    # small batches of large problems: 8 parts each while they fit the CUs
    big = np.full(16, 1024)
    assert (ragged.split_plan(big, big, cus=256) == 8).all()
    # 64 of them: 8-part problems fall back to 4 before anyone gives up parts
    big = np.full(64, 1024)
    assert (ragged.split_plan(big, big, cus=256) == 4).all()
    # size tiers and their row minima
    ns = np.array([1024, 600, 300, 200, 120, 60, 1024, 250])
    ms = np.array([1024, 700, 600, 700, 4000, 4000, 100, 4000])
    np.testing.assert_array_equal(ragged.split_plan(ns, ms, cus=256),
                                  [8, 4, 2, 2, 2, 1, 1, 4])
    # over budget: the smallest split problems give up parts first; with regs
    # the size is weighed as n m / reg
    ns = np.array([512, 512, 640])
    ms = np.array([512, 512, 640])
    np.testing.assert_array_equal(ragged.split_plan(ns, ms, cus=8), [2, 2, 4])
    np.testing.assert_array_equal(ragged.split_plan(ns, ms, cus=6), [1, 2, 4])
    np.testing.assert_array_equal(ragged.split_plan(ns, ms, cus=6, regs=[0.01, 0.1, 0.1]),
                                  [2, 1, 4])
    # batches of more than 2 problems per CU are not split
    assert (ragged.split_plan(np.full(9, 1024), np.full(9, 1024), cus=4) == 1).all()
    with pytest.raises(ValueError):
        ragged._check_parts([9], [1024])
    with pytest.raises(ValueError):
        ragged._check_parts([8], [255])


def test_points_and_network_argument_validation():
    # invalid arguments are rejected before any device call
    lib = _lib.load()
    P = ctypes.c_void_p(0x1000)
    Q = ctypes.c_void_p(0x1008)  # not 16-byte aligned

    def pts(**kw):
        a = dict(batch=1, max_n=8, max_m=8, d=2, normalize=1, max_iter=10, check=1, waves=0,
                 cmax=P, queue=None)
        a.update(kw)
        return lib.potamd_sinkhorn_log_batch_points(
            a["batch"], a["max_n"], a["max_m"], a["d"], P, P, P, P, P, P, P, a["normalize"],
            P, P, P, P, a["max_iter"], a["check"], 1e-9, 0.0, None, P, P, P, P, a["cmax"],
            a["waves"], a["queue"], None)

    assert pts(batch=0) == -1
    assert pts(max_iter=0) == -1
    assert pts(normalize=3) == -1
    assert pts(normalize=2, cmax=None) == -1
    assert pts(max_n=4096) == -3
    assert pts(d=17) == -3
    assert pts(waves=12) == -1
    assert lib.potamd_ragged_points_cmax(1, 8, 8, 2, P, P, P, P, P, P, None, None) == -1
    assert lib.potamd_ragged_points_cmax(1, 8, 8, 64, P, P, P, P, P, P, P, None) == -3

    def mlp(**kw):
        a = dict(mode=0, rows=1024, seg=1024, inp=P, pts=None)
        a.update(kw)
        return lib.potamd_mlp_fused(a["mode"], a["rows"], a["seg"], a["inp"], a["pts"], None,
                                    None, None, P, P, P, P, P, P, None, P, None)

    assert mlp(mode=2) == -1
    assert mlp(rows=0) == -1
    assert mlp(seg=100, rows=1000) == -3     # segments must be whole 64-row blocks
    assert mlp(inp=None) == -1               # no input rows and no points
    assert mlp(inp=Q) == -1                  # 16-byte loads need alignment


def test_noalloc_rejects_more_slots_than_its_scratch():
    # checked on the host before any GPU work: nsplit above the fixed maximum
    lib = _lib.load()
    P = ctypes.c_void_p(0x1000)
    assert lib.potamd_noalloc_scratch_bytes() >= 256 + 256 * 66048
    for nsplit, want in ((257, -4), (-1, -1)):
        rc = lib.potamd_sinkhorn_log_batch_points_split_noalloc(
            1, 64, 64, 16, P, P, P, P, P, P, P, 1, P, P, P, P, 10, 10, 1e-6, 0.0, 1, P, P,
            nsplit, P, P, P, P, P, ctypes.c_void_p(0))
        assert rc == want, (nsplit, rc)


def test_lazy_tensor_indexing_semantics():
    # test/test_utils.py:522-558 with a NumPy getitem
    import numpy as np

    from pot_amd.utils import LazyTensor
    rng = np.random.RandomState(42)
    x1, x2 = rng.randn(100, 2), rng.randn(200, 2)

    def getitem(i, j, x1, x2):
        return np.dot(x1[i], x2[j].T)

    T = LazyTensor((100, 200), getitem, x1=x1, x2=x2)
    assert T.shape == (100, 200)
    assert str(T) == "LazyTensor(shape=(100, 200),attributes=(x1,x2))"
    assert T.x1 is x1 and T.x2 is x2
    assert T[:].shape == (100, 200)
    assert T[1, 1] == np.dot(x1[1], x2[1].T)
    assert T[1].shape == (200,)
    assert T[::10, 5].shape == (10,)
    with pytest.raises(NotImplementedError):
        T["error"]


def test_bregman_namespace_has_the_path_functions():
    # the names ot.bregman exports for this path (ot/bregman/__init__.py:13-43)
    import pot_amd.bregman as b
    for name in ("sinkhorn", "sinkhorn2", "sinkhorn_knopp", "sinkhorn_log", "greenkhorn",
                 "sinkhorn_stabilized", "sinkhorn_epsilon_scaling", "empirical_sinkhorn",
                 "empirical_sinkhorn2"):
        assert callable(getattr(b, name)), name
    with pytest.raises(ValueError):
        b.sinkhorn([0.5, 0.5], [0.5, 0.5], [[0.0, 1.0], [1.0, 0.0]], 1.0, method="unknown")
