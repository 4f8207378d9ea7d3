"""GPU runs of the benchmark workloads, each checked against the oracle.

The benchmark times these shapes; here the same engines at the same shapes
(view blocks, entry splits, grid and slab counts) are compared with the FP64
oracle on fewer paths:
  * speckle image, 128 x 128 views (MCfield): replayed path sets <= 1e-12
    (FP64) / 2e-5 (mixed), max |u - u*| / max |u*|; a same-seed full run
    <= 1e-10 / 2e-5 (the GPU tracer reproduces the oracle's walks);
  * Fourier field, 511 x 511 grid, 256-bin three-lobe table
    (MCfieldFourierQuad): against the numpy-FFT restatement of
    MCfieldFourierQuad.m over the oracle's walks (itself pinned against the
    oracle's direct DFT), <= 1e-10 (FP64) / 2e-4 (mixed), same vertex count;
  * 4096 x 4096 covariance: the mixed result against the FP64 result on the
    same paths (test_gpu_parity.test_gpu_headline_size_mixed_vs_fp64).
"""
import os
import sys

import numpy as np
import pytest

from mcspeckle_amd import engines as E
from mcspeckle_amd import workloads as W
from test_oracle_quad import quad_numpy

pytestmark = pytest.mark.gpu

THREADS = min(16, os.cpu_count() or 1)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(u, u0):
    return float(np.abs(u - u0).max() / np.abs(u0).max())


_cache = {}


def cached(key, fn):
    """The oracle side of a case, computed once for both precisions."""
    if key not in _cache:
        _cache[key] = fn()
    return _cache[key]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-5)])
def test_speckle_image_128_replay(oracle, precision, tol):
    spec = W.speckle_image_128()
    assert spec.nv == 128 * 128 and spec.nl == 1
    n = 200
    say("speckle image replay", precision)
    ps = cached("image paths", lambda: oracle.trace(spec, n))
    u0 = cached("image replay", lambda: oracle.mcfield_replay(spec, ps, n))
    u = E.mcfield_replay(spec, ps, n, precision)
    err = rel(u, u0)
    say("  err", err)
    assert err <= tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 2e-5)])
def test_speckle_image_128_full_run(oracle, precision, tol):
    spec = W.speckle_image_128(seed=7)
    n = 1500
    say("speckle image full run", precision)
    V = np.prod(spec.box_max - spec.box_min)
    u0 = cached("image full", lambda: oracle.mcfield_sums(spec, 0, n, THREADS) * np.sqrt(V * spec.sigt / n))
    st = {}
    u = E.mcfield_spec(spec, n, precision, st)
    err = rel(u, u0)
    say("  err", err, "vertices", st["vertices"])
    assert st["paths"] == n
    assert err <= tol


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 2e-4)])
def test_fourier_field_511(oracle, precision, tol):
    a = W.fourier_field_511(seed=11)
    n = 300
    say("fourier field 511", precision)
    u0, K = cached("fourier", lambda: quad_numpy(
        oracle, a["sigt"], a["albedo"], a["box_min"], a["box_max"], np.array(a["l"], float), a["v_max"], a["v_stp"], n,
        a["lambda_"], a["doCBS"], a["smpFlg"], a["vsign"], a["seed"], sct_type=a["sct_type"], ampfunc=a["ampfunc"]))
    st = {}
    u = E.MCfieldFourierQuad(a["sigt"], a["albedo"], a["box_min"], a["box_max"], a["l"], a["v_max"], a["v_stp"], n,
                             a["lambda_"], a["doCBS"], a["smpFlg"], a["lmean0"], a["vsign"], sct_type=a["sct_type"],
                             ampfunc=a["ampfunc"], seed=a["seed"], precision=precision, stats=st)
    assert u.shape == (511, 511, 1) == u0.shape
    assert st["vertices"] == K
    err = rel(u, u0)
    say("  err", err, "vertices", K)
    assert err <= tol
