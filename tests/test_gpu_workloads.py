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


def paths_at_lateral_faces(ps, box_min, box_max, seed=5):
    """The same walks, each moved laterally by its own offset so that it
    touches a lateral face (x and y, min or max side, within 1e-3 .. 0.3) of
    the 1e5-wide slab: views leave such a vertex through the lateral face
    after a short distance, the case where an FP32 exit distance formed as
    W/a - (x - box_min)/a would lose every bit."""
    rng = np.random.default_rng(seed)
    ps = {k: v.copy() for k, v in ps.items()}
    off = ps["offset"]
    x1, x = ps["x1"].reshape(-1, 3), ps["x"].reshape(-1, 3)
    for p in range(off.size - 1):
        k0, k1 = off[p], off[p + 1]
        pts = np.vstack([x1[p:p + 1], x[k0:k1]])
        for j in (0, 1):
            gap = rng.uniform(1e-3, 0.3)
            d = (box_max[j] - gap - pts[:, j].max()) if rng.integers(2) else (box_min[j] + gap - pts[:, j].min())
            x1[p, j] += d
            x[k0:k1, j] += d
    return ps


@pytest.mark.parametrize("cbs,nl", [(False, 1), (True, 3)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-5)])
def test_speckle_image_paths_at_lateral_faces(oracle, precision, tol, cbs, nl):
    base = W.speckle_image_128()
    l = np.array([[0.0, 0.05, -0.03], [0.0, 0.0, 0.04], [1.0, 1.0, 1.0]])[:, :nl]
    l /= np.linalg.norm(l, axis=0)
    spec = E.make_problem(base.sigt, base.albedo, base.box_min, base.box_max, l, W.view_grid(64, 10), 1, 1, 1.0, cbs,
                          2, 3, 0.8, seed=3)
    n = 200
    say("speckle paths at lateral faces", precision, "cbs", cbs, "lights", nl)
    ps = cached(("edge paths", cbs, nl), lambda: paths_at_lateral_faces(oracle.trace(spec, n), spec.box_min,
                                                                        spec.box_max))
    u0 = cached(("edge replay", cbs, nl), lambda: oracle.mcfield_replay(spec, ps, n))
    u = E.mcfield_replay(spec, ps, n, precision)
    err = rel(u, u0)
    say("  err", err)
    assert err <= tol


@pytest.mark.parametrize("wave", [False, True])
def test_speckle_image_mixed_vs_fp64_many_paths(wave):
    """2e5 paths over the whole 1e5-wide slab (a few dozen vertices land
    within a view's exit distance of a lateral face): mixed against FP64 on
    the same paths, max |du| / max |u| <= 2e-5, for MCfield and for
    MCfieldOnWave (its first vertices go through the matrix-core kernel)."""
    base = W.speckle_image_128(seed=9)
    v = W.view_grid(32, 10)
    n = 200000
    say("speckle mixed vs fp64, wave", wave)
    if wave:
        rng = np.random.default_rng(1)
        l = np.vstack([rng.uniform(-0.1, 0.1, (2, 40)), np.ones((1, 40))])
        l /= np.linalg.norm(l, axis=0)
        wl = rng.normal(size=40) + 1j * rng.normal(size=40)
        args = (base.sigt, base.albedo, base.box_min, base.box_max, l, v, 1, 1, n, 1.0, True, 2, 3, 0.8)
        u64 = E.MCfieldOnWave(*args, Wl=wl, seed=9, precision="fp64")
        um = E.MCfieldOnWave(*args, Wl=wl, seed=9, precision="mixed")
    else:
        spec = E.make_problem(base.sigt, base.albedo, base.box_min, base.box_max, [[0], [0], [1]], v, 1, 1, 1.0,
                              False, 2, 3, 0.8, seed=9)
        u64 = E.mcfield_spec(spec, n, "fp64")
        um = E.mcfield_spec(spec, n, "mixed")
    err = rel(um, u64)
    say("  err", err)
    assert err <= 2e-5


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
