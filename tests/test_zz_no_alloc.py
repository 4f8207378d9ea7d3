"""scmc_mcfield_noalloc: MCfield with every device buffer in the library's
fixed scratch array.

CPU: a problem whose fixed part does not fit the scratch is rejected with
SCMC_E_ARG and a message naming the limit, before any device work.
GPU: the allocation-free call gives scmc_mcfield's field on the same walks
(<= 1e-12 FP64, <= 2e-5 mixed, max |u - u*| / max |u*|) on two calls in a
row, the second smaller than the first, and on walks long enough that a chunk's
paths overflow their vertex share (the prefix of paths that fits is emitted
and the rest traced in the next chunk).
"""
import sys

import numpy as np
import pytest

from mcspeckle_amd import _native as N
from mcspeckle_amd import engines as E
from mcspeckle_amd import workloads as W


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def rel(u, u0):
    return float(np.abs(u - u0).max() / np.abs(u0).max())


def test_too_large_a_problem_is_rejected_before_device_work():
    rng = np.random.default_rng(0)
    v = W.polar_dirs(rng, 200000, 10)
    spec = E.make_problem(2.0, 0.9, [-5, -5, 0], [5, 5, 1], [[0], [0], [1]], v, 1, 1, 1.0, False, 2, 3, 0.5)
    with pytest.raises(N.ScmcError) as ei:
        E.mcfield_noalloc_spec(spec, 10, "mixed")
    assert ei.value.code == N.SCMC_E_ARG
    assert "without allocation" in str(ei.value) and "scratch" in str(ei.value)


CALLS = [
    # (problem, paths): a larger call first, then a smaller one
    ("views64x64-lights2", 20000),
    ("views16-light1", 3000),
]


def problem(name):
    rng = np.random.default_rng(len(name))
    if name == "views64x64-lights2":
        v = W.view_grid(64, 8)
        l = W.polar_dirs(rng, 2, 5)
        return E.make_problem(6.0, 0.97, [-5e3, -5e3, 0], [5e3, 5e3, 1], l, v, 1, 1, 1.0, True, 2, 3, 0.8, seed=3)
    if name == "views16-light1":
        v = W.polar_dirs(rng, 16, 12)
        return E.make_problem(3.0, 0.9, [-5e3, -5e3, 0], [5e3, 5e3, 1.5], [[0], [0], [1]], v, 1, 1, 0.7, False, 2,
                              1, 0, seed=4)
    # long walks: a thick, almost conservative slab entered anywhere (smpFlg 1)
    v = W.polar_dirs(rng, 16, 10)
    return E.make_problem(60.0, 0.999, [-5e3, -5e3, 0], [5e3, 5e3, 1], [[0], [0], [1]], v, 1, 1, 1.0, True, 1, 1, 0,
                          seed=5)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-5)])
def test_noalloc_matches_mcfield_on_two_calls_in_a_row(precision, tol):
    for name, n in CALLS:
        spec = problem(name)
        say("noalloc", name, n, precision)
        st = {}
        u = E.mcfield_noalloc_spec(spec, n, precision, st)
        u0 = E.mcfield_spec(spec, n, precision)
        err = rel(u, u0)
        say("  err", err, "chunks", st["chunks"], "vertices", st["vertices"])
        assert st["paths"] == n
        assert err <= tol


@pytest.mark.gpu
def test_noalloc_long_walks_split_chunks_by_prefix():
    """~870 vertices per path against a share of 48: 4000 paths fit one chunk
    by count (fewer than the chunk's path capacity) but not by vertices, so
    more than one chunk means the prefix cut ran."""
    spec = problem("long walks")
    n = 4000
    say("noalloc long walks")
    st, st0 = {}, {}
    u = E.mcfield_noalloc_spec(spec, n, "fp64", st)
    u0 = E.mcfield_spec(spec, n, "fp64", st0)
    say("  vertices per path", st["vertices"] / n, "chunks", st["chunks"])
    assert st["vertices"] == st0["vertices"] and st["paths"] == n
    assert st["vertices"] / n > 48 and st["chunks"] >= 2
    assert rel(u, u0) <= 1e-12
