"""GPU MCCovHetro and MCsampleHetro (voxel media) against the oracle, same
Philox walks.

Tolerances (Cauchy-Schwarz metric for Ms, Mm; max |d| / max |ref| for mean1
and for the field u): FP64 <= 1e-10 (walks identical, transcendental ulps
aside), mixed precision <= 1e-4 for the covariance; mean1 is accumulated
from FP64 factors in both modes (<= 1e-10).  MCsampleHetro's u: FP64 <=
1e-10; mixed <= 2e-5 in the far field (FP64 march and phase reduction,
FP32 sin/cos/exp, as MCfield's mixed mode), FP64 in the near field
(<= 1e-10).  A uniform one-cell medium equals the GPU MCcov / MCfield to
1e-12.
"""
import sys

import numpy as np
import pytest

from conftest import cov_matrix, cs_scale, far_dirs
from mcspeckle_amd import engines as E
from test_oracle_engines import near_far_pair, small_spec, tab3d

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cs_err(M, Mref):
    C, R = cov_matrix(M), cov_matrix(Mref)
    return float(np.max(np.abs(C - R) / np.maximum(cs_scale(R), 1e-300)))


def _amp(spec):
    return spec.amp.g if spec.amp.type == 3 else (0 if spec.amp.type == 1 else tab3d())


def run_gpu(spec, grid, bb, n, precision, albedo=None, lrad=None):
    alb = spec.albedo if albedo is None else albedo
    return E.MCCovHetro(grid, alb, spec.box_min, spec.box_max, bb, spec.l, spec.v, spec.is_ff_l, spec.is_ff_v, n,
                        spec.lam, spec.do_cbs, spec.smp_flg, spec.amp.type, _amp(spec), lrad=lrad, seed=spec.seed,
                        precision=precision)


@pytest.mark.parametrize("kind,dim,cbs", [("hg", 3, True), ("iso", 2, False), ("tab", 3, True)])
def test_gpu_one_cell_medium_is_mccov(oracle, kind, dim, cbs):
    spec = small_spec(kind, dim, cbs, nv=4, nl=3, seed=5, smp=1, tab=tab3d() if kind == "tab" else None)
    say("one cell", kind, dim, cbs)
    n = 1500
    Ms, Mm, _ = run_gpu(spec, np.full((1,) * dim, spec.sigt), spec.box_max - spec.box_min, n, "fp64")
    Ms0, Mm0 = E.mccov_spec(spec, n, "fp64")
    assert cs_err(Ms, Ms0) <= 1e-12 and cs_err(Mm, Mm0) <= 1e-12


def grid3(seed=1, zeros=True):
    rng = np.random.default_rng(seed)
    g = rng.uniform(0.5, 4.0, (3, 4, 3))
    if zeros:
        g[1, 2, 0] = 0.0
        g[2, 0, 2] = 0.0
    return g


CASES = [pytest.param(*c, id="-".join(map(str, c))) for c in [
    ("hg", 3, True, 1, "far", None, None), ("hg", 3, True, 2, "far", None, None),
    ("iso", 3, False, 2, "far", "grid", None), ("tab", 3, True, 1, "far", None, None),
    ("hg", 3, True, 2, "both", None, None), ("hg", 3, False, 2, "far", None, 4.0),
    ("hg", 2, True, 2, "far", "grid", None), ("iso", 2, True, 1, "far", None, None)]]


def case_spec(kind, dim, cbs, smp, near):
    if near != "far":
        _, sp = near_far_pair(kind, dim, cbs, 4.0, near, nv=5, nl=3)
        return E.make_problem(sp.sigt, sp.albedo, sp.box_min, sp.box_max, sp.l, sp.v, sp.is_ff_l, sp.is_ff_v, sp.lam,
                              cbs, smp, sp.amp.type, _amp(sp), seed=sp.seed)
    return small_spec(kind, dim, cbs, nv=5, nl=3, seed=7, smp=smp, tab=tab3d() if kind == "tab" else None)


@pytest.mark.parametrize("kind,dim,cbs,smp,near,albg,lrad", CASES)
def test_gpu_hetro_matches_oracle(oracle, kind, dim, cbs, smp, near, albg, lrad):
    spec = case_spec(kind, dim, cbs, smp, near)
    if dim == 3:
        grid = grid3()
        bb = (spec.box_max - spec.box_min) / np.array(grid.shape)
    else:
        grid = np.random.default_rng(2).uniform(0.5, 4.0, (4, 3))
        grid[1, 1] = 0
        bb = (spec.box_max - spec.box_min) / np.array(grid.shape)
    alb = None
    if albg:
        alb = np.random.default_rng(3).uniform(0.6, 1.0, grid.shape)
    say("hetro", kind, dim, cbs, smp, near, albg, lrad)
    n = 800
    Ms0, Mm0, m0, K = oracle.mccov_hetro(spec, grid, bb, n, albedo=alb, lrad=lrad)
    assert K > 0
    for precision, tol in (("fp64", 1e-10), ("mixed", 1e-4)):
        Ms, Mm, m1 = run_gpu(spec, grid, bb, n, precision, albedo=alb, lrad=lrad)
        e_s, e_m = cs_err(Ms, Ms0), cs_err(Mm, Mm0)
        e_mean = np.abs(m1 - m0).max() / np.abs(m0).max()
        say("  ", precision, e_s, e_m, e_mean)
        assert e_s <= tol and e_m <= tol
        assert e_mean <= 1e-10


def test_gpu_hetro_lrad_smpflg1_one_light(oracle):
    rng = np.random.default_rng(4)
    l, v = far_dirs(1, 10, rng), far_dirs(4, 20, rng)
    spec = E.make_problem(2.0, 0.9, [-3, -3, 0], [3, 3, 1.5], l, v, 1, 1, 0.8, True, 1, 3, 0.7, seed=3)
    grid = grid3(5, zeros=False)
    bb = (spec.box_max - spec.box_min) / np.array(grid.shape)
    say("hetro lrad smpFlg 1")
    Ms0, Mm0, m0, _ = oracle.mccov_hetro(spec, grid, bb, 600, lrad=1.5)
    Ms, Mm, m1 = run_gpu(spec, grid, bb, 600, "fp64", lrad=1.5)
    assert cs_err(Ms, Ms0) <= 1e-10 and cs_err(Mm, Mm0) <= 1e-10
    assert np.abs(m1 - m0).max() <= 1e-10 * np.abs(m0).max()


# ---------------------------------------------------------------- MCsampleHetro

def run_gpu_sample(spec, grid, bb, n, precision, albedo=None):
    alb = spec.albedo if albedo is None else albedo
    amp0 = None if spec.amp0 is None else spec.amp0.g   # HG ampfunc0 in these tests
    return E.MCsampleHetro(grid, alb, spec.box_min, spec.box_max, bb, spec.l, spec.v, spec.is_ff_l, spec.is_ff_v, n,
                           spec.lam, spec.do_cbs, spec.smp_flg, spec.amp.type, _amp(spec), amp0,
                           lmean0=spec.lmean0, seed=spec.seed, precision=precision)


@pytest.mark.parametrize("kind,dim,cbs", [("hg", 3, True), ("iso", 2, False), ("tab", 3, True)])
def test_gpu_sample_one_cell_medium_is_mcfield(oracle, kind, dim, cbs):
    spec = small_spec(kind, dim, cbs, nv=4, nl=3, seed=5, smp=1, tab=tab3d() if kind == "tab" else None)
    say("sample one cell", kind, dim, cbs)
    n = 1500
    u = run_gpu_sample(spec, np.full((1,) * dim, spec.sigt), spec.box_max - spec.box_min, n, "fp64")
    u0 = E.mcfield_spec(spec, n, "fp64")
    assert np.abs(u - u0).max() <= 1e-12 * np.abs(u0).max()


SAMPLE_CASES = [pytest.param(*c, id="-".join(map(str, c))) for c in [
    ("hg", 3, True, 1, "far", None), ("hg", 3, True, 2, "far", None), ("iso", 3, False, 2, "far", "grid"),
    ("tab", 3, True, 1, "far", None), ("hg", 3, True, 2, "both", None), ("hg", 2, True, 2, "far", "grid"),
    ("iso", 2, True, 1, "far", None)]]


@pytest.mark.parametrize("kind,dim,cbs,smp,near,albg", SAMPLE_CASES)
def test_gpu_sample_hetro_matches_oracle(oracle, kind, dim, cbs, smp, near, albg):
    spec = case_spec(kind, dim, cbs, smp, near)
    if dim == 3:
        grid = grid3()
    else:
        grid = np.random.default_rng(2).uniform(0.5, 4.0, (4, 3))
        grid[1, 1] = 0
    bb = (spec.box_max - spec.box_min) / np.array(grid.shape)
    alb = np.random.default_rng(3).uniform(0.6, 1.0, grid.shape) if albg else None
    say("sample hetro", kind, dim, cbs, smp, near, albg)
    n = 800
    u0, K = oracle.mcfield_hetro(spec, grid, bb, n, albedo=alb)
    assert K > 0
    for precision, tol in (("fp64", 1e-10), ("mixed", 1e-10 if near != "far" else 2e-5)):
        u = run_gpu_sample(spec, grid, bb, n, precision, albedo=alb)
        err = np.abs(u - u0).max() / np.abs(u0).max()
        say("  ", precision, err)
        assert err <= tol


def test_gpu_sample_hetro_lmean0(oracle):
    rng = np.random.default_rng(6)
    l, v = far_dirs(2, 20, rng), far_dirs(5, 20, rng)
    spec = E.make_problem(2.0, 0.9, [-3, -3, 0], [3, 3, 1.5], l, v, 1, 1, 0.8, True, 2, 3, 0.7, 0.4,
                          lmean0=[0.3, 0.1, 1.0], seed=2)
    grid = grid3(7, zeros=False)
    bb = (spec.box_max - spec.box_min) / np.array(grid.shape)
    say("sample hetro lmean0")
    u0, _ = oracle.mcfield_hetro(spec, grid, bb, 600)
    u = run_gpu_sample(spec, grid, bb, 600, "fp64")
    assert np.abs(u - u0).max() <= 1e-10 * np.abs(u0).max()


def test_gpu_sample_hetro_rejects_lrad():
    import ctypes as C
    from mcspeckle_amd import _native as N
    spec = small_spec("hg", 3, True, nv=2, nl=2, seed=1, smp=2)
    grid = grid3(9, zeros=False)
    bb = (spec.box_max - spec.box_min) / np.array(grid.shape)
    m, keep = E.c_medium(spec, grid, bb, None, 1.0)
    p = E.c_problem(spec)
    re, im = np.empty(4), np.empty(4)
    rc = N.lib().scmc_mcfield_hetro(C.byref(p), C.byref(m), 10, 0, N.dptr(re), N.dptr(im), None)
    assert rc != 0
    assert b"lrad" in N.lib().scmc_last_error()
