"""GPU illumination-wave variants against the oracle (same Philox walks).

* MCfieldOnWave: max |u - u*| / max |u*| <= 1e-10 in FP64 and <= 2e-5 in
  mixed precision (the same bounds as MCfield); and, on the GPU itself, the
  wave field equals MCfield contracted with the weights (linearity of every
  term of MCfieldOnWave.m) to 1e-12.
* MCfieldFourierQuadOnWave: <= 1e-10 (FP64) / <= 2e-4 (mixed), as
  MCfieldFourierQuad, including view grids of even size (even transform).
"""
import sys

import numpy as np
import pytest

from mcspeckle_amd import engines as E
from test_gpu_parity import near_spec
from test_oracle_engines import small_spec, tab3d

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def weights(n, seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(0.2, 1, n) * np.exp(2j * np.pi * rng.uniform(size=n))


WAVE_CASES = [pytest.param(*c, id="-".join(map(str, c))) for c in
              [("hg", 3, True, "far"), ("tab", 3, False, "far"), ("iso", 3, True, "far"), ("hg", 2, True, "far"),
               ("hg", 3, True, "both"), ("iso", 2, False, "l")]]


def wave_spec(kind, dim, cbs, near, nv=70, nl=5):
    if near == "far":
        return small_spec(kind, dim, cbs, nv=nv, nl=nl, seed=12, tab=tab3d() if kind == "tab" else None)
    return near_spec(kind, dim, cbs, near, nv=nv, nl=nl, seed=12)


@pytest.mark.parametrize("kind,dim,cbs,near", WAVE_CASES)
def test_gpu_mcfield_wave_matches_oracle(oracle, kind, dim, cbs, near):
    spec = wave_spec(kind, dim, cbs, near)
    wl = weights(spec.nl)
    n = 800
    say("mcfield wave", kind, dim, cbs, near)
    u0 = oracle.mcfield_wave(spec, wl, n)
    for precision, tol in (("fp64", 1e-10), ("mixed", 2e-5)):
        u = E.MCfieldOnWave(spec.sigt, spec.albedo, spec.box_min, spec.box_max, spec.l, spec.v, spec.is_ff_l,
                            spec.is_ff_v, n, spec.lam, spec.do_cbs, spec.smp_flg, spec.amp.type,
                            _ampfunc(spec), Wl=wl, seed=spec.seed, precision=precision)
        err = np.abs(u - u0).max() / np.abs(u0).max()
        say("  ", precision, "err", err)
        assert err <= tol


def _ampfunc(spec):
    if spec.amp.type == 3:
        return spec.amp.g
    if spec.amp.type == 1:
        return 0
    return tab3d()


@pytest.mark.parametrize("kind,dim,cbs,near", WAVE_CASES[:3] + WAVE_CASES[4:5])
def test_gpu_mcfield_wave_is_the_contracted_field(kind, dim, cbs, near):
    spec = wave_spec(kind, dim, cbs, near)
    wl = weights(spec.nl, seed=8)
    n = 600
    say("mcfield wave linearity", kind, dim, cbs, near)
    u = E.mcfield_spec(spec, n, "fp64")
    uw = E.MCfieldOnWave(spec.sigt, spec.albedo, spec.box_min, spec.box_max, spec.l, spec.v, spec.is_ff_l,
                         spec.is_ff_v, n, spec.lam, spec.do_cbs, spec.smp_flg, spec.amp.type, _ampfunc(spec), Wl=wl,
                         seed=spec.seed, precision="fp64")
    want = u @ wl
    assert np.abs(uw - want).max() <= 1e-12 * np.abs(want).max()


L3 = np.array([[0.05, -0.1, 0.02], [0.0, 0.08, -0.04],
               [np.sqrt(1 - 0.05 ** 2), np.sqrt(1 - 0.1 ** 2 - 0.08 ** 2), np.sqrt(1 - 0.02 ** 2 - 0.04 ** 2)]])

QUAD_CASES = [
    # cbs, v_max, v_stp, lam, box half-width, T, vsign, smp   (0.325 / 0.05: 14 views, even transform)
    (False, 0.3, 0.05, 0.5, 2.0, 0.6, -1.0, 2),
    (True, 0.325, 0.05, 0.5, 2.0, 0.6, -1.0, 2),
    (True, 0.25, 0.01, 0.5, 4.0, 0.5, 1.0, 1),
    (False, 0.325, 0.05, 0.5, 3.0, 0.8, 1.0, 2),
    # even transforms that factor (two LDS passes): 64 = 8 x 8 padded, 80 = 10 x 8 cropped
    (True, 0.315, 0.01, 0.5, 4.0, 0.6, -1.0, 2),
    (False, 0.315, 0.01, 0.5, 30.0, 0.5, 1.0, 2),
]


@pytest.mark.parametrize("cbs,vmax,vstp,lam,hw,T,vsign,smp", QUAD_CASES)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 2e-4)])
def test_gpu_quad_wave_matches_oracle(oracle, cbs, vmax, vstp, lam, hw, T, vsign, smp, precision, tol):
    bmin, bmax = [-hw, -hw, 0.0], [hw, hw, T]
    wl = weights(3, seed=5)
    n = 300
    say("quad wave", cbs, vmax, vstp, lam, hw, T, vsign, smp, precision)
    g = oracle.quad_grid(np.array(bmin), np.array(bmax), vmax, vstp, lam, wave=True)
    assert E.quad_grid(bmin, bmax, vmax, vstp, lam, wave=True) == (g.nv0, g.nx, g.nz)
    u0, K = oracle.mcfield_fourier_quad_wave(4.0, 0.97, bmin, bmax, L3, wl, vmax, vstp, n, lam, cbs, smp, None,
                                             vsign, 9)
    st = {}
    u = E.MCfieldFourierQuadOnWave(4.0, 0.97, bmin, bmax, L3, vmax, vstp, n, lam, cbs, smp, None, wl, vsign, seed=9,
                                   precision=precision, stats=st)
    assert st["vertices"] == K
    assert u.shape == u0.shape
    err = np.abs(u - u0).max() / np.abs(u0).max()
    say("  err", err, "grid", u.shape)
    assert err <= tol


@pytest.mark.parametrize("nl,cbs", [(40, True), (600, False), (600, True)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 2e-4)])
def test_gpu_quad_wave_many_lights(oracle, nl, cbs, precision, tol):
    """Waves of many plane waves: up to 512 lights the walk and the scatter
    run in one pass with the lights staged in LDS; beyond that the walks are
    stored as vertex records and the vertex-parallel scatter reads the lights
    from HBM."""
    rng = np.random.default_rng(nl)
    lx, ly = rng.uniform(-0.05, 0.05, nl), rng.uniform(-0.05, 0.05, nl)
    l = np.vstack([lx, ly, np.sqrt(1 - lx ** 2 - ly ** 2)])
    wl = weights(nl, seed=6)
    bmin, bmax = [-2.0, -2.0, 0.0], [2.0, 2.0, 0.6]
    n = 200
    say("quad wave lights", nl, cbs, precision)
    u0, K = oracle.mcfield_fourier_quad_wave(4.0, 0.97, bmin, bmax, l, wl, 0.3, 0.05, n, 0.5, cbs, 2, None, -1.0, 9)
    st = {}
    u = E.MCfieldFourierQuadOnWave(4.0, 0.97, bmin, bmax, l, 0.3, 0.05, n, 0.5, cbs, 2, None, wl, -1.0, seed=9,
                                   precision=precision, stats=st)
    assert st["vertices"] == K
    err = np.abs(u - u0).max() / np.abs(u0).max()
    say("  err", err, "grid", u.shape)
    assert err <= tol
