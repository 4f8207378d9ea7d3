"""GPU MCfieldFourierQuad against the oracle (same Philox walks).

Tolerances: max |u - u*| / max |u*| <= 1e-10 (FP64 scatter with FP64 atomics
+ FP64 DFT GEMMs: only summation order differs) and <= 2e-4 in mixed mode
(FP32 grids, FP32 atomics and FP32 DFT GEMMs; FP64 phases).
"""
import sys

import numpy as np
import pytest

from mcspeckle_amd import engines as E

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def lights(n):
    rng = np.random.default_rng(n)
    lx, ly = rng.uniform(-0.1, 0.1, n), rng.uniform(-0.1, 0.1, n)
    return np.vstack([lx, ly, np.sqrt(1 - lx ** 2 - ly ** 2)])


CASES = [
    # cbs, v_max, v_stp, lam, box half-width, T, vsign, nl, smp
    (False, 0.3, 0.05, 0.5, 2.0, 0.6, -1.0, 2, 2),
    (True, 0.3, 0.05, 0.5, 2.0, 0.6, -1.0, 1, 2),
    (True, 0.25, 0.01, 0.5, 4.0, 0.5, 1.0, 2, 1),
    (False, 0.2, 0.004, 0.2, 5.0, 1.0, -1.0, 1, 2),
    # six lights: stored vertex records and the vertex-parallel scatter
    # (up to four, the walk and the scatter run in one pass)
    (True, 0.3, 0.05, 0.5, 2.0, 0.6, -1.0, 6, 2),
    # transforms that factor (the slab DFT in two LDS passes instead of the
    # DFT-matrix GEMMs): 63 = 9 x 7 with the cell grid padded (13 cells), and
    # 77 = 11 x 7 with the transform cropped to the 63 views
    (True, 0.31, 0.01, 0.5, 4.0, 0.6, -1.0, 2, 2),
    (False, 0.31, 0.01, 0.5, 30.0, 0.6, 1.0, 1, 2),
]


@pytest.mark.parametrize("cbs,vmax,vstp,lam,hw,T,vsign,nl,smp", CASES)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 2e-4)])
def test_gpu_quad_matches_oracle(oracle, cbs, vmax, vstp, lam, hw, T, vsign, nl, smp, precision, tol):
    l = lights(nl)
    bmin, bmax = [-hw, -hw, 0.0], [hw, hw, T]
    n = 300
    say("quad", cbs, vmax, vstp, lam, hw, T, vsign, nl, smp, precision)
    u0, K = oracle.mcfield_fourier_quad(4.0, 0.97, bmin, bmax, l, vmax, vstp, n, lam, cbs, smp, None, vsign, 9)
    st = {}
    u = E.MCfieldFourierQuad(4.0, 0.97, bmin, bmax, l, vmax, vstp, n, lam, cbs, smp, None, vsign, seed=9,
                             precision=precision, stats=st)
    assert st["vertices"] == K
    assert u.shape == u0.shape
    err = np.abs(u - u0).max() / np.abs(u0).max()
    say("  err", err, "grid", u.shape)
    assert err <= tol


PHASE_CASES = [
    # sct_type, nl (wave: None), cbs, angle bins: HG and tabulated walks, the fused
    # pass (nl <= 4), stored records (six lights), the wave variant, and a table
    # too large for LDS (read from HBM)
    (3, 2, True, 0), (2, 2, True, 256), (2, 1, False, 256), (2, 6, True, 256), (3, None, True, 0),
    (2, None, False, 256), (2, 2, True, 2001),
]


@pytest.mark.parametrize("sct_type,nl,cbs,nbins", PHASE_CASES)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-10), ("mixed", 2e-4)])
def test_gpu_quad_walk_phase_functions(oracle, sct_type, nl, cbs, nbins, precision, tol):
    """HG and tabulated walks (the Fourier benchmark walks with a tabulated
    three-lobe phase function; the reference only walks isotropically)."""
    from test_oracle_quad import tab_phase
    amp = 0.85 if sct_type == 3 else tab_phase(nbins)
    ph = E._phase_from_arg(sct_type, amp, "ampfunc")
    bmin, bmax = [-2.0, -2.0, 0.0], [2.0, 2.0, 0.6]
    n = 300
    say("quad phase", sct_type, nl, cbs, precision)
    st = {}
    if nl is None:
        l = lights(3)
        wl = np.array([0.7 * np.exp(0.3j), 0.4 * np.exp(-1.1j), 0.9 * np.exp(2.0j)])
        u0, K = oracle.mcfield_fourier_quad_wave(4.0, 0.97, bmin, bmax, l, wl, 0.3, 0.05, n, 0.5, cbs, 2, None, -1.0,
                                                 9, amp=ph)
        u = E.MCfieldFourierQuadOnWave(4.0, 0.97, bmin, bmax, l, 0.3, 0.05, n, 0.5, cbs, 2, None, wl, -1.0,
                                       sct_type=sct_type, ampfunc=amp, seed=9, precision=precision, stats=st)
    else:
        l = lights(nl)
        u0, K = oracle.mcfield_fourier_quad(4.0, 0.97, bmin, bmax, l, 0.3, 0.05, n, 0.5, cbs, 2, None, -1.0, 9,
                                            amp=ph)
        u = E.MCfieldFourierQuad(4.0, 0.97, bmin, bmax, l, 0.3, 0.05, n, 0.5, cbs, 2, None, -1.0, sct_type=sct_type,
                                 ampfunc=amp, seed=9, precision=precision, stats=st)
    assert st["vertices"] == K
    assert u.shape == u0.shape
    err = np.abs(u - u0).max() / np.abs(u0).max()
    say("  err", err, "grid", u.shape)
    assert err <= tol


def test_gpu_quad_split_and_interleaved_outputs_agree():
    """The C-ABI writes split complex (real and imaginary arrays) or, with a
    NULL imaginary pointer, interleaved complex; the Python wrapper uses the
    latter.  Both hold the same field (FP64 grid: only the atomics' order
    differs between the two runs)."""
    import ctypes as C
    from mcspeckle_amd import _native as N
    l = lights(2)
    bmin, bmax = [-2.0, -2.0, 0.0], [2.0, 2.0, 0.6]
    u = E.MCfieldFourierQuad(4.0, 0.97, bmin, bmax, l, 0.3, 0.05, 300, 0.5, True, 2, None, -1.0, seed=9,
                             precision="fp64")
    spec = E.make_problem(4.0, 0.97, bmin, bmax, l, np.zeros((3, 0)), 1, 1, 0.5, True, 2, 1, 0, None, seed=9,
                          need_views=False)
    n = u.size
    re, im = np.empty(n), np.empty(n)
    p = E.c_problem(spec)
    N.check(N.lib().scmc_mcfield_fourier_quad(C.byref(p), 0.3, 0.05, 300, -1.0, N.FP64, N.dptr(re), N.dptr(im), None))
    v = (re + 1j * im).reshape(u.shape, order="F")
    assert np.abs(v - u).max() <= 1e-12 * np.abs(u).max()
