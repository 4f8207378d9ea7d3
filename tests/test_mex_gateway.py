"""MEX gateways (MCcov/MCfield/MCfieldFourierQuad and the illumination-wave
variants MCfieldOnWave/MCfieldFourierQuadOnWave) under a fake libmx.

CPU: the argument contract of MCcov.m:45-70 / MCfield.m:41-66 (narginchk,
'Invalid box size', 'Invalid light source input', 'Invalid view source
input') raises with the reference's messages before any device work.
GPU: a gateway call returns exactly what the engine returns for the Philox key
the gateway draws from MATLAB's stream (rng(seed) fixes it).
"""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "mex"))
from harness import Mex  # noqa: E402

from mcspeckle_amd import engines as E  # noqa: E402


@pytest.fixture(scope="module")
def mex():
    return Mex()


def base_args():
    l = np.array([[0.0, 0.1], [0.0, 0.0], [1.0, np.sqrt(1 - 0.01)]])
    v = np.array([[0.05, -0.05, 0.0], [0.0, 0.02, 0.1], [0.9987, 0.9985, 0.995]])
    return [2.0, 0.9, np.array([[-10.0], [-10.0], [0.0]]), np.array([[10.0], [10.0], [1.0]]), l, v, 1, 1, 300, 1.0,
            1, 2, 3, 0.7]


def test_narginchk(mex):
    _, err = mex.call("MCcov", 2, *base_args()[:13])
    assert err == ("MATLAB:narginchk:notEnoughInputs", "Not enough input arguments.")
    _, err = mex.call("MCcov", 2, *(base_args() + [0.5, 0.5]))
    assert err[0] == "MATLAB:narginchk:tooManyInputs"
    _, err = mex.call("MCfield", 1, *(base_args() + [0.5, np.array([[0], [0], [1.0]]), 7]))
    assert err[0] == "MATLAB:narginchk:tooManyInputs"


@pytest.mark.parametrize("gw", ["MCcov", "MCfield"])
def test_reference_argument_errors(mex, gw):
    a = base_args()
    a[3] = np.array([[10.0, 1.0], [10.0, 1.0]])
    assert mex.call(gw, 1, *a)[1][1] == "Invalid box size"
    a = base_args()
    a[4] = np.zeros((2, 2))
    assert mex.call(gw, 1, *a)[1][1] == "Invalid light source input"
    a = base_args()
    a[5] = np.zeros((4, 1))
    assert mex.call(gw, 1, *a)[1][1] == "Invalid view source input"


def test_lmean0_is_checked_before_it_is_read(mex):
    """lmean0 must be a real double vector of dim (MCfield, MCfieldOnWave) or
    3 (the Fourier gateways) elements; a short, complex or non-double value
    raises instead of being read past its end."""
    a = base_args()
    bad = [np.array([[0.0], [1.0]]), np.array([[0.0], [0.0], [1.0], [0.0]]), np.array([[0.0], [0.0], [1.0 + 1j]]),
           {"x": 1.0}]
    for lm in bad:
        _, err = mex.call("MCfield", 1, *(a + [[], lm]))
        assert err is not None and "lmean0" in err[1], (lm, err)
        _, err = mex.call("MCfieldOnWave", 1, *(a + [[], lm, np.array([[1.0, 0.5]])]))
        assert err is not None and "lmean0" in err[1], (lm, err)
    q = [2.0, 0.9, np.array([[-5.0], [-5.0], [0.0]]), np.array([[5.0], [5.0], [1.0]]), np.array([[0.0], [0.0], [1.0]]),
         0.1, 0.01, 100, 1.0, 0, 2]
    for lm in bad:
        _, err = mex.call("MCfieldFourierQuad", 1, *(q + [lm]))
        assert err is not None and "lmean0" in err[1], (lm, err)
        _, err = mex.call("MCfieldFourierQuadOnWave", 1, *(q + [lm, np.array([[1.0]])]))
        assert err is not None and "lmean0" in err[1], (lm, err)


@pytest.mark.gpu
def test_mccov_gateway_equals_engine(mex):
    a = base_args()
    key = mex.seed_for_next_call(77)
    (Ms, Mm), err = mex.call("MCcov", 2, *a)
    assert err is None, err
    spec = E.make_problem(*a[:8], a[9], a[10], a[11], a[12], a[13], seed=key)
    Ms0, Mm0 = E.mccov_spec(spec, a[8], "fp64")
    assert Ms.shape == (3, 3, 2, 2)
    np.testing.assert_array_equal(Ms, Ms0)
    np.testing.assert_array_equal(Mm, Mm0)


@pytest.mark.gpu
def test_mcfield_gateway_2d_angles_equal_engine(mex):
    # 2-D problem with 1 x N angle rows (MCfield.m:52-66)
    a = [2.0, 0.9, np.array([[-10.0], [0.0]]), np.array([[10.0], [1.0]]), np.array([[0.0, 0.2]]),
         np.array([[0.1, -0.1, 0.05, 0.0]]), 1, 1, 300, 1.0, 0, 2, 3, 0.5]
    key = mex.seed_for_next_call(5)
    (u,), err = mex.call("MCfield", 1, *a)
    assert err is None, err
    spec = E.make_problem(*a[:8], a[9], a[10], a[11], a[12], a[13], seed=key)
    u0 = E.mcfield_spec(spec, 300, "fp64")
    np.testing.assert_array_equal(u, u0)


@pytest.mark.gpu
def test_quad_gateway_runs(mex):
    l = np.array([[0.0], [0.0], [1.0]])
    key = mex.seed_for_next_call(9)
    (u,), err = mex.call("MCfieldFourierQuad", 1, 4.0, 0.97, np.array([[-2.0], [-2.0], [0.0]]),
                         np.array([[2.0], [2.0], [0.6]]), l, 0.3, 0.05, 200, 0.5, 0, 2, np.zeros((0, 0)), -1.0)
    assert err is None, err
    u0 = E.MCfieldFourierQuad(4.0, 0.97, [-2, -2, 0], [2, 2, 0.6], l, 0.3, 0.05, 200, 0.5, 0, 2, None, -1.0,
                              seed=key, precision="fp64")
    # the scatter uses FP64 atomics, so only the summation order differs
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-12 * np.abs(u0).max())


WL = np.array([0.8 * np.exp(0.4j), 0.5 * np.exp(-2.0j)])


def test_wave_gateway_argument_contract(mex):
    a = base_args() + [np.zeros((0, 0)), np.zeros((0, 0))]
    _, err = mex.call("MCfieldOnWave", 1, *a)
    assert err[0] == "MATLAB:narginchk:notEnoughInputs"
    _, err = mex.call("MCfieldOnWave", 1, *(a + [WL[:1]]))
    assert err == ("scmc:MCfieldOnWave", "Wl must have one weight per light")
    b = a + [WL]
    b[4] = np.zeros((2, 2))
    assert mex.call("MCfieldOnWave", 1, *b)[1][1] == "Invalid light source input"
    q = [4.0, 0.97, np.array([[-2.0], [-2.0], [0.0]]), np.array([[2.0], [2.0], [0.6]]), np.eye(3)[:, :2], 0.3, 0.05,
         100, 0.5, 0, 2, np.zeros((0, 0))]
    _, err = mex.call("MCfieldFourierQuadOnWave", 1, *q)
    assert err[0] == "MATLAB:narginchk:notEnoughInputs"
    _, err = mex.call("MCfieldFourierQuadOnWave", 1, *(q + [WL, -1.0, 7]))
    assert err[0] == "MATLAB:narginchk:tooManyInputs"


@pytest.mark.gpu
def test_mcfield_wave_gateway_equals_engine(mex):
    a = base_args() + [np.zeros((0, 0)), np.zeros((0, 0)), WL.reshape(1, -1)]
    key = mex.seed_for_next_call(21)
    (u,), err = mex.call("MCfieldOnWave", 1, *a)
    assert err is None, err
    u0 = E.MCfieldOnWave(*a[:14], None, None, WL, seed=key, precision="fp64")
    assert u.shape == (3, 1)
    np.testing.assert_array_equal(u[:, 0], u0)


@pytest.mark.gpu
def test_quad_wave_gateway_equals_engine(mex):
    l = np.array([[0.0, 0.05], [0.0, -0.03], [1.0, np.sqrt(1 - 0.05 ** 2 - 0.03 ** 2)]])
    key = mex.seed_for_next_call(13)
    (u,), err = mex.call("MCfieldFourierQuadOnWave", 1, 4.0, 0.97, np.array([[-2.0], [-2.0], [0.0]]),
                         np.array([[2.0], [2.0], [0.6]]), l, 0.325, 0.05, 200, 0.5, 1, 2, np.zeros((0, 0)),
                         WL.reshape(-1, 1), 1.0)
    assert err is None, err
    u0 = E.MCfieldFourierQuadOnWave(4.0, 0.97, [-2, -2, 0], [2, 2, 0.6], l, 0.325, 0.05, 200, 0.5, 1, 2, None, WL,
                                    1.0, seed=key, precision="fp64")
    assert u.shape == (14, 14)
    np.testing.assert_allclose(u, u0, rtol=0, atol=1e-12 * np.abs(u0).max())


def hetro_args():
    grid = np.random.default_rng(1).uniform(0.5, 3.0, (2, 3, 2))
    l = np.array([[0.0, 0.1], [0.0, 0.0], [1.0, np.sqrt(1 - 0.01)]])
    v = np.array([[0.05, -0.05, 0.0], [0.0, 0.02, 0.1], [0.9987, 0.9985, 0.995]])
    return [grid, 0.9, np.array([[-2.0], [-3.0], [0.0]]), np.array([[2.0], [3.0], [1.0]]),
            np.array([[2.0], [2.0], [0.5]]), l, v, 1, 1, 300, 1.0, 1, 2, 3, 0.7]


def test_hetro_gateway_argument_contract(mex):
    a = hetro_args()
    a[0] = a[0].reshape(2, 6, order="F")
    _, err = mex.call("MCCovHetro", 3, *a[:14])
    assert err[0] == "MATLAB:narginchk:notEnoughInputs"
    _, err = mex.call("MCCovHetro", 3, *(a + [np.zeros((0, 0)), 1.0, 5]))
    assert err[0] == "MATLAB:narginchk:tooManyInputs"
    b = list(a)
    b[0] = np.ones((2, 2))
    assert mex.call("MCCovHetro", 3, *b)[1] == ("scmc:MCCovHetro", "sigt does not match the cell grid")


@pytest.mark.gpu
def test_hetro_gateway_equals_engine(mex):
    a = hetro_args()
    g = list(a)
    g[0] = a[0].reshape(2, 6, order="F")   # the fake libmx holds 2-D arrays; numel and order are kept
    key = mex.seed_for_next_call(31)
    out, err = mex.call("MCCovHetro", 3, *g)
    assert err is None, err
    Ms, Mm, m1 = out
    Ms0, Mm0, m10 = E.MCCovHetro(*a[:15], seed=key, precision="fp64")
    assert Ms.shape == (3, 3, 2, 2) and m1.shape == (3, 2)
    np.testing.assert_array_equal(Ms, Ms0)
    np.testing.assert_array_equal(Mm, Mm0)
    np.testing.assert_array_equal(m1, m10)


def test_sample_hetro_gateway_argument_contract(mex):
    a = hetro_args()
    a[0] = a[0].reshape(2, 6, order="F")
    _, err = mex.call("MCsampleHetro", 1, *a[:14])
    assert err[0] == "MATLAB:narginchk:notEnoughInputs"
    _, err = mex.call("MCsampleHetro", 1, *(a + [np.zeros((0, 0)), np.zeros((0, 0)), np.zeros((0, 0)), 5]))
    assert err[0] == "MATLAB:narginchk:tooManyInputs"
    _, err = mex.call("MCsampleHetro", 2, *a)
    assert err == ("scmc:MCsampleHetro", "dbgpair / corrM are not provided by this build")
    b = list(a)
    b[0] = np.ones((2, 2))
    assert mex.call("MCsampleHetro", 1, *b)[1] == ("scmc:MCsampleHetro", "sigt does not match the cell grid")


@pytest.mark.gpu
def test_sample_hetro_gateway_equals_engine(mex):
    a = hetro_args()
    g = list(a)
    g[0] = a[0].reshape(2, 6, order="F")   # the fake libmx holds 2-D arrays; numel and order are kept
    lm0 = np.array([[0.05], [0.0], [1.0]])
    key = mex.seed_for_next_call(37)
    out, err = mex.call("MCsampleHetro", 1, *(g + [np.zeros((0, 0)), lm0]))
    assert err is None, err
    (u,) = out
    u0 = E.MCsampleHetro(*a[:15], None, lm0.ravel(), seed=key, precision="fp64")
    assert u.shape == (3, 2)
    np.testing.assert_array_equal(u, u0)
