"""Pins of the oracle's sampling/evaluation primitives against closed forms,
published known-answer vectors and invariants (CPU only)."""
import ctypes as C

import numpy as np
import pytest
from scipy import integrate, stats

from oracle import oracle as O


def _d(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a, a.ctypes.data_as(C.POINTER(C.c_double))


def test_philox_random123_kat(oracle):
    # Random123 kat_vectors, philox4x32 with 10 rounds
    assert oracle.philox([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert oracle.philox([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]
    assert oracle.philox([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == [
        0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def test_uniform_pairs_range_and_independence(oracle):
    u = np.array([oracle.uniform_pair(7, p, v, g) for p in range(200) for v in range(5) for g in range(4)])
    assert u.min() >= 0 and u.max() < 1
    flat = u.reshape(-1)
    # uniform: KS against U[0,1)
    assert stats.kstest(flat, "uniform").pvalue > 1e-3
    # distinct slots give distinct numbers
    assert np.unique(flat).size == flat.size


def test_cube_dist_slab_closed_form(oracle):
    L = oracle.lib()
    T = 2.0
    bmin, pmin = _d([-1e6, -1e6, 0.0])
    bmax, pmax = _d([1e6, 1e6, T])
    rng = np.random.default_rng(1)
    for _ in range(200):
        z = rng.uniform(0, T)
        th = rng.uniform(0.01, np.pi - 0.01)
        ph = rng.uniform(0, 2 * np.pi)
        d = np.array([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
        x, px = _d([rng.uniform(-10, 10), rng.uniform(-10, 10), z])
        dd, pd = _d(d)
        got = L.orc_cube_dist(px, pmin, pmax, pd, 3)
        want = (T - z) / d[2] if d[2] > 0 else z / -d[2]
        assert got == pytest.approx(want, rel=1e-12)


def test_phase_att_transmission_closed_form(oracle):
    """|E(x,d)| = exp(-sigma/2 * z/d_z) (enters through the bottom face) and
    arg E = k x.d for a far-field direction d with d_z > 0."""
    L = oracle.lib()
    sigma, lam, T = 3.0, 0.7, 1.5
    bmin, pmin = _d([-1e5, -1e5, 0.0])
    bmax, pmax = _d([1e5, 1e5, T])
    rng = np.random.default_rng(2)
    dirs = []
    for _ in range(50):
        th = rng.uniform(0, 1.2)
        ph = rng.uniform(0, 2 * np.pi)
        dirs.append([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])
    D = np.array(dirs)
    dd, pd = _d(D.reshape(-1))
    x, px = _d([rng.uniform(-5, 5), rng.uniform(-5, 5), rng.uniform(0, T)])
    re, pre = _d(np.zeros(len(D)))
    im, pim = _d(np.zeros(len(D)))
    L.orc_eval_phase_att(px, pd, len(D), 3, 1, 1.0, sigma, lam, pmin, pmax, pre, pim)
    e = re + 1j * im
    mag = np.exp(-sigma / 2 * x[2] / D[:, 2])
    np.testing.assert_allclose(np.abs(e), mag, rtol=1e-12)
    ph = 2 * np.pi / lam * (D @ x)
    np.testing.assert_allclose(e / np.abs(e), np.exp(1j * ph), atol=1e-9)


@pytest.mark.parametrize("dim", [3, 2])
def test_phase_att_near_field_closed_form(oracle, dim):
    """Point source p below the slab z in [0, T], x inside: the straight
    segment x -> p runs inside the medium for the fraction z/(z + h) of its
    length nd, so |E| = exp(-sigma/2 * nd z/(z + h)) / (nd + 0.01 lambda)^((dim-1)/2)
    and arg E = k nd.  A source inside the box: the whole segment is in the
    medium (|E| uses nd itself)."""
    L = oracle.lib()
    sigma, lam, T = 2.5, 0.6, 1.0
    W = 1e5
    bmin, pmin = _d([-W, 0.0] if dim == 2 else [-W, -W, 0.0])
    bmax, pmax = _d([W, T] if dim == 2 else [W, W, T])
    rng = np.random.default_rng(8)
    for inside in (False, True):
        pts = []
        for _ in range(40):
            lat = rng.uniform(-3, 3, dim - 1)
            h = rng.uniform(0.5, 4.0)
            pts.append(np.r_[lat, rng.uniform(0.05, T - 0.05) if inside else -h])
        P = np.array(pts)
        pp, ppt = _d(P.reshape(-1))
        x = np.r_[rng.uniform(-1, 1, dim - 1), rng.uniform(0.1, T - 0.1)]
        xx, px = _d(x)
        re, pre = _d(np.zeros(len(P)))
        im, pim = _d(np.zeros(len(P)))
        L.orc_eval_phase_att(px, ppt, len(P), dim, 0, 1.0, sigma, lam, pmin, pmax, pre, pim)
        e = re + 1j * im
        nd = np.linalg.norm(x - P, axis=1)
        zf = x[-1]
        path_in = nd if inside else nd * zf / (zf - P[:, -1])
        mag = np.exp(-sigma / 2 * path_in) / (nd + 0.01 * lam) ** ((dim - 1) / 2)
        np.testing.assert_allclose(np.abs(e), mag, rtol=1e-12)
        np.testing.assert_allclose(e / np.abs(e), np.exp(2j * np.pi / lam * nd), atol=1e-9)


@pytest.mark.parametrize("g", [-0.5, 0.0, 0.3, 0.9, 0.95])
def test_hg_normalisation(oracle, g):
    L = oracle.lib()
    # 3-D: 2 pi int_{-1}^{1} HG(c) dc = 1
    val, _ = integrate.quad(lambda c: L.orc_evaluate_hg(c, g, 3), -1, 1, limit=200, points=[1.0])
    assert 2 * np.pi * val == pytest.approx(1.0, rel=1e-8)
    # 2-D: int_0^{2 pi} HG2(cos t) dt = 1
    val2, _ = integrate.quad(lambda t: L.orc_evaluate_hg(np.cos(t), g, 2), 0, 2 * np.pi, limit=400)
    assert val2 == pytest.approx(1.0, rel=1e-8)


def test_isotropic_amplitude_normalisation(oracle):
    L = oracle.lib()
    ph = O.OPhase()
    ph.type = 1
    for dim, meas in ((3, 4 * np.pi), (2, 2 * np.pi)):
        re, im = C.c_double(), C.c_double()
        L.orc_eval_amp(0.3, C.byref(ph), dim, C.byref(re), C.byref(im))
        assert re.value ** 2 * meas == pytest.approx(1.0, rel=1e-14)
        assert im.value == 0


@pytest.mark.parametrize("g", [0.0005, 0.2, 0.8, 0.9, -0.4])
def test_hg_sampler_mean_and_distribution(oracle, g):
    L = oracle.lib()
    u = (np.arange(40000) + 0.5) / 40000   # stratified uniforms
    c = np.array([L.orc_sample_hg(g, 3, x) for x in u])
    assert np.all(np.abs(c) <= 1 + 1e-12)
    geff = 0.0 if abs(g) < 0.001 else g
    assert c.mean() == pytest.approx(geff, abs=2e-4)
    # CDF of HG3 in cos: F(c) = (1-g^2)/(2g) [1/sqrt(1+g^2-2gc) - 1/(1+g)]
    if abs(g) >= 0.001:
        cdf = lambda x: (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * x) - 1 / (1 + g))
        assert stats.kstest(c, cdf).pvalue > 1e-3
    # 2-D (flatland HG): E[cos theta] = g
    c2 = np.array([L.orc_sample_hg(g, 2, x) for x in u])
    assert c2.mean() == pytest.approx(geff, abs=2e-3)


def test_rotate_by_theta_invariants(oracle):
    L = oracle.lib()
    rng = np.random.default_rng(3)
    for k in range(300):
        if k < 20:
            ow = np.array([0.0, 0.0, 1.0 if k % 2 else -1.0])   # pole branch
        else:
            ow = rng.normal(size=3)
            ow /= np.linalg.norm(ow)
        c = rng.uniform(-1, 1)
        o, po = _d(ow)
        w, pw = _d(np.zeros(3))
        L.orc_rotate_by_theta(po, c, 3, rng.uniform(), pw)
        assert np.linalg.norm(w) == pytest.approx(1.0, abs=1e-12)
        assert w @ ow == pytest.approx(c, abs=1e-12)
        # 2-D
        o2 = ow[:2] / np.linalg.norm(ow[:2]) if k >= 20 else np.array([0.0, 1.0])
        o, po = _d(o2)
        w2, pw2 = _d(np.zeros(2))
        L.orc_rotate_by_theta(po, c, 2, rng.uniform(), pw2)
        assert np.linalg.norm(w2) == pytest.approx(1.0, abs=1e-12)
        assert w2 @ o2 == pytest.approx(c, abs=1e-12)


@pytest.mark.parametrize("lz", [1.0, 0.8, -0.9])
def test_exp_smp_x_is_a_density_ratio(oracle, lz):
    """px = p(z) * dmax for the truncated exponential p, so E_p[1/px] = 1 and
    E_p[z/px] = (zmin+zmax)/2 (the uniform mean)."""
    L = oracle.lib()
    bmin, pmin = _d([-2.0, -3.0, 0.5])
    bmax, pmax = _d([2.0, 3.0, 2.5])
    lm, plm = _d([np.sqrt(1 - lz * lz), 0.0, lz])
    n = 20000
    inv, zs = np.empty(n), np.empty(n)
    rng = np.random.default_rng(4)
    for i in range(n):
        u, pu = _d(rng.uniform(size=3))
        x, px_ = _d(np.zeros(3))
        p = C.c_double()
        L.orc_exp_smp_x(pmin, pmax, plm, 4.0, 3, pu, px_, C.byref(p))
        assert bmin[0] <= x[0] <= bmax[0] and bmin[1] <= x[1] <= bmax[1] and bmin[2] <= x[2] <= bmax[2]
        inv[i] = 1 / p.value
        zs[i] = x[2] / p.value
    se = inv.std() / np.sqrt(n)
    assert abs(inv.mean() - 1) < 4 * se
    se2 = zs.std() / np.sqrt(n)
    assert abs(zs.mean() - 1.5) < 4 * se2


@pytest.mark.parametrize("dim", [3, 2])
def test_isotropic_directions_are_uniform(oracle, dim):
    """smpampfunc_general.m:7-8 / MCcov.m:148: randn(dim,1) normalised is uniform
    on the sphere (circle): unit length, zero mean, second moment I/dim, and in
    3-D every coordinate uniform on [-1, 1] (Archimedes), in 2-D the angle
    uniform on [0, 2 pi).  The oracle's Box-Muller pairs feed it."""
    L = oracle.lib()
    ph = O.OPhase()
    ph.type = 1
    ow = np.ascontiguousarray([0.0, 0.0, 1.0][:dim])
    rng = np.random.default_rng(dim)
    n = 20000
    W = np.empty((n, dim))
    w = np.zeros(dim)
    for i in range(n):
        u = np.ascontiguousarray(rng.uniform(size=4))
        L.orc_sample_dir(O._p(ow), C.byref(ph), dim, O._p(u), O._p(w))
        W[i] = w
    assert np.allclose(np.linalg.norm(W, axis=1), 1, atol=1e-12)
    assert np.abs(W.mean(axis=0)).max() < 4.5 / np.sqrt(n)
    assert np.abs(W.T @ W / n - np.eye(dim) / dim).max() < 5 / np.sqrt(n)
    if dim == 3:
        for j in range(3):
            assert stats.kstest(W[:, j], stats.uniform(-1, 2).cdf).pvalue > 1e-3
    else:
        ang = np.mod(np.arctan2(W[:, 1], W[:, 0]), 2 * np.pi)
        assert stats.kstest(ang, stats.uniform(0, 2 * np.pi).cdf).pvalue > 1e-3
