"""The measured workloads, defined once for the benchmark and the tests.

Each builder returns the inputs of one engine call with fixed seeds, so the
benchmark times and the GPU tests check exactly the same problems:

    plain_cov_64        MCcov, 64 x 64 covariance (8 lights x 8 views)
    speckle_image_128   MCfield, 128 x 128 far-field views, one light
    cov_4096            MCcov, 4096 x 4096 covariance (64 lights x 64 views)
    fourier_field_511   MCfieldFourierQuad, 511 x 511 grid, tabulated phase
    request_stream      covariance requests for the serving loop

The far-field conventions follow the reference's examples:
renderSpecklesImages.m:34-41 (view grid of normalised [sin tx; sin ty; 1]),
test3DFourierSmpQuad.m:30-60 (Fourier grid arguments).
"""
import numpy as np

from . import engines as E
from .interface import tabulatedAmplitudeScatter


def cone_dirs(rng, n, deg):
    """n directions uniform in solid angle on a cone of half-angle deg about +z."""
    ct = rng.uniform(np.cos(np.deg2rad(deg)), 1.0, n)
    st = np.sqrt(1 - ct ** 2)
    ph = rng.uniform(0, 2 * np.pi, n)
    return np.vstack([st * np.cos(ph), st * np.sin(ph), ct])


def polar_dirs(rng, n, max_deg):
    """n directions with polar angle uniform in [0, max_deg] about +z."""
    th = np.deg2rad(rng.uniform(0, max_deg, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    return np.vstack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)])


def plain_cov_64(seed=0, seed_dirs=0):
    """Slab of thickness 1 (x, y in +-5e4: V = prod(box size) enters the
    normalisation, MCcov.m:297-299), sigma_s = 5, sigma_a = 0.1, isotropic,
    lambda = 1, 8 lights and 8 views with polar angles uniform in [0, 20] deg,
    CBS on, exponential first-scatter sampling: a 64 x 64 covariance."""
    rng = np.random.default_rng(seed_dirs)
    l, v = polar_dirs(rng, 8, 20), polar_dirs(rng, 8, 20)
    sigt = 5.1
    return E.make_problem(sigt, 5 / sigt, [-5e4, -5e4, 0], [5e4, 5e4, 1], l, v, 1, 1, 1.0, True, 2, 1, 0, seed=seed)


def view_grid(n, deg):
    """n x n far-field views, normalised [sin tx; sin ty; 1] on a uniform
    angle grid over +-deg (renderSpecklesImages.m:34-41)."""
    th = np.deg2rad(np.linspace(-deg, deg, n))
    X, Y = np.meshgrid(th, th)
    v = np.vstack([np.sin(X.ravel()), np.sin(Y.ravel()), np.ones(X.size)])
    return v / np.linalg.norm(v, axis=0)


def speckle_image_128(seed=0):
    """Slab of thickness 2 (x, y in +-5e4), sigma_s = 10, sigma_a = 0.05, HG
    g = 0.8 (isotropic first direction), one plane wave at normal incidence,
    128 x 128 views over +-10 deg, CBS off."""
    sigt = 10.05
    return E.make_problem(sigt, 10 / sigt, [-5e4, -5e4, 0], [5e4, 5e4, 2], [[0], [0], [1]], view_grid(128, 10), 1, 1,
                          1.0, False, 2, 3, 0.8, seed=seed)


def cov_4096(seed=0, seed_dirs=2024):
    """Slab of thickness 1 (x, y in +-5e4), sigma = 20, albedo 1, HG g = 0.9,
    64 lights and 64 views uniform in solid angle on a 15 deg cone, CBS on,
    lambda = 1: a 4096 x 4096 covariance."""
    rng = np.random.default_rng(seed_dirs)
    l, v = cone_dirs(rng, 64, 15), cone_dirs(rng, 64, 15)
    return E.make_problem(20.0, 1.0, [-5e4, -5e4, 0], [5e4, 5e4, 1], l, v, 1, 1, 1.0, True, 2, 3, 0.9, seed=seed)


def three_lobe_table(n_bins=256, seed=3):
    """Tabulated phase function: sqrt of a mixture of three HG lobes, g_i
    uniform in [-0.3, 0.95], weights Dirichlet(1, 1, 1), on n_bins angles in
    [0, pi], built by tabulatedAmplitudeScatter.  Returns (ampfunc, g, w)."""
    rng = np.random.default_rng(seed)
    gs, ws = rng.uniform(-0.3, 0.95, 3), rng.dirichlet([1, 1, 1])
    th = np.linspace(0, np.pi, n_bins)
    th[-1] = np.pi
    hg = sum(w * (1 - g * g) / (4 * np.pi * (1 + g * g - 2 * g * np.cos(th)) ** 1.5) for g, w in zip(gs, ws))
    return tabulatedAmplitudeScatter(th, np.sqrt(hg))["ampfunc"], gs, ws


def fourier_field_511(seed=0):
    """Keyword arguments of engines.MCfieldFourierQuad: box 10 x 10 x 1,
    sigma_s = 8, sigma_a = 0.02, the 256-bin three-lobe table, one light at
    normal incidence, CBS off, a 511 x 511 grid (v_max = 0.25, 511 steps; the
    reference's grids are odd), lambda such that the cell grid has the same
    size."""
    tab, _, _ = three_lobe_table()
    v_max, v_stp = 0.25, 0.5 / 510
    sigt = 8.02
    return dict(sigt=sigt, albedo=8 / sigt, box_min=[-5, -5, 0], box_max=[5, 5, 1], l=[[0], [0], [1]], v_max=v_max,
                v_stp=v_stp, lambda_=10 * (2 * v_max + v_stp) / 511, doCBS=False, smpFlg=2, lmean0=None, vsign=-1,
                sct_type=2, ampfunc=tab, seed=seed)


def request_stream(n_req, rate, path_scale, seed=5):
    """Covariance requests: Nv, Nl uniform in {4..64}, optical depth uniform
    in [1, 20] (slab T = 1, so sigma = OD, albedo 1), HG g uniform in
    [0, 0.95], paths log-uniform in [1e5, 1e7] times path_scale, directions on
    a 15 deg cone, Poisson arrivals at `rate` per second.  Returns a list of
    (arrival time s, ProblemSpec, paths)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(n_req):
        nv, nl = int(rng.integers(4, 65)), int(rng.integers(4, 65))
        od, g = rng.uniform(1, 20), rng.uniform(0, 0.95)
        paths = max(1, int(round(np.exp(rng.uniform(np.log(1e5), np.log(1e7))) * path_scale)))
        l, v = cone_dirs(rng, nl, 15), cone_dirs(rng, nv, 15)
        spec = E.make_problem(od, 1.0, [-5e4, -5e4, 0], [5e4, 5e4, 1], l, v, 1, 1, 1.0, True, 2, 3, g,
                              seed=1000 + i)
        out.append((t, spec, paths))
        t += rng.exponential(1.0 / rate)
    return out
