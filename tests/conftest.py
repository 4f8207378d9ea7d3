import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP kernels)")


def far_dirs(n, max_deg, rng, sign=1.0):
    """n unit directions with polar angle uniform in [0, max_deg] (+z or -z)."""
    th = np.deg2rad(rng.uniform(0, max_deg, n))
    ph = rng.uniform(0, 2 * np.pi, n)
    return np.vstack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), sign * np.cos(th)])


def cone_dirs(n, max_deg, rng):
    """n directions uniform in solid angle on a cone of half-angle max_deg."""
    ct = rng.uniform(np.cos(np.deg2rad(max_deg)), 1.0, n)
    st = np.sqrt(1 - ct ** 2)
    ph = rng.uniform(0, 2 * np.pi, n)
    return np.vstack([st * np.cos(ph), st * np.sin(ph), ct])


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.lib()
    return O


@pytest.fixture(scope="session")
def amp_fixture():
    d = np.load(os.path.join(ROOT, "tests", "golden", "scatteringAmplitude.npz"))
    return d["theta"], d["farField_re"] + 1j * d["farField_im"]


def cov_matrix(M):
    """nv x nv x nl x nl (MATLAB layout) -> N x N matrix with row (v1,l1)."""
    nv, _, nl, _ = M.shape
    return np.transpose(M, (0, 2, 1, 3)).reshape(nv * nl, nv * nl, order="F")


def cs_scale(Cref):
    """Cauchy-Schwarz scale sqrt(C_ii C_jj) of a PSD covariance."""
    d = np.sqrt(np.abs(np.real(np.diag(Cref))))
    return np.outer(d, d)
