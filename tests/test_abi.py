"""The C-ABI library loads and exports everything include/scmc.h declares;
argument errors come back as codes with the reference's messages.  No compute
calls (runs without a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from mcspeckle_amd import _native as N
from mcspeckle_amd import engines as E


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "scmc.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(scmc_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = N.lib()
    syms = declared_symbols()
    assert len(syms) >= 15
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(N.EXPORTS)


def test_version_string():
    assert b"gfx950" in N.lib().scmc_version()


def _problem(**over):
    spec = E.make_problem(2.0, 0.9, [-1, -1, 0], [1, 1, 1], np.array([[0.0], [0.0], [1.0]]),
                          np.array([[0.0], [0.0], [1.0]]), 1, 1, 1.0, True, 2, 3, 0.5)
    p = E.c_problem(spec)
    for k, v in over.items():
        setattr(p, k, v)
    return spec, p


@pytest.mark.parametrize("field,val,msg", [
    ("dim", 4, "Invalid box size"),
    ("nl", 0, "Invalid light source input"),
    ("nv", 0, "Invalid view source input"),
    ("smp_flg", 3, "smpFlg"),
    ("sigt", -1.0, "sigt"),
])
def test_argument_errors_before_device(field, val, msg):
    spec, p = _problem(**{field: val})
    out = np.empty(4)
    rc = N.lib().scmc_mccov(C.byref(p), 10, 0, N.dptr(out), N.dptr(out), N.dptr(out), N.dptr(out), None)
    assert rc == N.SCMC_E_ARG
    assert msg in N.lib().scmc_last_error().decode()


def test_bad_iteration_count():
    spec, p = _problem()
    out = np.empty(4)
    assert N.lib().scmc_mccov(C.byref(p), 0, 0, N.dptr(out), None, None, None, None) == N.SCMC_E_ARG
    assert "Invalid photons number" in N.lib().scmc_last_error().decode()


def test_python_argument_conventions():
    # MCcov.m:48-51, 56-70
    with pytest.raises(E.ScmcArgumentError, match="Invalid box size"):
        E.make_problem(1, 1, [0, 0, 0, 0], [1, 1, 1, 1], [[0], [0], [1]], [[0], [0], [1]], 1, 1, 1, 0, 2, 1, 0)
    with pytest.raises(E.ScmcArgumentError, match="Invalid light source input"):
        E.make_problem(1, 1, [0, 0, 0], [1, 1, 1], [[0], [1]], [[0], [0], [1]], 1, 1, 1, 0, 2, 1, 0)
    with pytest.raises(E.ScmcArgumentError, match="Invalid view source input"):
        E.make_problem(1, 1, [0, 0, 0], [1, 1, 1], [[0], [0], [1]], [0.1, 0.2], 1, 1, 1, 0, 2, 1, 0)
    # 2-D angles -> [sin; cos]
    s = E.make_problem(1, 1, [0, 0], [1, 1], [0.1, 0.2], [0.3], 1, 1, 1, 0, 2, 1, 0)
    np.testing.assert_allclose(s.l, [[np.sin(0.1), np.sin(0.2)], [np.cos(0.1), np.cos(0.2)]])
    assert s.nv == 1 and s.dim == 2


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(N, "_lib", None)
    monkeypatch.setattr(N, "LIB_PATH", str(tmp_path / "nope.so"))
    with pytest.raises(ImportError, match="no CPU fallback"):
        N.lib()
