"""The ctypes mirrors of the public structs (mcspeckle_amd/_native.py) match
include/scmc.h: gcc compiles a probe that prints every struct's size and
each member's offset, and the test compares them with ctypes.  CPU only."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

from mcspeckle_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
# ctypes field names that differ from the C member (Python keywords)
ALIAS = {("scmc_problem", "lam"): "lambda"}
STRUCTS = [("scmc_phase", N.Phase), ("scmc_problem", N.Problem), ("scmc_medium", N.Medium),
           ("scmc_stats", N.Stats), ("scmc_pathset", N.PathSet), ("scmc_server_config", N.ServerConfig)]


@pytest.fixture(scope="module")
def layout(tmp_path_factory):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    d = tmp_path_factory.mktemp("abi")
    lines = ["#include <stddef.h>", "#include <stdio.h>", '#include "scmc.h"', "int main(void) {"]
    for cname, cls in STRUCTS:
        lines.append('printf("%s size %%zu\\n", sizeof(%s));' % (cname, cname))
        for f in cls._fields_:
            cm = ALIAS.get((cname, f[0]), f[0])
            lines.append('printf("%s.%s %%zu\\n", offsetof(%s, %s));' % (cname, f[0], cname, cm))
    lines += ["return 0;", "}"]
    src = d / "probe.c"
    src.write_text("\n".join(lines))
    exe = d / "probe"
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    vals = {}
    for line in out.splitlines():
        k, *rest = line.split()
        vals[k if not rest[:1] == ["size"] else k + ".__size__"] = int(rest[-1])
    return vals


@pytest.mark.parametrize("cname,cls", STRUCTS, ids=[s[0] for s in STRUCTS])
def test_ctypes_layout_matches_header(layout, cname, cls):
    assert C.sizeof(cls) == layout[cname + ".__size__"], cname
    for f in cls._fields_:
        assert getattr(cls, f[0]).offset == layout["%s.%s" % (cname, f[0])], (cname, f[0])
