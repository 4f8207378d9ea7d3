"""GPU parity: the HIP engines (through the C-ABI) against the FP64 oracle.

Tolerances (error metric for covariances: max_ij |C_ij - C*_ij| /
sqrt(C*_ii C*_jj), the Cauchy-Schwarz scale, so near-zero off-diagonal
entries do not blow up the relative error):
  * replayed path sets, FP64 GPU path ............ <= 1e-12
  * replayed path sets, mixed (FP16 hi/lo MFMA) .. <= 1e-4
  * full runs, same seed, FP64 .................... <= 1e-10 (the GPU tracer
    reproduces the oracle's walks; only transcendental ulps differ)
  * fields: max |u - u*| / max |u*| with the same bounds.
"""
import sys

import numpy as np
import pytest

from conftest import cone_dirs, cov_matrix, cs_scale, far_dirs
from mcspeckle_amd import engines as E
from test_oracle_engines import small_spec, tab3d

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cs_err(M, Mref):
    C, R = cov_matrix(M), cov_matrix(Mref)
    s = cs_scale(R)
    s = np.maximum(s, 1e-300)
    return float(np.max(np.abs(C - R) / s))


def field_err(u, uref):
    return float(np.abs(u - uref).max() / np.abs(uref).max())


def spec_cases():
    out = []
    for kind, dim, cbs, smp, amp0 in [("iso", 3, True, 2, None), ("hg", 3, False, 2, None), ("hg", 3, True, 1, 0.4),
                                      ("tab", 3, True, 2, None), ("hg", 2, True, 2, None), ("iso", 2, False, 1, None)]:
        out.append(pytest.param(kind, dim, cbs, smp, amp0, id=f"{kind}-{dim}d-cbs{int(cbs)}-smp{smp}-a0{amp0}"))
    return out


def build(kind, dim, cbs, smp, amp0, nv=5, nl=3, seed=11):
    tab = tab3d() if kind == "tab" else None
    return small_spec(kind, dim, cbs, nv=nv, nl=nl, seed=seed, smp=smp, amp0=amp0, tab=tab)


@pytest.mark.parametrize("kind,dim,cbs,smp,amp0", spec_cases())
def test_gpu_trace_reproduces_oracle_walks(oracle, kind, dim, cbs, smp, amp0):
    spec = build(kind, dim, cbs, smp, amp0)
    say("trace", kind, dim, cbs, smp, amp0)
    n = 3000
    g = E.trace(spec, n)
    o = oracle.trace(spec, n)
    same = np.diff(g["offset"]) == np.diff(o["offset"])
    assert same.mean() > 0.999
    assert np.array_equal(g["offset"], o["offset"])
    np.testing.assert_allclose(g["x"], o["x"], rtol=0, atol=1e-10)
    np.testing.assert_allclose(g["ow"], o["ow"], rtol=0, atol=1e-12)
    np.testing.assert_array_equal(g["weight"], o["weight"])
    np.testing.assert_array_equal(g["xi"], o["xi"])
    np.testing.assert_allclose(g["px"], o["px"], rtol=1e-12)
    np.testing.assert_allclose(g["w0p_re"], o["w0p_re"], rtol=1e-12)


@pytest.mark.parametrize("n", [1, 2047, 2048, 2049, 6145, 120001, 700001])
def test_gpu_trace_offsets_across_scan_tiles(oracle, n):
    """The vertex offsets (a tiled exclusive scan of the per-path counts,
    2048 counts per tile) at path counts on and beside the tile edges, and
    the walks themselves at a chunk size of the Fourier field's order."""
    spec = build("hg", 3, True, 2, None)
    say("trace offsets", n)
    g = E.trace(spec, n)
    o = oracle.trace(spec, n)
    assert g["offset"][0] == 0
    assert np.array_equal(g["offset"], o["offset"])
    if n > 100000:
        np.testing.assert_array_equal(g["weight"], o["weight"])
        np.testing.assert_array_equal(g["xi"], o["xi"])
        np.testing.assert_allclose(g["x"], o["x"], rtol=0, atol=1e-10)
        np.testing.assert_allclose(g["px"], o["px"], rtol=1e-12)


@pytest.mark.parametrize("kind,dim,cbs,smp,amp0", spec_cases())
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-4)])
def test_gpu_mccov_replay(oracle, kind, dim, cbs, smp, amp0, precision, tol):
    spec = build(kind, dim, cbs, smp, amp0)
    say("mccov replay", kind, dim, cbs, smp, amp0, precision)
    n = 400
    ps = oracle.trace(spec, n)
    Ms0, Mm0 = oracle.mccov_replay(spec, ps, n)
    Ms, Mm = E.mccov_replay(spec, ps, n, precision)
    assert cs_err(Ms, Ms0) <= tol
    assert cs_err(Mm, Mm0) <= tol
    assert cs_err(Ms + Mm, Ms0 + Mm0) <= tol


@pytest.mark.parametrize("nv,nl", [(1, 1), (7, 5), (64, 3), (130, 1), (20, 13)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-4)])
def test_gpu_mccov_tile_edges(oracle, nv, nl, precision, tol):
    """Sizes around the 64/128 tile edges, incl. N=1 and N not a multiple."""
    rng = np.random.default_rng(nv * 100 + nl)
    l, v = far_dirs(nl, 15, rng), far_dirs(nv, 15, rng)
    spec = E.make_problem(6.0, 0.95, [-50, -50, 0], [50, 50, 1], l, v, 1, 1, 1.0, True, 2, 3, 0.8, seed=5)
    say("tile edges", nv, nl, precision)
    n = 60
    ps = oracle.trace(spec, n)
    Ms0, Mm0 = oracle.mccov_replay(spec, ps, n)
    Ms, Mm = E.mccov_replay(spec, ps, n, precision)
    assert cs_err(Ms, Ms0) <= tol
    assert cs_err(Mm, Mm0) <= tol
    # exactly Hermitian with a real diagonal (diagonal tiles included)
    for M in (Ms, Mm):
        C = cov_matrix(M)
        assert np.array_equal(C, C.conj().T)


@pytest.mark.parametrize("nv,nl", [(8, 8), (9, 13)])
def test_gpu_mccov_fp64_several_ksplits(oracle, nv, nl):
    """Enough vertices (600 paths at optical depth 12) that the FP64 update
    of one chunk runs as several k-splits, each a contiguous run of 8-vertex
    stages of the expanded rows, the last one padded with zeros: equal to
    the oracle to 1e-12."""
    rng = np.random.default_rng(nv * 31 + nl)
    l, v = far_dirs(nl, 15, rng), far_dirs(nv, 15, rng)
    spec = E.make_problem(12.0, 0.97, [-50, -50, 0], [50, 50, 1], l, v, 1, 1, 1.0, True, 2, 3, 0.7, seed=11)
    n = 600
    ps = oracle.trace(spec, n)
    nvert = int(ps["offset"][-1])
    say("fp64 k-splits", nv, nl, "vertices", nvert)
    assert nvert >= 2 * 2048   # at least two k-splits of >= 2048 vertices
    Ms0, Mm0 = oracle.mccov_replay(spec, ps, n)
    Ms, Mm = E.mccov_replay(spec, ps, n, "fp64")
    assert cs_err(Ms, Ms0) <= 1e-12
    assert cs_err(Mm, Mm0) <= 1e-12


@pytest.mark.parametrize("kind,dim,cbs,smp,amp0", spec_cases())
def test_gpu_mccov_full_run_same_seed(oracle, kind, dim, cbs, smp, amp0):
    spec = build(kind, dim, cbs, smp, amp0)
    say("mccov full", kind, dim, cbs, smp, amp0)
    n = 2000
    Ms0, Mm0 = oracle.mccov(spec, n)
    Ms, Mm = E.mccov_spec(spec, n, "fp64")
    assert cs_err(Ms, Ms0) <= 1e-10
    assert cs_err(Mm, Mm0) <= 1e-10
    Ms1, Mm1 = E.mccov_spec(spec, n, "mixed")
    assert cs_err(Ms1 + Mm1, Ms0 + Mm0) <= 1e-4


@pytest.mark.parametrize("kind,dim,cbs,smp,amp0", spec_cases())
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 2e-5)])
def test_gpu_mcfield_replay(oracle, kind, dim, cbs, smp, amp0, precision, tol):
    spec = build(kind, dim, cbs, smp, amp0, nv=37, nl=3)
    say("mcfield replay", kind, dim, cbs, smp, amp0, precision)
    n = 300
    ps = oracle.trace(spec, n)
    u0 = oracle.mcfield_replay(spec, ps, n)
    u = E.mcfield_replay(spec, ps, n, precision)
    assert field_err(u, u0) <= tol


@pytest.mark.parametrize("kind,dim,cbs,smp,amp0", spec_cases())
def test_gpu_mcfield_full_run_same_seed(oracle, kind, dim, cbs, smp, amp0):
    spec = build(kind, dim, cbs, smp, amp0, nv=300, nl=2)
    say("mcfield full", kind, dim, cbs, smp, amp0)
    n = 1000
    u0 = oracle.mcfield(spec, n)
    u = E.mcfield_spec(spec, n, "fp64")
    assert field_err(u, u0) <= 1e-10
    u1 = E.mcfield_spec(spec, n, "mixed")
    assert field_err(u1, u0) <= 2e-5


def test_gpu_headline_size_mixed_vs_fp64(oracle):
    """4096^2 covariance (64 lights x 64 views on a 15 degree cone, HG 0.9,
    slab OD 20): mixed vs FP64 on the same paths <= 1e-4, and the FP64 GPU
    path vs the oracle on a few replayed paths <= 1e-12."""
    rng = np.random.default_rng(2024)
    l, v = cone_dirs(64, 15, rng), cone_dirs(64, 15, rng)
    spec = E.make_problem(20.0, 1.0, [-5e4, -5e4, 0], [5e4, 5e4, 1], l, v, 1, 1, 1.0, True, 2, 3, 0.9, seed=0)
    say("headline 4096 replay fp64 vs oracle (3 paths)")
    ps = oracle.trace(spec, 3)
    Ms0, Mm0 = oracle.mccov_replay(spec, ps, 3)
    Ms, Mm = E.mccov_replay(spec, ps, 3, "fp64")
    assert cs_err(Mm, Mm0) <= 1e-12
    assert cs_err(Ms, Ms0) <= 1e-12
    say("headline 4096 mixed vs fp64 (300 paths)")
    Msd, Mmd = E.mccov_spec(spec, 300, "fp64")
    Msm, Mmm = E.mccov_spec(spec, 300, "mixed")
    err = cs_err(Msm + Mmm, Msd + Mmd)
    say("headline mixed cs err", err)
    assert err <= 1e-4


def test_gpu_headline_size_multi_chunk_pipeline():
    """12 000 paths at 4096^2 run as several chunks in one call (the tracer of
    chunk k+1 beside the update of chunk k, the two buffer slots reused):
    mixed vs FP64 on the same paths <= 1e-4, and FP64 in one call vs the same
    paths in four session slices <= 1e-12."""
    rng = np.random.default_rng(2024)
    l, v = cone_dirs(64, 15, rng), cone_dirs(64, 15, rng)
    spec = E.make_problem(20.0, 1.0, [-5e4, -5e4, 0], [5e4, 5e4, 1], l, v, 1, 1, 1.0, True, 2, 3, 0.9, seed=7)
    n = 12000
    say("headline multi-chunk fp64")
    with E.Session(spec, "cov", "fp64") as s:
        s.run(0, n)
        Msd, Mmd = s.result(n)
        assert s.stats()["chunks"] >= 3
    say("headline multi-chunk mixed")
    with E.Session(spec, "cov", "mixed") as s:
        s.run(0, n)
        Msm, Mmm = s.result(n)
        assert s.stats()["chunks"] >= 2
    err = cs_err(Msm + Mmm, Msd + Mmd)
    say("  mixed cs err", err)
    assert err <= 1e-4
    say("headline fp64 in slices")
    with E.Session(spec, "cov", "fp64") as s:
        for a in range(0, n, 3000):
            s.run(a, 3000)
        Ms1, Mm1 = s.result(n)
    assert cs_err(Ms1, Msd) <= 1e-12 and cs_err(Mm1, Mmd) <= 1e-12


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_agrees_with_oracle_within_3_standard_errors(oracle, precision):
    """Independent seeds: per covariance entry (real and imaginary parts of the
    64 x 64 plain-MCcov configuration: slab T=1, sigma_s=5, sigma_a=0.1,
    isotropic, 8 lights x 8 views in [0, 20] deg, CBS), the GPU estimate and
    the oracle estimate differ by less than 3 combined standard errors
    (batch means, 20 batches each) for >= 98% of the entries (the t_19
    distribution puts ~0.7% beyond 3)."""
    rng = np.random.default_rng(0)
    l, v = far_dirs(8, 20, rng), far_dirs(8, 20, rng)
    sigt = 5.1
    spec = E.make_problem(sigt, 5 / sigt, [-5e4, -5e4, 0], [5e4, 5e4, 1], l, v, 1, 1, 1.0, True, 2, 1, 0, seed=1)
    B, ng, no = 20, 4000, 1000
    say("3-SE gpu", precision)
    gb = []
    with E.Session(spec, "cov", precision) as s:
        for b in range(B):
            s2 = E.Session(spec, "cov", precision)
            s2.run(b * ng, ng)
            Ms, Mm = s2.result(ng)
            s2.close()
            gb.append(cov_matrix(Ms + Mm))
    say("3-SE oracle")
    ospec = E.make_problem(sigt, 5 / sigt, [-5e4, -5e4, 0], [5e4, 5e4, 1], l, v, 1, 1, 1.0, True, 2, 1, 0, seed=2)
    ob = []
    V = np.prod(ospec.box_max - ospec.box_min)
    for b in range(B):
        Ms, Mm = oracle.mccov_sums(ospec, b * no, no, 16)
        ob.append(cov_matrix(Ms + Mm / 2) * V * sigt / no)
    gb, ob = np.array(gb), np.array(ob)
    mg, mo = gb.mean(0), ob.mean(0)
    # the results are exactly Hermitian with a real diagonal in both modes
    for Cb in gb:
        assert np.array_equal(Cb, Cb.conj().T)
    z = []
    for part in (np.real, np.imag):
        # entries that are identically zero in the oracle (Im C_ii) are exactly
        # zero on the GPU too, and carry no statistics
        zero = np.all(part(ob) == 0, axis=0)
        assert np.all(part(gb)[:, zero] == 0)
        seg = part(gb).std(0, ddof=1) / np.sqrt(B)
        seo = part(ob).std(0, ddof=1) / np.sqrt(B)
        se = np.sqrt(seg ** 2 + seo ** 2)
        z.append((np.abs(part(mg) - part(mo))[~zero] / se[~zero]).ravel())
    z = np.concatenate(z)
    say("  fraction within 3 SE", np.mean(z < 3), "max z", z.max())
    assert np.mean(z < 3) >= 0.98
    assert z.max() < 6


# ------------------------------------------------------ near-field sources
NEAR_CASES = [pytest.param(*c, id="-".join(map(str, c))) for c in
              [("hg", 3, True, "both"), ("iso", 3, False, "l"), ("tab", 3, True, "v"), ("hg", 2, True, "both"),
               ("iso", 2, False, "v")]]


def near_spec(kind, dim, cbs, which, nv=5, nl=3, seed=11, smp=2):
    """small_spec with point sources: lights 1-3 below the slab, views 1-3
    above it (MCcov.m:127-142, evalphaseatt.m:10-16)."""
    f = build(kind, dim, cbs, smp, None, nv=nv, nl=nl, seed=seed)
    rng = np.random.default_rng(seed + 100)
    top = f.box_max[-1]

    def points(n, z):
        return np.vstack([rng.uniform(-2, 2, (dim - 1, n)), z])

    near_l, near_v = which in ("l", "both"), which in ("v", "both")
    l = points(nl, -rng.uniform(1, 3, nl)) if near_l else f.l
    v = points(nv, top + rng.uniform(1, 3, nv)) if near_v else f.v
    amp = f.amp.g if f.amp.type == 3 else (0 if f.amp.type == 1 else tab3d())
    return E.make_problem(f.sigt, f.albedo, f.box_min, f.box_max, l, v, not near_l, not near_v, f.lam, cbs, smp,
                          f.amp.type, amp, seed=seed)


@pytest.mark.parametrize("kind,dim,cbs,which", NEAR_CASES)
def test_gpu_near_field_trace_and_replay(oracle, kind, dim, cbs, which):
    spec = near_spec(kind, dim, cbs, which)
    say("near trace+replay", kind, dim, cbs, which)
    n = 400
    g = E.trace(spec, n)
    ps = oracle.trace(spec, n)
    assert np.array_equal(g["offset"], ps["offset"])
    np.testing.assert_allclose(g["x"], ps["x"], rtol=0, atol=1e-10)
    Ms0, Mm0 = oracle.mccov_replay(spec, ps, n)
    for precision, tol in (("fp64", 1e-12), ("mixed", 1e-4)):
        Ms, Mm = E.mccov_replay(spec, ps, n, precision)
        assert cs_err(Ms, Ms0) <= tol, precision
        assert cs_err(Mm, Mm0) <= tol, precision
    u0 = oracle.mcfield_replay(spec, ps, n)
    for precision, tol in (("fp64", 1e-12), ("mixed", 2e-5)):
        u = E.mcfield_replay(spec, ps, n, precision)
        assert field_err(u, u0) <= tol, precision


@pytest.mark.parametrize("kind,dim,cbs,which", NEAR_CASES)
def test_gpu_near_field_full_run_same_seed(oracle, kind, dim, cbs, which):
    spec = near_spec(kind, dim, cbs, which, nv=6, nl=4)
    say("near full", kind, dim, cbs, which)
    n = 1500
    Ms0, Mm0 = oracle.mccov(spec, n)
    Ms, Mm = E.mccov_spec(spec, n, "fp64")
    assert cs_err(Ms, Ms0) <= 1e-10
    assert cs_err(Mm, Mm0) <= 1e-10
    Ms1, Mm1 = E.mccov_spec(spec, n, "mixed")
    assert cs_err(Ms1 + Mm1, Ms0 + Mm0) <= 1e-4
    u0 = oracle.mcfield(spec, n)
    assert field_err(E.mcfield_spec(spec, n, "fp64"), u0) <= 1e-10
    assert field_err(E.mcfield_spec(spec, n, "mixed"), u0) <= 2e-5


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-4)])
def test_gpu_near_field_tile_edges(oracle, precision, tol):
    """Near-field single-scatter rows at N = 130 (crosses the 64/128 tiles)."""
    spec = near_spec("hg", 3, True, "both", nv=26, nl=5, seed=3)
    say("near tile edges", precision)
    n = 60
    ps = oracle.trace(spec, n)
    Ms0, Mm0 = oracle.mccov_replay(spec, ps, n)
    Ms, Mm = E.mccov_replay(spec, ps, n, precision)
    assert cs_err(Ms, Ms0) <= tol
    assert cs_err(Mm, Mm0) <= tol


# ------------------------------------------------ sessions and tiny runs
@pytest.mark.parametrize("kind", ["cov", "field"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-4)])
def test_gpu_session_slices_equal_one_run(kind, precision, tol):
    """A solve resumed in slices of paths (sessions: run(first, n) adds paths
    [first, first + n)) equals the same paths in one call; the slices cut
    through chunks at other places, so FP64 agrees to rounding."""
    spec = build("hg", 3, True, 2, None, nv=6, nl=4, seed=3)
    say("session slices", kind, precision)
    n, cuts = 2500, [0, 700, 701, 1900, 2500]
    with E.Session(spec, kind, precision) as s:
        for a, b in zip(cuts[:-1], cuts[1:]):
            s.run(a, b - a)
        out = s.result(n)
    if kind == "cov":
        Ms0, Mm0 = E.mccov_spec(spec, n, precision)
        assert cs_err(out[0], Ms0) <= tol and cs_err(out[1], Mm0) <= tol
    else:
        u0 = E.mcfield_spec(spec, n, precision)
        assert field_err(out, u0) <= (1e-12 if precision == "fp64" else 2e-5)


@pytest.mark.parametrize("kind,dim,cbs,smp,amp0", spec_cases()[:3])
def test_gpu_single_path_runs(oracle, kind, dim, cbs, smp, amp0):
    """maxItr = 1 (MCcov.m:293-299 and MCfield.m:253-254 normalise by it)."""
    spec = build(kind, dim, cbs, smp, amp0)
    say("single path", kind, dim, cbs, smp, amp0)
    Ms0, Mm0 = oracle.mccov(spec, 1)
    Ms, Mm = E.mccov_spec(spec, 1, "fp64")
    scale = max(np.abs(Ms0).max(), np.abs(Mm0).max(), 1e-300)
    assert np.abs(Ms - Ms0).max() <= 1e-10 * scale and np.abs(Mm - Mm0).max() <= 1e-10 * scale
    u0 = oracle.mcfield(spec, 1)
    u = E.mcfield_spec(spec, 1, "fp64")
    assert np.abs(u - u0).max() <= 1e-10 * max(np.abs(u0).max(), 1e-300)


# ------------------------------------------------------ non-finite inputs
@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gpu_nonfinite_path_weights_are_flagged(oracle, precision):
    """A NaN path weight injected into a replayed path set is detected: the
    solve reports the vertices / paths whose factors came out non-finite
    (scmc_stats.nonfinite) and, as the reference would, the result holds
    NaN; the clean set reports 0."""
    spec = build("hg", 3, True, 2, None)
    n = 200
    ps = oracle.trace(spec, n)
    st = {}
    Ms, Mm = E.mccov_replay(spec, ps, n, precision, stats=st)
    assert st["nonfinite"] == 0 and np.isfinite(Ms).all() and np.isfinite(Mm).all()
    u = E.mcfield_replay(spec, ps, n, precision, stats=st)
    assert st["nonfinite"] == 0 and np.isfinite(u).all()
    off = ps["offset"]
    p = int(np.argmax(np.diff(off) >= 3))          # a path with >= 3 vertices
    bad = dict(ps)
    bad["weight"] = ps["weight"].copy()
    bad["weight"][off[p] + 1] = np.nan            # its second vertex
    say("nonfinite", precision, "path", p)
    st = {}
    Ms, Mm = E.mccov_replay(spec, bad, n, precision, stats=st)
    assert st["nonfinite"] == 1
    assert not np.isfinite(Mm).all()
    st = {}
    u = E.mcfield_replay(spec, bad, n, precision, stats=st)
    assert st["nonfinite"] == 1
    assert not np.isfinite(u).all()


@pytest.mark.parametrize("kind,dim,cbs", [("hg", 3, True), ("hg", 2, False), ("tab", 3, True), ("iso", 3, False)])
def test_gpu_mcfield_one_light(oracle, kind, dim, cbs):
    """One light: the two-views-per-lane field instance (HG amplitudes have their
    own instance), 700 views so a lane's second view block is ragged."""
    spec = build(kind, dim, cbs, 2, None, nv=700, nl=1)
    say("mcfield one light", kind, dim, cbs)
    n = 400
    ps = oracle.trace(spec, n)
    u0 = oracle.mcfield_replay(spec, ps, n)
    assert field_err(E.mcfield_replay(spec, ps, n, "mixed"), u0) <= 2e-5
    assert field_err(E.mcfield_replay(spec, ps, n, "fp64"), u0) <= 1e-12
    u1 = oracle.mcfield(spec, n)
    assert field_err(E.mcfield_spec(spec, n, "mixed"), u1) <= 2e-5


@pytest.mark.parametrize("kind,dim", [("hg", 3), ("tab", 3), ("hg", 2)])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("mixed", 1e-4)])
def test_gpu_cbs_reciprocity(kind, dim, precision, tol):
    """The GPU's CBS update is reciprocal: the exchanged problem (lights -v,
    views -l) walks the same paths (smpFlg = 1) and gives the covariances
    with the view and light index pairs exchanged (MCcov.m:244-254)."""
    from test_oracle_engines import swap_lights_and_views
    spec = build(kind, dim, True, 1, None, nv=6, nl=4, seed=23)
    sw = swap_lights_and_views(spec)
    say("cbs reciprocity", kind, dim, precision)
    n = 3000
    Ms, Mm = E.mccov_spec(spec, n, precision)
    Ms2, Mm2 = E.mccov_spec(sw, n, precision)
    assert cs_err(Ms2, Ms.transpose(2, 3, 0, 1)) <= tol
    assert cs_err(Mm2, Mm.transpose(2, 3, 0, 1)) <= tol
