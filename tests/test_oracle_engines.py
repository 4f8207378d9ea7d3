"""Pins of the oracle's MCcov / MCfield restatement (CPU only).

* Brute force: the covariance is recomputed from an exported path set in the
  rank-1 form C = sum_k w_k u_k u_k^H with u = f (x) g + f' (x) g' (numpy,
  written from the mathematics, not from MCcov.m's M00/Mnn/Mn0/M0n loops) and
  must equal the oracle to 1e-12.
* Single-scatter known answer: E[Ms(v,v,l,l)] = sigma*albedo*A*a^2*I_z for a
  wide slab (closed-form z integral), within 3 standard errors.
* MCfield second moment: over the random phases, E[u u^H] equals the
  covariance of the same paths (MCfield.m:207-222 vs MCcov.m:232-264).
* Invariants: Hermitian, positive semi-definite.
* MCfieldOnWave: equals MCfield contracted with the weights Wl on the same
  paths (every term of MCfieldOnWave.m is linear in the lights).
* Near-field sources: placed at distance R along the far-field directions,
  the near-field covariance and field tend to the far-field ones (factor
  R^(dim-1) per near side, e^{ikR} per near side in the field).
"""
import numpy as np
import pytest

from conftest import cone_dirs, cov_matrix, cs_scale, far_dirs
from mcspeckle_amd.engines import make_problem


# ------------------------------------------------------- numpy restatement

def np_E(x, D, sigma, lam, bmin, bmax):
    """exp(i k x.d - sigma/2 * dist(x -> boundary along -d)), far field."""
    dim = D.shape[0]
    ph = 2 * np.pi / lam * (x[:dim] @ D)
    nd = -D
    dist = np.full(D.shape[1], np.inf)
    for j in range(dim):
        a = np.maximum(np.abs(nd[j]), 0.1 ** 7)
        dj = np.where(nd[j] <= 0, (x[j] - bmin[j]) / a, (bmax[j] - x[j]) / a)
        dist = np.minimum(dist, dj)
    return np.exp(1j * ph - sigma / 2 * dist)


def np_amp(c, spec):
    ph, dim = spec.amp, spec.dim
    c = np.asarray(c, dtype=float)
    if ph.type == 1:
        return np.full(c.shape, 1 / np.sqrt(2 ** (dim - 1) * np.pi), dtype=complex)
    if ph.type == 3:
        g = ph.g
        if dim == 3:
            hg = (1 - g * g) / (4 * np.pi * (1 + g * g - 2 * g * c) ** 1.5)
        else:
            hg = (1 - g * g) / (2 * np.pi * (1 + g * g - 2 * g * c))
        return np.sqrt(hg).astype(complex)
    n = ph.eval_amp.size
    ang = np.arccos(np.clip(c, -1, 1))
    scale = (n - 1) / (2 * np.pi) if dim == 2 else (n - 1) / np.pi
    idx = np.clip(np.floor(ang * scale + 0.5).astype(int), 0, n - 1)
    return ph.eval_amp[idx]


def np_cov_from_paths(spec, ps, max_itr):
    """Rank-1 form of the covariance from a path set."""
    dim, nv, nl = spec.dim, spec.nv, spec.nl
    N = nv * nl
    L, V = spec.l, spec.v
    bmin, bmax = spec.box_min, spec.box_max
    Cs = np.zeros((N, N), complex)
    Cm = np.zeros((N, N), complex)
    A = np.stack([np_amp(L[:, j] @ V, spec) for j in range(nl)], axis=1)   # nv x nl
    off = ps["offset"]
    for p in range(off.size - 1):
        x1 = ps["x1"][3 * p:3 * p + dim]
        w1 = ps["w1"][3 * p:3 * p + dim]
        w0p = ps["w0p_re"][p] + 1j * ps["w0p_im"][p]
        px = ps["px"][p]
        el0 = np_E(x1, L, spec.sigt, spec.lam, bmin, bmax)
        ev0 = np_E(x1, -V, spec.sigt, spec.lam, bmin, bmax)
        g = el0 * np_amp(w1 @ L, spec) / w0p            # e_l0_ms
        f2 = ev0 * np_amp(-w1 @ V, spec) / w0p          # e_v0_ms
        for k in range(off[p], off[p + 1]):
            W = ps["weight"][k]
            if k == off[p]:
                s = (np.outer(ev0, el0) * A).reshape(-1, order="F")   # index v + nv*l
                Cs += W / px * np.outer(s, s.conj())
                continue
            x = ps["x"][3 * k:3 * k + dim]
            ow = ps["ow"][3 * k:3 * k + dim]
            f = np_E(x, -V, spec.sigt, spec.lam, bmin, bmax) * np_amp(ow @ V, spec)
            u = np.outer(f, g).reshape(-1, order="F")
            if spec.do_cbs:
                g2 = np_E(x, L, spec.sigt, spec.lam, bmin, bmax) * np_amp(-ow @ L, spec)
                u = u + np.outer(f2, g2).reshape(-1, order="F")
            Cm += W / px * np.outer(u, u.conj())
    Vol = np.prod(bmax - bmin)
    norm = Vol * spec.sigt / max_itr
    return Cs * norm, Cm * norm * (0.5 if spec.do_cbs else 1.0)


def small_spec(kind, dim=3, cbs=True, nv=3, nl=2, seed=11, smp=2, amp0=None, tab=None):
    rng = np.random.default_rng(seed)
    if dim == 3:
        l, v = far_dirs(nl, 25, rng), far_dirs(nv, 25, rng)
        bmin, bmax = [-3, -4, 0], [3, 4, 1.2]
    else:
        l, v = rng.uniform(-0.3, 0.3, nl), rng.uniform(-0.4, 0.4, nv)   # angles (MCcov.m:56-70)
        bmin, bmax = [-3, 0], [3, 1.2]
    if kind == "iso":
        st, af = 1, 0
    elif kind == "hg":
        st, af = 3, 0.7
    else:
        st, af = 2, tab
    return make_problem(2.5, 0.9, bmin, bmax, l, v, 1, 1, 0.8, cbs, smp, st, af, amp0, seed=seed)


def tab3d(g=0.6, n=181):
    from mcspeckle_amd.interface import tabulatedAmplitudeScatter
    th = np.linspace(0, np.pi, n)
    f = np.sqrt((1 - g * g) / (1 + g * g - 2 * g * np.cos(th)) ** 1.5) * np.exp(0.3j * th)
    return tabulatedAmplitudeScatter(th, f)["ampfunc"]


CASES = [
    ("iso", 3, True, 2, None), ("iso", 3, False, 1, None), ("hg", 3, True, 2, None), ("hg", 3, False, 2, 0.5),
    ("hg", 2, True, 2, None), ("tab", 3, True, 2, None), ("tab", 3, False, 1, "tab0"), ("iso", 2, False, 1, None),
]


@pytest.mark.parametrize("kind,dim,cbs,smp,amp0", CASES)
def test_mccov_bruteforce_rank1_form(oracle, kind, dim, cbs, smp, amp0):
    tab = tab3d() if kind == "tab" else None
    a0 = tab3d(0.3) if amp0 == "tab0" else amp0
    spec = small_spec(kind, dim, cbs, smp=smp, amp0=a0, tab=tab)
    n = 12
    ps = oracle.trace(spec, n)
    Ms, Mm = oracle.mccov_replay(spec, ps, n)
    Cs, Cm = np_cov_from_paths(spec, ps, n)
    for got, want in ((cov_matrix(Ms), Cs), (cov_matrix(Mm), Cm)):
        scale = np.abs(want).max()
        assert scale > 0
        np.testing.assert_allclose(got, want, atol=1e-12 * scale, rtol=0)
    # full run over the same paths is the replay
    Ms2, Mm2 = oracle.mccov(spec, n)
    np.testing.assert_array_equal(Ms, Ms2)
    np.testing.assert_array_equal(Mm, Mm2)


def test_mccov_hermitian_psd(oracle):
    spec = small_spec("hg", 3, True, nv=4, nl=3)
    Ms, Mm = oracle.mccov(spec, 400)
    for M in (Ms, Mm, Ms + Mm):
        C = cov_matrix(M)
        np.testing.assert_allclose(C, C.conj().T, atol=1e-13 * np.abs(C).max())
        ev = np.linalg.eigvalsh((C + C.conj().T) / 2)
        assert ev.min() > -1e-10 * ev.max()
        assert np.all(np.real(np.diag(C)) > 0)


def test_single_scatter_known_answer(oracle):
    """E[Ms(v,v,l,l)] = sigma*albedo*A*a^2 * int_0^T exp(-sigma((T-z)/v_z + z/l_z)) dz
    for a slab wide enough that lateral exits are negligible."""
    rng = np.random.default_rng(5)
    l, v = far_dirs(2, 20, rng), far_dirs(2, 20, rng)
    sigma, alb, T, W = 2.0, 0.8, 1.0, 200.0
    spec = make_problem(sigma, alb, [-W, -W, 0], [W, W, T], l, v, 1, 1, 1.0, False, 2, 1, 0, seed=3)
    B, n = 16, 400
    est = []
    for b in range(B):
        spec.path_offset = b * n
        ps = oracle.trace(spec, n)
        Ms, _ = oracle.mccov_replay(spec, ps, n)
        est.append(Ms)
    est = np.array(est)
    mean, se = est.mean(0), est.std(0, ddof=1) / np.sqrt(B)
    a2 = 1 / (4 * np.pi)
    A = (2 * W) ** 2
    for i in range(2):
        for j in range(2):
            vz, lz = v[2, i], l[2, j]
            q = sigma * (1 / vz - 1 / lz)
            Iz = np.exp(-sigma * T / vz) * ((np.exp(q * T) - 1) / q if abs(q) > 1e-12 else T)
            want = sigma * alb * A * a2 * Iz
            got = mean[i, i, j, j]
            assert abs(got.imag) < 1e-9 * abs(want)
            assert abs(got.real - want) < 3.5 * se[i, i, j, j].real, (got, want, se[i, i, j, j])


@pytest.mark.parametrize("cbs", [False, True])
def test_mcfield_second_moment_is_the_covariance(oracle, cbs):
    """Over the per-vertex random phases xi, E[u u^H] = Ms + Mm of the same paths."""
    spec = small_spec("hg", 3, cbs, nv=3, nl=2, seed=21)
    n = 20
    ps = oracle.trace(spec, n)
    Ms, Mm = oracle.mccov_replay(spec, ps, n)
    C = cov_matrix(Ms + Mm)
    R = 3000
    rng = np.random.default_rng(9)
    acc = np.zeros_like(C)
    acc2 = np.zeros(C.shape)
    for _ in range(R):
        q = dict(ps)
        q["xi"] = rng.uniform(size=ps["xi"].size)
        u = oracle.mcfield_replay(spec, q, n).reshape(-1, order="F")
        o = np.outer(u, u.conj())
        acc += o
        acc2 += np.abs(o) ** 2
    mean = acc / R
    se = np.sqrt(np.maximum(acc2 / R - np.abs(mean) ** 2, 0) / R)
    z = np.abs(mean - C) / np.maximum(se, 1e-300)
    assert np.mean(z < 3) > 0.95
    assert np.abs(mean - C).max() < 6 * se.max()


def test_mcfield_first_vertex_term_exact(oracle):
    """A path set whose every path has one vertex: u = sum_p sqrt(W/px) e^{2 pi i xi} s_p."""
    spec = small_spec("iso", 3, False, nv=3, nl=2, seed=4)
    ps = oracle.trace(spec, 40)
    keep = np.flatnonzero(np.diff(ps["offset"]) == 1)[:10]
    assert keep.size >= 3
    sub = {k: [] for k in ps}
    off = [0]
    for p in keep:
        o = ps["offset"][p]
        for k in ("x1", "w1"):
            sub[k].append(ps[k][3 * p:3 * p + 3])
        for k in ("px", "w0p_re", "w0p_im"):
            sub[k].append([ps[k][p]])
        for k in ("x", "ow"):
            sub[k].append(ps[k][3 * o:3 * o + 3])
        for k in ("weight", "xi"):
            sub[k].append([ps[k][o]])
        off.append(off[-1] + 1)
    q = {k: np.concatenate(v) for k, v in sub.items() if k != "offset"}
    q["offset"] = np.array(off, dtype=np.int64)
    u = oracle.mcfield_replay(spec, q, 7)
    A = np.stack([np_amp(spec.l[:, j] @ spec.v, spec) for j in range(spec.nl)], axis=1)
    want = np.zeros((spec.nv, spec.nl), complex)
    for p in range(keep.size):
        x1 = q["x1"][3 * p:3 * p + 3]
        s = np.outer(np_E(x1, -spec.v, spec.sigt, spec.lam, spec.box_min, spec.box_max),
                     np_E(x1, spec.l, spec.sigt, spec.lam, spec.box_min, spec.box_max)) * A
        want += np.sqrt(q["weight"][p] / q["px"][p]) * np.exp(2j * np.pi * q["xi"][p]) * s
    want *= np.sqrt(np.prod(spec.box_max - spec.box_min) * spec.sigt / 7)
    np.testing.assert_allclose(u, want, rtol=1e-12, atol=1e-13 * np.abs(want).max())


def near_far_pair(kind, dim, cbs, R, which="both", nv=3, nl=2, seed=31):
    """A far-field problem and the same problem with its sources placed at
    distance R along the directions (lights at -R l, views at +R v).  Uniform
    first vertices and an isotropic first direction make the walks of the two
    identical (neither depends on the sources)."""
    f = small_spec(kind, dim, cbs, nv=nv, nl=nl, seed=seed, smp=1, tab=tab3d() if kind == "tab" else None)
    near_l = which in ("both", "l")
    near_v = which in ("both", "v")
    l = -R * f.l if near_l else f.l
    v = R * f.v if near_v else f.v
    amp = f.amp.g if f.amp.type == 3 else (0 if f.amp.type == 1 else tab3d())
    n = make_problem(f.sigt, f.albedo, f.box_min, f.box_max, l, v, not near_l, not near_v, f.lam, cbs, 1,
                     f.amp.type, amp, seed=f.seed)
    return f, n


@pytest.mark.parametrize("kind,dim,cbs,which", [("hg", 3, True, "both"), ("iso", 3, False, "l"),
                                                ("tab", 3, True, "v"), ("hg", 2, True, "both")])
def test_near_field_sources_tend_to_far_field(oracle, kind, dim, cbs, which):
    """Sources at distance R: every near-field factor is the far-field one times
    e^{ikR} / R^((dim-1)/2) up to O(|x|/R) in directions and O(k|x|^2/R) in phase,
    so R^(dim-1) per near side times the covariance tends to the far-field one
    (the e^{ikR} cancel in u u^H), and the field picks up e^{ikR} per near side."""
    R = 1e8
    f, nr = near_far_pair(kind, dim, cbs, R, which)
    sides = 2 if which == "both" else 1
    n = 30
    psf, psn = oracle.trace(f, n), oracle.trace(nr, n)
    for key in ("offset", "x", "ow", "weight", "x1", "w1", "px"):
        np.testing.assert_array_equal(psf[key], psn[key])
    Msf, Mmf = oracle.mccov_replay(f, psf, n)
    Msn, Mmn = oracle.mccov_replay(nr, psn, n)
    sc = (R + 0.01 * f.lam) ** ((dim - 1) * sides)
    for near, far in ((Msn, Msf), (Mmn, Mmf)):
        C, Cf = cov_matrix(near) * sc, cov_matrix(far)
        assert np.max(np.abs(C - Cf) / cs_scale(Cf)) < 2e-5
    uf = oracle.mcfield_replay(f, psf, n)
    un = oracle.mcfield_replay(nr, psn, n)
    k = 2 * np.pi / f.lam
    un = un * np.exp(-1j * k * R * sides) * (R + 0.01 * f.lam) ** ((dim - 1) / 2 * sides)
    assert np.abs(un - uf).max() < 2e-5 * np.abs(uf).max()


def test_near_field_covariance_hermitian_psd(oracle):
    _, nr = near_far_pair("hg", 3, True, 6.0, "both", nv=4, nl=3)
    Ms, Mm = oracle.mccov(nr, 300)
    for M in (Ms, Mm, Ms + Mm):
        C = cov_matrix(M)
        np.testing.assert_allclose(C, C.conj().T, atol=1e-13 * np.abs(C).max())
        ev = np.linalg.eigvalsh((C + C.conj().T) / 2)
        assert ev.min() > -1e-10 * ev.max()


@pytest.mark.parametrize("kind,dim,cbs,near", [("hg", 3, True, False), ("tab", 3, False, False),
                                               ("iso", 3, True, False), ("hg", 2, True, False),
                                               ("hg", 3, True, True)])
def test_mcfield_wave_is_mcfield_contracted_with_the_weights(oracle, kind, dim, cbs, near):
    if near:
        _, spec = near_far_pair(kind, dim, cbs, 5.0, "both", nv=4, nl=3)
    else:
        spec = small_spec(kind, dim, cbs, nv=4, nl=3, seed=12, tab=tab3d() if kind == "tab" else None)
    rng = np.random.default_rng(3)
    wl = rng.uniform(0.2, 1, spec.nl) * np.exp(2j * np.pi * rng.uniform(size=spec.nl))
    n = 200
    u = oracle.mcfield(spec, n)
    uw = oracle.mcfield_wave(spec, wl, n)
    want = u @ wl
    np.testing.assert_allclose(uw, want, rtol=0, atol=1e-12 * np.abs(want).max())


def swap_lights_and_views(spec):
    """The exchanged problem: lights -v, views -l (same walks when the first
    vertex is drawn uniformly, smpFlg = 1, and the first direction
    isotropically)."""
    import dataclasses
    return dataclasses.replace(spec, l=np.ascontiguousarray(-spec.v), v=np.ascontiguousarray(-spec.l))


def cs_rel(A, B):
    from conftest import cov_matrix, cs_scale
    C, R = cov_matrix(A), cov_matrix(B)
    return float(np.max(np.abs(C - R) / np.maximum(cs_scale(R), 1e-300)))


@pytest.mark.parametrize("kind,dim", [("iso", 3), ("hg", 3), ("tab", 3), ("hg", 2)])
def test_cbs_reciprocity(oracle, kind, dim):
    """With CBS (MCcov.m:244-254) each vertex adds (f + r)(f + r)^H, and the
    reversed path r of the problem (l, v) is the forward path f of the
    exchanged problem (lights -v, views -l): e_l(x)[-v] = e_v(x)[v],
    e_v(x)[-l] = e_l(x)[l], a(ow.(-l)) = a(-ow.l), a(w1.(-v)) = a(-w1.v).  On
    the same walks, Ms and Mm of the exchanged problem are therefore those of
    the original with the view and light index pairs exchanged,
    M'(l1, l2, v1, v2) = M(v1, v2, l1, l2).  Without CBS Mm has no such
    symmetry (the check has the power to see a missing reversed path)."""
    tab = tab3d() if kind == "tab" else None
    spec = small_spec(kind, dim, True, nv=4, nl=3, seed=21, smp=1, tab=tab)
    sw = swap_lights_and_views(spec)
    n = 400
    Ms, Mm = oracle.mccov(spec, n)
    Ms2, Mm2 = oracle.mccov(sw, n)
    assert cs_rel(Ms2, Ms.transpose(2, 3, 0, 1)) <= 1e-12
    assert cs_rel(Mm2, Mm.transpose(2, 3, 0, 1)) <= 1e-12
    import dataclasses
    no = dataclasses.replace(spec, do_cbs=False)
    _, Mm0 = oracle.mccov(no, n)
    _, Mm0s = oracle.mccov(swap_lights_and_views(no), n)
    assert cs_rel(Mm0s, Mm0.transpose(2, 3, 0, 1)) > 1e-3
