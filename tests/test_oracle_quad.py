"""Pins of the oracle's MCfieldFourierQuad and MCfieldFourierQuadOnWave
restatements (CPU only).

A second restatement written with numpy's FFT routines (fftshift, ifftshift,
fft2 and flips exactly as MCfieldFourierQuad.m:276-280 spells them) and the
oracle's own walks (traced in the enlarged box) must reproduce the oracle's
direct-DFT result, at odd and (OnWave only) even transform sizes; the wave
variant must be linear in its weights; the grid sizes must match the C-ABI's
scmc_quad_grid."""
import numpy as np
import pytest

from mcspeckle_amd import engines as E


def quad_numpy(oracle, sigt, albedo, bmin, bmax, l, v_max, v_stp, n_paths, lam, cbs, smp, vsign, seed, wl=None,
               sct_type=1, ampfunc=0):
    """MCfieldFourierQuad.m, or with wl MCfieldFourierQuadOnWave.m, in numpy.
    sct_type/ampfunc: the walk's phase function (isotropic in the reference)."""
    wave = wl is not None
    g = oracle.quad_grid(np.array(bmin, float), np.array(bmax, float), v_max, v_stp, lam, wave=wave)
    Bmin, Bmax = np.array(g.box_min[:]), np.array(g.box_max[:])
    spec = E.make_problem(sigt, albedo, Bmin, Bmax, l, np.zeros((3, 0)), 1, 1, lam, cbs, smp, sct_type, ampfunc,
                          lmean0=[0, 0, 1], seed=seed, need_views=False)
    ps = oracle.trace(spec, n_paths)
    return quad_sum(g, ps, sigt, l, v_max, v_stp, lam, cbs, vsign, wl)


def quad_sum(g, ps, sigt, l, v_max, v_stp, lam, cbs, vsign, wl=None):
    """The vertex store, slab scatter and transforms of MCfieldFourierQuad.m:
    218-303 (MCfieldFourierQuadOnWave.m with wl) over a given path set in the
    grid's box g.  Returns (u, K)."""
    wave = wl is not None
    nv0, nx = g.nv0, g.nx
    Bmin, Bmax = np.array(g.box_min[:]), np.array(g.box_max[:])
    n_paths = ps["offset"].size - 1
    off = ps["offset"]
    xL, x1L, pxL = [], [], []
    for p in range(n_paths):
        for k in range(off[p], off[p + 1]):
            xL.append(ps["x"][3 * k:3 * k + 3])
            x1L.append(ps["x1"][3 * p:3 * p + 3])
            c = np.sqrt(ps["weight"][k] / ps["px"][p]) * np.exp(2j * np.pi * ps["xi"][k])
            if cbs and k > off[p]:
                c /= np.sqrt(2)
            pxL.append(c)
    xL, x1L, pxL = np.array(xL).T, np.array(x1L).T, np.array(pxL)
    K = pxL.size
    theta_v = -v_max + np.arange(nv0) * v_stp
    vx, vy = np.meshgrid(-theta_v, -theta_v, indexing="ij")
    vz = vsign * np.sqrt(1 - vx ** 2 - vy ** 2)
    gz = np.concatenate([Bmin[2] + np.arange(g.nz - 1) * g.z_bin, [Bmax[2]]])
    diratt0 = np.where(vz < 0, 1 / np.abs(vz), 0)
    idiratt0 = np.where(vz > 0, 1 / np.abs(vz), 0)
    lz = l[2, -1]
    l_diratt = 1 / abs(lz) if lz > 0 else 0
    l_idiratt = 1 / abs(lz) if lz < 0 else 0
    # MCfieldFourierQuad.m:134 / MCfieldFourierQuadOnWave.m:143
    zexp0 = (2j if wave else -2j) * np.pi / lam * vz * vsign - sigt / 2 * (diratt0 - idiratt0)
    boxexp0 = np.exp(-sigt / 2 * (-Bmin[2] * diratt0 + Bmax[2] * idiratt0))
    boxexp0 = boxexp0 * np.exp(2j * np.pi / lam * (-Bmin[2] * (vz < 0) * vz + Bmax[2] * (vz > 0) * vz))
    nl = l.shape[1]
    nout = 1 if wave else nl
    u = np.zeros((nv0, nv0, nout), complex)
    pad, crop = g.nx_pad, g.nva_pad
    for tt in range(2):
        if tt == 1:
            if not cbs:
                continue
            xL, x1L = x1L, xL
        order = np.argsort(x1L[2], kind="stable")
        xs, xl, pw = x1L[:, order], xL[:, order], pxL[order]
        slab = np.searchsorted(gz[1:], xs[2], side="right")   # first j with z < gz[j+1]
        for j1 in range(g.nz - 1):
            sel = np.flatnonzero(slab == j1)
            if sel.size == 0:
                continue
            tz = (gz[j1] + gz[j1 + 1]) / 2
            ix = np.clip(np.floor((xs[0, sel] - g.gx0) / g.box_stp + 0.5).astype(int), 0, nx - 1)
            iy = np.clip(np.floor((xs[1, sel] - g.gy0) / g.box_stp + 0.5).astype(int), 0, nx - 1)
            atl = np.exp(-sigt / 2 * ((xl[2, sel] - Bmin[2]) * l_diratt + (Bmax[2] - xl[2, sel]) * l_idiratt))
            for j2 in range(nout):
                if wave:
                    twl = np.asarray(wl).reshape(-1) @ np.exp(vsign * 2j * np.pi / lam * (l.T @ xl[:, sel]))
                    term = pw[sel] * np.exp(vsign * 2j * np.pi / lam * (-xs[2, sel] + tz)) * twl * atl
                else:
                    term = pw[sel] * np.exp(-2j * np.pi / lam * (-xs[2, sel] + tz)) * np.exp(
                        2j * np.pi / lam * (l[:, j2] @ xl[:, sel])) * atl
                uxy = np.zeros((nx, nx), complex)
                np.add.at(uxy, (ix, iy), term)
                uxy = np.pad(uxy, pad)
                tfu = np.fliplr(np.flipud(np.fft.fftshift(np.fft.fft2(np.fft.ifftshift(uxy)))))
                if crop:
                    tfu = tfu[crop:-crop, crop:-crop]
                u[:, :, j2] += tfu * np.exp(zexp0 * tz)
    u *= boxexp0[:, :, None]
    V = (nx * g.box_stp) ** 2 * (Bmax[2] - Bmin[2])
    u = u * np.sqrt(V * sigt / K)
    return (u[:, :, 0] if wave else u), K


@pytest.mark.parametrize("cbs,vmax,vsign", [(False, 0.3, -1.0), (True, 0.3, -1.0), (True, 0.25, 1.0)])
def test_oracle_quad_matches_numpy_fft_restatement(oracle, cbs, vmax, vsign):
    l = np.array([[0.05, -0.1], [0.0, 0.08], [np.sqrt(1 - 0.05 ** 2), np.sqrt(1 - 0.1 ** 2 - 0.08 ** 2)]])
    args = dict(sigt=3.0, albedo=0.95, bmin=[-2.0, -2.0, 0.0], bmax=[2.0, 2.0, 0.6], l=l, v_max=vmax, v_stp=0.05,
                n_paths=150, lam=0.5, cbs=cbs, smp=2, vsign=vsign, seed=4)
    u_np, K = quad_numpy(oracle, **args)
    u_or, K2 = oracle.mcfield_fourier_quad(args["sigt"], args["albedo"], args["bmin"], args["bmax"], l, vmax, 0.05,
                                           150, 0.5, cbs, 2, None, vsign, 4)
    assert K == K2
    np.testing.assert_allclose(u_or, u_np, rtol=0, atol=1e-10 * np.abs(u_np).max())


def tab_phase(n=256, seed=3):
    """A tabulated phase function like the Fourier benchmark's: a seeded
    mixture of three HG lobes on n angle bins."""
    from mcspeckle_amd.interface import tabulatedAmplitudeScatter
    rng = np.random.default_rng(seed)
    g, w = rng.uniform(-0.3, 0.95, 3), rng.dirichlet(np.ones(3))
    th = np.linspace(0, np.pi, n)
    hg = sum(wi * (1 - gi ** 2) / (1 + gi ** 2 - 2 * gi * np.cos(th)) ** 1.5 for gi, wi in zip(g, w))
    return tabulatedAmplitudeScatter(th, np.sqrt(hg))["ampfunc"]


@pytest.mark.parametrize("sct_type,wave", [(3, False), (2, False), (3, True), (2, True)])
def test_oracle_quad_walk_phase_functions(oracle, sct_type, wave):
    """HG and tabulated walks (an extension: the reference walks isotropically):
    the same numpy restatement over the oracle's walks with that phase function."""
    amp = 0.8 if sct_type == 3 else tab_phase()
    ph = E._phase_from_arg(sct_type, amp, "ampfunc")
    args = dict(sigt=3.0, albedo=0.95, bmin=[-2.0, -2.0, 0.0], bmax=[2.0, 2.0, 0.6], l=L2 if wave else L2[:, :2],
                v_max=0.3, v_stp=0.05, n_paths=120, lam=0.5, cbs=True, smp=2, vsign=-1.0, seed=8)
    u_np, K = quad_numpy(oracle, wl=WL if wave else None, sct_type=sct_type, ampfunc=amp, **args)
    a = [args[k] for k in ("sigt", "albedo", "bmin", "bmax", "l")]
    if wave:
        u_or, K2 = oracle.mcfield_fourier_quad_wave(*a, WL, 0.3, 0.05, 120, 0.5, True, 2, None, -1.0, 8, amp=ph)
    else:
        u_or, K2 = oracle.mcfield_fourier_quad(*a, 0.3, 0.05, 120, 0.5, True, 2, None, -1.0, 8, amp=ph)
    assert K == K2
    np.testing.assert_allclose(u_or, u_np, rtol=0, atol=1e-10 * np.abs(u_np).max())
    # the phase function reaches the walk: the isotropic walk stores another vertex set
    if wave:
        _, K0 = oracle.mcfield_fourier_quad_wave(*a, WL, 0.3, 0.05, 120, 0.5, True, 2, None, -1.0, 8)
    else:
        _, K0 = oracle.mcfield_fourier_quad(*a, 0.3, 0.05, 120, 0.5, True, 2, None, -1.0, 8)
    assert K0 != K


@pytest.mark.parametrize("bw,vmax,vstp,lam", [(4.0, 0.3, 0.05, 0.5), (10.0, 0.25, 0.5 / 510, 0.0098),
                                              (2.0, 0.5, 0.01, 1.0)])
def test_quad_grid_sizes_match(oracle, bw, vmax, vstp, lam):
    bmin, bmax = [-bw / 2, -bw / 2, 0.0], [bw / 2, bw / 2, 1.0]
    g = oracle.quad_grid(np.array(bmin), np.array(bmax), vmax, vstp, lam)
    nv0, nx, nz = E.quad_grid(bmin, bmax, vmax, vstp, lam)
    assert (nv0, nx, nz) == (g.nv0, g.nx, g.nz)
    assert nv0 % 2 == 1 and nx % 2 == 1


L2 = np.array([[0.05, -0.1, 0.02], [0.0, 0.08, -0.04],
               [np.sqrt(1 - 0.05 ** 2), np.sqrt(1 - 0.1 ** 2 - 0.08 ** 2), np.sqrt(1 - 0.02 ** 2 - 0.04 ** 2)]])
WL = np.array([0.7 * np.exp(0.3j), 0.4 * np.exp(-1.1j), 0.9 * np.exp(2.0j)])


@pytest.mark.parametrize("cbs,vmax,vsign", [(False, 0.3, -1.0), (True, 0.325, -1.0), (True, 0.3, 1.0),
                                            (False, 0.325, 1.0)])
def test_oracle_quad_wave_matches_numpy_fft_restatement(oracle, cbs, vmax, vsign):
    """v_max 0.325 / v_stp 0.05 gives 14 views per axis: the OnWave grid then
    has an even lateral size and an even transform."""
    args = dict(sigt=3.0, albedo=0.95, bmin=[-2.0, -2.0, 0.0], bmax=[2.0, 2.0, 0.6], l=L2, v_max=vmax, v_stp=0.05,
                n_paths=120, lam=0.5, cbs=cbs, smp=2, vsign=vsign, seed=6)
    g = oracle.quad_grid(np.array(args["bmin"]), np.array(args["bmax"]), vmax, 0.05, 0.5, wave=True)
    assert (g.nv0 - g.nx) % 2 == 0
    u_np, K = quad_numpy(oracle, wl=WL, **args)
    u_or, K2 = oracle.mcfield_fourier_quad_wave(args["sigt"], args["albedo"], args["bmin"], args["bmax"], L2, WL,
                                                vmax, 0.05, 120, 0.5, cbs, 2, None, vsign, 6)
    assert K == K2
    np.testing.assert_allclose(u_or, u_np, rtol=0, atol=1e-10 * np.abs(u_np).max())


def test_oracle_quad_wave_even_transform_size(oracle):
    g = oracle.quad_grid(np.array([-2.0, -2.0, 0.0]), np.array([2.0, 2.0, 0.6]), 0.325, 0.05, 0.5, wave=True)
    assert g.nv0 == 14 and g.nx % 2 == 0 and g.nfft % 2 == 0


def test_oracle_quad_wave_is_linear_in_the_weights(oracle):
    args = (3.0, 0.95, [-2.0, -2.0, 0.0], [2.0, 2.0, 0.6], L2)
    rest = (0.3, 0.05, 80, 0.5, True, 2, None, -1.0, 9)
    parts = [oracle.mcfield_fourier_quad_wave(*args, np.eye(3)[j], *rest)[0] for j in range(3)]
    whole = oracle.mcfield_fourier_quad_wave(*args, WL, *rest)[0]
    want = sum(WL[j] * parts[j] for j in range(3))
    np.testing.assert_allclose(whole, want, rtol=0, atol=1e-12 * np.abs(want).max())


def _subset(ps, idx):
    """The paths idx of a path set (oracle layout)."""
    off = ps["offset"]
    out = {k: [] for k in ("x1", "px", "w1", "w0p_re", "w0p_im", "x", "ow", "weight", "xi")}
    counts = []
    for p in idx:
        a, b = off[p], off[p + 1]
        counts.append(b - a)
        for k in ("x1", "w1"):
            out[k].append(ps[k][3 * p:3 * p + 3])
        for k in ("px", "w0p_re", "w0p_im"):
            out[k].append(ps[k][p:p + 1])
        for k in ("x", "ow"):
            out[k].append(ps[k][3 * a:3 * b])
        for k in ("weight", "xi"):
            out[k].append(ps[k][a:b])
    r = {k: np.ascontiguousarray(np.concatenate(v)) for k, v in out.items()}
    r["offset"] = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return r


def _paths_on_cell_centres(g, slabs, n_paths, seed):
    """A path set whose first vertices sit exactly on cell centres of the
    Fourier grid (x, y) and on slab centres (z), the positions the Fourier
    scatter quantises to; later vertices anywhere near the box centre."""
    rng = np.random.default_rng(seed)
    Bmin, Bmax = np.array(g.box_min[:]), np.array(g.box_max[:])
    gz = np.concatenate([Bmin[2] + np.arange(g.nz - 1) * g.z_bin, [Bmax[2]]])
    tz = (gz[:-1] + gz[1:]) / 2
    c = (g.nx - 1) // 2
    a_iso = 1 / np.sqrt(4 * np.pi)
    ps = {k: [] for k in ("x1", "px", "w1", "w0p_re", "w0p_im", "x", "ow", "weight", "xi")}
    counts, slab_of = [], []
    for p in range(n_paths):
        j = int(rng.choice(slabs))
        slab_of.append(j)
        ix, iy = c + rng.integers(-1, 2), c + rng.integers(-1, 2)
        x1 = np.array([g.gx0 + ix * g.box_stp, g.gy0 + iy * g.box_stp, tz[j]])
        nvtx = int(rng.integers(1, 6))
        counts.append(nvtx)
        w1 = rng.normal(size=3)
        ps["x1"].append(x1)
        ps["w1"].append(w1 / np.linalg.norm(w1))
        ps["px"].append([rng.uniform(0.5, 2.0)])
        ps["w0p_re"].append([a_iso])
        ps["w0p_im"].append([0.0])
        for n in range(nvtx):
            xn = x1 if n == 0 else np.array([rng.uniform(-1, 1), rng.uniform(-1, 1),
                                             rng.uniform(Bmin[2] + 0.02, Bmax[2] - 0.02)])
            ow = rng.normal(size=3)
            ps["x"].append(xn)
            ps["ow"].append(ow / np.linalg.norm(ow))
            ps["weight"].append([rng.uniform(0.2, 1.0)])
            ps["xi"].append([rng.uniform()])
    r = {k: np.ascontiguousarray(np.concatenate(v)) for k, v in ps.items()}
    r["offset"] = np.concatenate([[0], np.cumsum(counts)]).astype(np.int64)
    return r, np.array(slab_of), tz


@pytest.mark.parametrize("vsign", [1.0, -1.0])
def test_fourier_field_equals_direct_field_on_cell_centres(oracle, vsign):
    """MCfieldFourierQuad against MCfield on one path set: when every first
    vertex lies on a cell centre and a slab centre (where the Fourier scatter
    puts it), the Fourier field is the direct field after the normalisation
    map, with lights and views exchanged (the Fourier sum puts the view phase
    at the first vertex and the light at the current one, MCfield.m's
    reversed-path term; by reciprocity that is MCfield with lights -v and view
    -l):

        u_F(o) = sqrt(maxItr / K) / a_iso * e^{i phi_j(o)} * u_MC,j(o)  summed over slabs j

    sqrt(maxItr/K): the vertex-count normalisation (MCfieldFourierQuad.m:214,
    302-303 vs MCfield.m:252-254); 1/a_iso: the isotropic amplitude the Fourier
    sum leaves out; phi_j: k vz zmax (vsign = +1, viewing towards +z: one
    constant per pixel), or -2 k |vz| t_j (vsign = -1: the reference's
    slab-centre phase -k|vz| t_j (zexp0, :134) has the opposite sign of the
    in-slab phase +k(z1 - t_j) and of MCfield's +k|vz| z1 for a view towards
    -z).  This checks the DFT index folding (fftshift/flip/crop), the lateral
    and slab phases, both attenuations and the normalisation independently of
    a transliteration of the Fourier code."""
    lam, v_max, v_stp = 0.5, 0.3, 0.05
    bmin, bmax = [-2.0, -2.0, 0.0], [2.0, 2.0, 1.2]
    sigt, albedo = 1.5, 0.9
    l = np.array([[0.04], [-0.03], [np.sqrt(1 - 0.04 ** 2 - 0.03 ** 2)]])
    g = oracle.quad_grid(np.array(bmin), np.array(bmax), v_max, v_stp, lam)
    assert g.nfft == g.nv0 == 13 and g.nz - 1 >= 2      # no spatial padding beyond the view grid
    n_paths = 60
    ps, slab_of, tz = _paths_on_cell_centres(g, [0, 1, g.nz - 2], n_paths, seed=3 if vsign > 0 else 4)
    u_F, K = quad_sum(g, ps, sigt, l, v_max, v_stp, lam, False, vsign)
    assert K == ps["offset"][-1]
    # the view grid of MCfieldFourierQuad.m:103-105 and the exchanged problem
    theta_v = -v_max + np.arange(g.nv0) * v_stp
    vx, vy = np.meshgrid(-theta_v, -theta_v, indexing="ij")
    vz = vsign * np.sqrt(1 - vx ** 2 - vy ** 2)
    vdirs = np.vstack([vx.ravel(order="F"), vy.ravel(order="F"), vz.ravel(order="F")])
    Bmin, Bmax = np.array(g.box_min[:]), np.array(g.box_max[:])
    spec = E.make_problem(sigt, albedo, Bmin, Bmax, -vdirs, -l, 1, 1, lam, False, 2, 1, 0, seed=0)
    k = 2 * np.pi / lam
    a_iso = 1 / np.sqrt(4 * np.pi)
    pred = np.zeros((g.nv0, g.nv0), complex)
    for j in np.unique(slab_of):
        u_mc = oracle.mcfield_replay(spec, _subset(ps, np.flatnonzero(slab_of == j)), n_paths)
        u_mc = u_mc.reshape(-1).reshape((g.nv0, g.nv0), order="F")
        if vsign > 0:
            phi = k * vz * Bmax[2]
        else:
            phi = -2 * k * np.abs(vz) * tz[j]
        pred += np.sqrt(n_paths / K) / a_iso * np.exp(1j * phi) * u_mc
    np.testing.assert_allclose(u_F[:, :, 0], pred, rtol=0, atol=1e-11 * np.abs(pred).max())
