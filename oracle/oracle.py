"""ctypes wrapper of the FP64 CPU oracle (oracle/_build/libscmc_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, as the checker.  The product never imports it.
Accepts the same ProblemSpec objects as mcspeckle_amd.engines so a test runs
the GPU engine and the oracle on identical inputs.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libscmc_oracle.so")

_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)


class OPhase(C.Structure):
    _fields_ = [("type", C.c_int), ("g", C.c_double), ("n_amp", C.c_int), ("amp_re", _dp), ("amp_im", _dp),
                ("n_icdf", C.c_int), ("icdf", _dp)]


class OProblem(C.Structure):
    _fields_ = [("dim", C.c_int), ("sigt", C.c_double), ("albedo", C.c_double), ("lam", C.c_double),
                ("box_min", C.c_double * 3), ("box_max", C.c_double * 3), ("nl", C.c_int), ("nv", C.c_int),
                ("l", _dp), ("v", _dp), ("is_ff_l", C.c_int), ("is_ff_v", C.c_int), ("do_cbs", C.c_int),
                ("smp_flg", C.c_int), ("amp", OPhase), ("has_amp0", C.c_int), ("amp0", OPhase),
                ("has_lmean0", C.c_int), ("lmean0", C.c_double * 3), ("seed", C.c_uint64)]


class OPathSet(C.Structure):
    _fields_ = [("n_paths", C.c_int64), ("n_vertices", C.c_int64), ("offset", _i64p), ("x1", _dp), ("px", _dp),
                ("w1", _dp), ("w0p_re", _dp), ("w0p_im", _dp), ("x", _dp), ("ow", _dp), ("weight", _dp),
                ("xi", _dp)]


class OGrid(C.Structure):
    _fields_ = [("nv", C.c_int), ("nv0", C.c_int), ("nx", C.c_int), ("nz", C.c_int), ("nfft", C.c_int),
                ("nx_pad", C.c_int), ("nva_pad", C.c_int), ("box_stp", C.c_double), ("z_bin", C.c_double),
                ("box_min", C.c_double * 3), ("box_max", C.c_double * 3), ("gx0", C.c_double), ("gy0", C.c_double)]


_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        P = C.POINTER(OProblem)
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32)] * 3
        L.orc_uniform_pair.argtypes = [C.c_uint64, C.c_uint64, C.c_uint32, C.c_uint32, _dp]
        L.orc_cube_dist.argtypes = [_dp, _dp, _dp, _dp, C.c_int]
        L.orc_cube_dist.restype = C.c_double
        L.orc_eval_phase_att.argtypes = [_dp, _dp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                         _dp, _dp, _dp, _dp]
        L.orc_eval_amp.argtypes = [C.c_double, C.POINTER(OPhase), C.c_int, _dp, _dp]
        L.orc_evaluate_hg.argtypes = [C.c_double, C.c_double, C.c_int]
        L.orc_evaluate_hg.restype = C.c_double
        L.orc_sample_hg.argtypes = [C.c_double, C.c_int, C.c_double]
        L.orc_sample_hg.restype = C.c_double
        L.orc_rotate_by_theta.argtypes = [_dp, C.c_double, C.c_int, C.c_double, _dp]
        L.orc_sample_dir.argtypes = [_dp, C.POINTER(OPhase), C.c_int, _dp, _dp]
        L.orc_exp_smp_x.argtypes = [_dp, _dp, _dp, C.c_double, C.c_int, _dp, _dp, _dp]
        L.orc_trace_count.argtypes = [P, C.c_uint64, C.c_int64, _i64p]
        L.orc_trace_count.restype = C.c_int64
        L.orc_trace_fill.argtypes = [P, C.c_uint64, C.c_int64, C.POINTER(OPathSet)]
        L.orc_mccov.argtypes = [P, C.c_int64, _dp, _dp, _dp, _dp]
        L.orc_mccov_mt.argtypes = [P, C.c_int64, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_mccov_sums.argtypes = [P, C.c_uint64, C.c_int64, C.c_int, _dp, _dp, _dp, _dp]
        L.orc_mccov_replay.argtypes = [P, C.POINTER(OPathSet), C.c_int64, _dp, _dp, _dp, _dp]
        L.orc_mcfield.argtypes = [P, C.c_int64, _dp, _dp]
        L.orc_mcfield_sums.argtypes = [P, C.c_uint64, C.c_int64, C.c_int, _dp, _dp]
        L.orc_mcfield_replay.argtypes = [P, C.POINTER(OPathSet), C.c_int64, _dp, _dp]
        L.orc_quad_grid_setup.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, C.POINTER(OGrid)]
        L.orc_mcfield_fourier_quad.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp, C.c_int, C.c_double,
                                               C.c_double, C.c_int64, C.c_double, C.c_int, C.c_int, _dp,
                                               C.c_double, C.POINTER(OPhase), C.c_uint64, _dp, _dp]
        L.orc_mcfield_fourier_quad.restype = C.c_int64
        L.orc_mcfield_wave.argtypes = [P, _dp, _dp, C.c_int64, _dp, _dp]
        L.orc_quad_grid_setup_wave.argtypes = [_dp, _dp, C.c_double, C.c_double, C.c_double, C.POINTER(OGrid)]
        L.orc_mcfield_fourier_quad_wave.argtypes = [C.c_double, C.c_double, _dp, _dp, _dp, C.c_int, _dp, _dp,
                                                    C.c_double, C.c_double, C.c_int64, C.c_double, C.c_int,
                                                    C.c_int, _dp, C.c_double, C.POINTER(OPhase), C.c_uint64, _dp,
                                                    _dp]
        L.orc_mcfield_fourier_quad_wave.restype = C.c_int64
        L.orc_tabulated_scatter.argtypes = [_dp, _dp, _dp, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]
        L.orc_inv_cdf.argtypes = [_dp, C.c_int, C.c_int, _dp]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_dp)


def _phase(ph, keep):
    o = OPhase()
    if ph is None:
        return o
    o.type, o.g = ph.type, ph.g
    if ph.type == 2:
        re = np.ascontiguousarray(ph.eval_amp.real, dtype=np.float64)
        im = np.ascontiguousarray(ph.eval_amp.imag, dtype=np.float64)
        ic = np.ascontiguousarray(ph.sample_icdf, dtype=np.float64)
        keep += [re, im, ic]
        o.n_amp, o.amp_re, o.amp_im = re.size, _p(re), _p(im)
        o.n_icdf, o.icdf = ic.size, _p(ic)
    return o


def problem(spec):
    """OProblem from a mcspeckle_amd.engines.ProblemSpec; returns (struct, keepalive)."""
    keep = []
    p = OProblem()
    p.dim, p.sigt, p.albedo, p.lam = spec.dim, spec.sigt, spec.albedo, spec.lam
    for j in range(spec.dim):
        p.box_min[j], p.box_max[j] = spec.box_min[j], spec.box_max[j]
    lc, vc = np.ascontiguousarray(spec.l.T), np.ascontiguousarray(spec.v.T)
    keep += [lc, vc]
    p.nl, p.nv = spec.l.shape[1], spec.v.shape[1]
    p.l = _p(lc)
    p.v = _p(vc) if vc.size else None
    p.is_ff_l, p.is_ff_v, p.do_cbs, p.smp_flg = int(spec.is_ff_l), int(spec.is_ff_v), int(spec.do_cbs), spec.smp_flg
    p.amp = _phase(spec.amp, keep)
    p.has_amp0 = int(spec.amp0 is not None)
    p.amp0 = _phase(spec.amp0, keep)
    if spec.lmean0 is not None:
        p.has_lmean0 = 1
        for j in range(spec.dim):
            p.lmean0[j] = spec.lmean0[j]
    p.seed = spec.seed
    return p, keep


def _offset_spec(spec):
    # the oracle traces paths [p0, p0+n) with p0 passed explicitly
    return spec.path_offset


def trace(spec, n_paths):
    L = lib()
    p, keep = problem(spec)
    p0 = _offset_spec(spec)
    counts = np.empty(n_paths, dtype=np.int64)
    total = L.orc_trace_count(C.byref(p), p0, n_paths, counts.ctypes.data_as(_i64p))
    off = np.zeros(n_paths + 1, dtype=np.int64)
    off[1:] = np.cumsum(counts)
    ps = {"offset": off, "x1": np.empty(3 * n_paths), "px": np.empty(n_paths), "w1": np.empty(3 * n_paths),
          "w0p_re": np.empty(n_paths), "w0p_im": np.empty(n_paths), "x": np.empty(3 * total),
          "ow": np.empty(3 * total), "weight": np.empty(total), "xi": np.empty(total)}
    c = _pathset(ps)
    L.orc_trace_fill(C.byref(p), p0, n_paths, C.byref(c))
    return ps


def _pathset(ps):
    c = OPathSet()
    c.n_paths = ps["offset"].size - 1
    c.n_vertices = int(ps["offset"][-1])
    c.offset = ps["offset"].ctypes.data_as(_i64p)
    for k in ("x1", "px", "w1", "w0p_re", "w0p_im", "x", "ow", "weight", "xi"):
        setattr(c, k, _p(ps[k]))
    return c


def mccov(spec, max_itr, nthreads=1):
    """code/MCcov.m restated, paths [0, max_itr) of spec.seed (path_offset must be 0);
    nthreads > 1 sums the paths in OpenMP shards."""
    L = lib()
    p, keep = problem(spec)
    ne = (spec.nv * spec.nl) ** 2
    a = [np.empty(ne) for _ in range(4)]
    if nthreads > 1:
        L.orc_mccov_mt(C.byref(p), int(max_itr), int(nthreads), *[_p(x) for x in a])
    else:
        L.orc_mccov(C.byref(p), int(max_itr), *[_p(x) for x in a])
    shape = (spec.nv, spec.nv, spec.nl, spec.nl)
    return (a[0] + 1j * a[1]).reshape(shape, order="F"), (a[2] + 1j * a[3]).reshape(shape, order="F")


def mccov_sums(spec, p0, n_paths, nthreads=1):
    """Unnormalised MCcov sums over paths [p0, p0+n) (CPU baseline timing)."""
    L = lib()
    p, keep = problem(spec)
    ne = (spec.nv * spec.nl) ** 2
    a = [np.empty(ne) for _ in range(4)]
    L.orc_mccov_sums(C.byref(p), int(p0), int(n_paths), int(nthreads), *[_p(x) for x in a])
    shape = (spec.nv, spec.nv, spec.nl, spec.nl)
    return (a[0] + 1j * a[1]).reshape(shape, order="F"), (a[2] + 1j * a[3]).reshape(shape, order="F")


def mccov_replay(spec, ps, max_itr):
    L = lib()
    p, keep = problem(spec)
    ne = (spec.nv * spec.nl) ** 2
    a = [np.empty(ne) for _ in range(4)]
    c = _pathset(ps)
    L.orc_mccov_replay(C.byref(p), C.byref(c), int(max_itr), *[_p(x) for x in a])
    shape = (spec.nv, spec.nv, spec.nl, spec.nl)
    return (a[0] + 1j * a[1]).reshape(shape, order="F"), (a[2] + 1j * a[3]).reshape(shape, order="F")


def mcfield(spec, max_itr):
    L = lib()
    p, keep = problem(spec)
    re, im = np.empty(spec.nv * spec.nl), np.empty(spec.nv * spec.nl)
    L.orc_mcfield(C.byref(p), int(max_itr), _p(re), _p(im))
    return (re + 1j * im).reshape((spec.nv, spec.nl), order="F")


def mcfield_sums(spec, p0, n_paths, nthreads=1):
    L = lib()
    p, keep = problem(spec)
    re, im = np.empty(spec.nv * spec.nl), np.empty(spec.nv * spec.nl)
    L.orc_mcfield_sums(C.byref(p), int(p0), int(n_paths), int(nthreads), _p(re), _p(im))
    return (re + 1j * im).reshape((spec.nv, spec.nl), order="F")


def mcfield_replay(spec, ps, max_itr):
    L = lib()
    p, keep = problem(spec)
    re, im = np.empty(spec.nv * spec.nl), np.empty(spec.nv * spec.nl)
    c = _pathset(ps)
    L.orc_mcfield_replay(C.byref(p), C.byref(c), int(max_itr), _p(re), _p(im))
    return (re + 1j * im).reshape((spec.nv, spec.nl), order="F")


def mcfield_wave(spec, wl, max_itr):
    """MCfieldOnWave: the field (nv,) of the wave sum_l wl[l] * light l."""
    L = lib()
    p, keep = problem(spec)
    wl = np.asarray(wl, dtype=complex).reshape(-1)
    wr, wi = np.ascontiguousarray(wl.real), np.ascontiguousarray(wl.imag)
    re, im = np.empty(spec.nv), np.empty(spec.nv)
    L.orc_mcfield_wave(C.byref(p), _p(wr), _p(wi), int(max_itr), _p(re), _p(im))
    return re + 1j * im


def quad_grid(box_min, box_max, v_max, v_stp, lam, wave=False):
    g = OGrid()
    bmin = np.ascontiguousarray(box_min, dtype=np.float64)
    bmax = np.ascontiguousarray(box_max, dtype=np.float64)
    f = lib().orc_quad_grid_setup_wave if wave else lib().orc_quad_grid_setup
    f(_p(bmin), _p(bmax), v_max, v_stp, lam, C.byref(g))
    return g


def _quad_phase(amp, keep):
    """The walk's phase function (a PhaseSpec of mcspeckle_amd.engines, or
    None for the reference's isotropic walk) as a pointer for the quad calls."""
    if amp is None:
        return None
    o = _phase(amp, keep)
    keep.append(o)
    return C.pointer(o)


def mcfield_fourier_quad_wave(sigt, albedo, box_min, box_max, l, wl, v_max, v_stp, max_itr, lam, do_cbs, smp_flg,
                              lmean0=None, vsign=-1.0, seed=0, amp=None):
    """MCfieldFourierQuadOnWave: (nv0, nv0) field of the wave sum_l wl[l] * light l."""
    L = lib()
    keep = []
    ph = _quad_phase(amp, keep)
    g = quad_grid(box_min, box_max, v_max, v_stp, lam, wave=True)
    l = np.asarray(l, dtype=np.float64).reshape(3, -1)
    nl = l.shape[1]
    lc = np.ascontiguousarray(l.T)
    wl = np.asarray(wl, dtype=complex).reshape(-1)
    wr, wi = np.ascontiguousarray(wl.real), np.ascontiguousarray(wl.imag)
    bmin = np.ascontiguousarray(box_min, dtype=np.float64)
    bmax = np.ascontiguousarray(box_max, dtype=np.float64)
    lm = None if lmean0 is None else np.ascontiguousarray(lmean0, dtype=np.float64)
    n = g.nv0 * g.nv0
    re, im = np.empty(n), np.empty(n)
    nstore = L.orc_mcfield_fourier_quad_wave(sigt, albedo, _p(bmin), _p(bmax), _p(lc), nl, _p(wr), _p(wi), v_max,
                                             v_stp, int(max_itr), lam, int(do_cbs), int(smp_flg),
                                             _p(lm) if lm is not None else None, float(vsign), ph, int(seed),
                                             _p(re), _p(im))
    return (re + 1j * im).reshape((g.nv0, g.nv0), order="F"), nstore


def mcfield_fourier_quad(sigt, albedo, box_min, box_max, l, v_max, v_stp, max_itr, lam, do_cbs, smp_flg,
                         lmean0=None, vsign=-1.0, seed=0, amp=None):
    L = lib()
    keep = []
    ph = _quad_phase(amp, keep)
    g = quad_grid(box_min, box_max, v_max, v_stp, lam)
    l = np.asarray(l, dtype=np.float64).reshape(3, -1)
    nl = l.shape[1]
    lc = np.ascontiguousarray(l.T)
    bmin = np.ascontiguousarray(box_min, dtype=np.float64)
    bmax = np.ascontiguousarray(box_max, dtype=np.float64)
    lm = None if lmean0 is None else np.ascontiguousarray(lmean0, dtype=np.float64)
    n = g.nv0 * g.nv0 * nl
    re, im = np.empty(n), np.empty(n)
    nstore = L.orc_mcfield_fourier_quad(sigt, albedo, _p(bmin), _p(bmax), _p(lc), nl, v_max, v_stp, int(max_itr),
                                        lam, int(do_cbs), int(smp_flg), _p(lm) if lm is not None else None,
                                        float(vsign), ph, int(seed), _p(re), _p(im))
    return (re + 1j * im).reshape((g.nv0, g.nv0, nl), order="F"), nstore


def philox(ctr, key):
    L = lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.orc_philox4x32_10(c, k, o)
    return list(o)


def uniform_pair(seed, path, vertex, group):
    o = np.empty(2)
    lib().orc_uniform_pair(seed, path, vertex, group, _p(o))
    return o


def tabulated_scatter(theta, f):
    L = lib()
    theta = np.ascontiguousarray(theta, dtype=np.float64)
    fr = np.ascontiguousarray(np.real(f), dtype=np.float64)
    fi = np.ascontiguousarray(np.imag(f), dtype=np.float64)
    n = theta.size
    cs = C.c_double()
    out = {k: np.empty(n) for k in ("samplePdf", "sampleCdf", "sampleIcdf", "evalPdf", "re", "im")}
    D = L.orc_tabulated_scatter(_p(theta), _p(fr), _p(fi), n, C.byref(cs), _p(out["samplePdf"]),
                                _p(out["sampleCdf"]), _p(out["sampleIcdf"]), _p(out["evalPdf"]), _p(out["re"]),
                                _p(out["im"]))
    out["D"] = D
    out["cs"] = cs.value
    out["evalAmp"] = out.pop("re") + 1j * out.pop("im")
    return out


# ------------------------------------------------ heterogeneous media

class OMedium(C.Structure):
    _fields_ = [("nb", C.c_int * 3), ("bin", C.c_double * 3), ("sigt", _dp), ("albedo", _dp),
                ("has_lrad", C.c_int), ("lrad", C.c_double)]


def medium(spec, sigt, box_bin, albedo=None, lrad=None):
    """orc_medium for sigt (an array of cells, x fastest like MATLAB) over box_bin cells."""
    dim = spec.dim
    sig = np.ascontiguousarray(np.asarray(sigt, dtype=np.float64).reshape(-1, order="F"))
    bb = np.asarray(box_bin, dtype=np.float64).reshape(-1)
    nb = np.rint((spec.box_max - spec.box_min) / bb).astype(int)
    assert nb.prod() == sig.size, (nb, sig.size)
    m = OMedium()
    for j in range(dim):
        m.nb[j] = int(nb[j])
        m.bin[j] = float(bb[j])
    for j in range(dim, 3):
        m.nb[j] = 1
        m.bin[j] = 1.0
    m.sigt = _p(sig)
    alb = None
    if albedo is not None and np.size(albedo) > 1:
        alb = np.ascontiguousarray(np.asarray(albedo, dtype=np.float64).reshape(-1, order="F"))
        m.albedo = _p(alb)
    else:
        m.albedo = C.cast(None, _dp)
    m.has_lrad = 0 if lrad is None else 1
    m.lrad = 0.0 if lrad is None else float(lrad)
    return m, (sig, alb)


def _hetro_sigs(L):
    MP, P = C.POINTER(OMedium), C.POINTER(OProblem)
    L.orc_mccov_hetro.argtypes = [P, MP, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_mccov_hetro.restype = C.c_int64
    L.orc_mccov_hetro_sums.argtypes = [P, MP, C.c_uint64, C.c_int64, _dp, _dp, _dp, _dp, _dp, _dp]
    L.orc_mccov_hetro_sums.restype = C.c_int64
    L.orc_eval_phase_att_hetro.argtypes = [P, MP, _dp, _dp, C.c_int, C.c_int, C.c_double, _dp, _dp]
    L.orc_woodcock.argtypes = [P, MP, _dp, _dp, C.c_uint64]
    L.orc_woodcock.restype = C.c_double
    L.orc_hetro_icdf.argtypes = [_dp, C.c_int, C.c_int, C.POINTER(C.c_int32)]
    L.orc_hetro_icdf.restype = C.c_int
    L.orc_mcfield_hetro.argtypes = [P, MP, C.c_int64, _dp, _dp]
    L.orc_mcfield_hetro.restype = C.c_int64
    L.orc_mcfield_hetro_sums.argtypes = [P, MP, C.c_uint64, C.c_int64, _dp, _dp]
    L.orc_mcfield_hetro_sums.restype = C.c_int64


def mccov_hetro(spec, sigt, box_bin, max_itr, albedo=None, lrad=None, sums_from=None):
    """MCCovHetro: (Ms, Mm, mean1, vertex count).  With sums_from=p0 the
    unnormalised sums over path ids [p0, p0 + max_itr)."""
    L = lib()
    _hetro_sigs(L)
    p, keep = problem(spec)
    m, keep2 = medium(spec, sigt, box_bin, albedo, lrad)
    nv, nl = spec.nv, spec.nl
    ne, nf = nv * nv * nl * nl, nv * nl
    a, b, c, d = np.empty(ne), np.empty(ne), np.empty(ne), np.empty(ne)
    e, f = np.empty(nf), np.empty(nf)
    if sums_from is None:
        k = L.orc_mccov_hetro(C.byref(p), C.byref(m), int(max_itr), _p(a), _p(b), _p(c), _p(d), _p(e), _p(f))
    else:
        k = L.orc_mccov_hetro_sums(C.byref(p), C.byref(m), int(sums_from), int(max_itr), _p(a), _p(b), _p(c), _p(d),
                                   _p(e), _p(f))
    sh = (nv, nv, nl, nl)
    return ((a + 1j * b).reshape(sh, order="F"), (c + 1j * d).reshape(sh, order="F"),
            (e + 1j * f).reshape((nv, nl), order="F"), k)


def mcfield_hetro(spec, sigt, box_bin, max_itr, albedo=None, sums_from=None):
    """MCsampleHetro: (u nv x nl, vertex count).  With sums_from=p0 the
    unnormalised sum over path ids [p0, p0 + max_itr)."""
    L = lib()
    _hetro_sigs(L)
    p, keep = problem(spec)
    m, keep2 = medium(spec, sigt, box_bin, albedo)
    nv, nl = spec.nv, spec.nl
    re, im = np.empty(nv * nl), np.empty(nv * nl)
    if sums_from is None:
        k = L.orc_mcfield_hetro(C.byref(p), C.byref(m), int(max_itr), _p(re), _p(im))
    else:
        k = L.orc_mcfield_hetro_sums(C.byref(p), C.byref(m), int(sums_from), int(max_itr), _p(re), _p(im))
    if k < 0:
        raise ValueError("invalid medium")
    return (re + 1j * im).reshape((nv, nl), order="F"), k


def phase_att_hetro(spec, sigt, box_bin, x, dirs, is_ff, sign=1.0, albedo=None):
    L = lib()
    _hetro_sigs(L)
    p, keep = problem(spec)
    m, keep2 = medium(spec, sigt, box_bin, albedo)
    dirs = np.asarray(dirs, dtype=np.float64).reshape(spec.dim, -1)
    dc = np.ascontiguousarray(dirs.T)
    xx = np.ascontiguousarray(np.r_[np.asarray(x, dtype=np.float64).reshape(-1), np.zeros(3)][:3])
    n = dirs.shape[1]
    re, im = np.empty(n), np.empty(n)
    L.orc_eval_phase_att_hetro(C.byref(p), C.byref(m), _p(xx), _p(dc), n, int(is_ff), float(sign), _p(re), _p(im))
    return re + 1j * im


def woodcock(spec, sigt, box_bin, x, w, paths):
    L = lib()
    _hetro_sigs(L)
    p, keep = problem(spec)
    m, keep2 = medium(spec, sigt, box_bin)
    xx = np.ascontiguousarray(np.r_[np.asarray(x, float).reshape(-1), np.zeros(3)][:3])
    ww = np.ascontiguousarray(np.r_[np.asarray(w, float).reshape(-1), np.zeros(3)][:3])
    return np.array([L.orc_woodcock(C.byref(p), C.byref(m), _p(xx), _p(ww), int(q)) for q in paths])


def hetro_icdf(sigt, n_out):
    L = lib()
    _hetro_sigs(L)
    sig = np.ascontiguousarray(np.asarray(sigt, dtype=np.float64).reshape(-1, order="F"))
    out = np.empty(n_out, dtype=np.int32)
    L.orc_hetro_icdf(_p(sig), sig.size, n_out, out.ctypes.data_as(C.POINTER(C.c_int32)))
    return out
