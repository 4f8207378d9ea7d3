"""The speckle Monte Carlo engines with the reference's positional-argument
conventions, running on the MI355X through libscmc.so.

    [Ms, Mm] = MCcov(sigt, albedo, box_min, box_max, l, v, is_ff_l, is_ff_v,
                     maxItr, lambda, doCBS, smpFlg, sct_type, ampfunc[, ampfunc0])
                                                      (code/MCcov.m:1, :45-70)
    u = MCfield(...same..., [ampfunc0[, lmean0]])     (code/MCfield.m:1, :41-66)
    u = MCfieldFourierQuad(sigt, albedo, box_min, box_max, l, v_max, v_stp,
                           maxItr, lambda, doCBS, smpFlg, lmean0, vsign)
                                                      (code/MCfieldFourierQuad.m:1)
    [Ms, Mm, mean1] = MCCovHetro(sigt, albedo, box_min, box_max, box_bin, l, v,
                                 is_ff_l, is_ff_v, maxItr, lambda, doCBS, smpFlg,
                                 sct_type, ampfunc[, ampfunc0[, lrad]])
                                                      (code/MCCovHetro.m:1)
    u = MCfieldOnWave(...MCfield's 14..., ampfunc0, lmean0, Wl)
                                                      (code/MCfieldOnWave.m:1)
    u = MCfieldFourierQuadOnWave(sigt, albedo, box_min, box_max, l, v_max,
                                 v_stp, maxItr, lambda, doCBS, smpFlg, lmean0,
                                 Wl, vsign)           (code/MCfieldFourierQuadOnWave.m:1)

Differences from MATLAB are keyword-only: ``seed`` and ``path_offset``
(Philox key and first path id; MATLAB uses the global ``rng``), ``precision``
("fp64" or "mixed") and ``stats`` (a dict receiving timing and counts).
Outputs are numpy complex128 arrays in MATLAB's column-major layout
(``Ms[v1, v2, l1, l2]``, ``u[v, l]``).
"""
from dataclasses import dataclass, field
from typing import Optional

import atexit
import ctypes as C

import numpy as np

from . import _native as N


class ScmcArgumentError(ValueError):
    """Raised with the reference's error message for invalid arguments."""


@dataclass
class PhaseSpec:
    type: int                       # 1 isotropic, 2 tabulated, 3 HG
    g: float = 0.0
    eval_amp: Optional[np.ndarray] = None      # complex, tabulated
    sample_icdf: Optional[np.ndarray] = None   # real, tabulated


@dataclass
class ProblemSpec:
    dim: int
    sigt: float
    albedo: float
    lam: float
    box_min: np.ndarray
    box_max: np.ndarray
    l: np.ndarray                   # dim x nl
    v: np.ndarray                   # dim x nv (may be empty for the quad engine)
    is_ff_l: bool
    is_ff_v: bool
    do_cbs: bool
    smp_flg: int
    amp: PhaseSpec
    amp0: Optional[PhaseSpec] = None
    lmean0: Optional[np.ndarray] = None
    seed: int = 0
    path_offset: int = 0
    _keep: list = field(default_factory=list, repr=False)

    @property
    def nl(self):
        return self.l.shape[1]

    @property
    def nv(self):
        return self.v.shape[1]


def _phase_from_arg(sct_type, ampfunc, name):
    if sct_type == 1:
        return PhaseSpec(1)
    if sct_type == 3:
        g = float(np.asarray(ampfunc).reshape(-1)[0])
        return PhaseSpec(3, g=g)
    if sct_type == 2:
        try:
            ea = ampfunc["evalAmp"]
            ic = ampfunc["sampleIcdf"]
        except (TypeError, KeyError):
            ea = getattr(ampfunc, "evalAmp", None)
            ic = getattr(ampfunc, "sampleIcdf", None)
        if ea is None or ic is None:
            raise ScmcArgumentError(f"{name} must hold evalAmp and sampleIcdf for a tabulated scatterer")
        return PhaseSpec(2, eval_amp=np.asarray(ea, dtype=np.complex128).reshape(-1),
                         sample_icdf=np.asarray(ic, dtype=np.float64).reshape(-1))
    raise ScmcArgumentError("sct_type must be 1 (isotropic), 2 (tabulated) or 3 (HG)")


def _sources(x, dim, what):
    """MCcov.m:56-70: dim x N, or a 1 x N row of angles in 2-D."""
    a = np.asarray(x, dtype=np.float64)
    if a.ndim == 1:
        a = a.reshape(1, -1)
    if a.shape[0] != dim:
        if dim == 2 and a.shape[0] == 1:
            a = np.vstack([np.sin(a[0]), np.cos(a[0])])
        else:
            raise ScmcArgumentError(f"Invalid {what} source input")
    return np.ascontiguousarray(a)


def make_problem(sigt, albedo, box_min, box_max, l, v, is_ff_l, is_ff_v, lambda_, doCBS, smpFlg, sct_type, ampfunc,
                 ampfunc0=None, lmean0=None, seed=0, path_offset=0, need_views=True):
    bmin = np.asarray(box_min, dtype=np.float64).reshape(-1)
    bmax = np.asarray(box_max, dtype=np.float64).reshape(-1)
    # MCcov.m:48-51
    if bmax.size not in (2, 3) or bmax.size != bmin.size:
        raise ScmcArgumentError("Invalid box size")
    dim = bmax.size
    L = _sources(l, dim, "light")
    V = _sources(v, dim, "view") if need_views else np.zeros((dim, 0))
    if smpFlg not in (1, 2):
        raise ScmcArgumentError("smpFlg must be 1 or 2")
    amp = _phase_from_arg(int(sct_type), ampfunc, "ampfunc")
    amp0 = _phase_from_arg(int(sct_type), ampfunc0, "ampfunc0") if ampfunc0 is not None else None
    lm = None
    if lmean0 is not None and np.asarray(lmean0).size:
        lm = np.asarray(lmean0, dtype=np.float64).reshape(-1)
    return ProblemSpec(dim=dim, sigt=float(sigt), albedo=float(albedo), lam=float(lambda_), box_min=bmin,
                       box_max=bmax, l=L, v=V, is_ff_l=bool(is_ff_l), is_ff_v=bool(is_ff_v), do_cbs=bool(doCBS),
                       smp_flg=int(smpFlg), amp=amp, amp0=amp0, lmean0=lm, seed=int(seed),
                       path_offset=int(path_offset))


def _c_phase(ph, keep):
    out = N.Phase()
    if ph is None:
        return out
    out.type = ph.type
    out.g = ph.g
    if ph.type == 2:
        re = np.ascontiguousarray(ph.eval_amp.real, dtype=np.float64)
        im = np.ascontiguousarray(ph.eval_amp.imag, dtype=np.float64)
        ic = np.ascontiguousarray(ph.sample_icdf, dtype=np.float64)
        keep += [re, im, ic]
        out.n_amp = re.size
        out.amp_re = N.dptr(re)
        out.amp_im = N.dptr(im)
        out.n_icdf = ic.size
        out.icdf = N.dptr(ic)
    return out


def c_problem(spec):
    """ctypes scmc_problem for a ProblemSpec (arrays kept alive on spec)."""
    keep = []
    p = N.Problem()
    p.dim = spec.dim
    p.sigt, p.albedo, p.lam = spec.sigt, spec.albedo, spec.lam
    for j in range(spec.dim):
        p.box_min[j] = spec.box_min[j]
        p.box_max[j] = spec.box_max[j]
    lc = np.ascontiguousarray(spec.l.T)   # dim x nl column-major
    vc = np.ascontiguousarray(spec.v.T)
    keep += [lc, vc]
    p.nl, p.nv = spec.l.shape[1], spec.v.shape[1]
    p.l = N.dptr(lc)
    p.v = N.dptr(vc) if vc.size else None
    p.is_ff_l, p.is_ff_v = int(spec.is_ff_l), int(spec.is_ff_v)
    p.do_cbs, p.smp_flg = int(spec.do_cbs), spec.smp_flg
    p.amp = _c_phase(spec.amp, keep)
    p.has_amp0 = int(spec.amp0 is not None)
    p.amp0 = _c_phase(spec.amp0, keep)
    if spec.lmean0 is not None:
        p.has_lmean0 = 1
        for j in range(spec.dim):
            p.lmean0[j] = spec.lmean0[j]
    p.seed = spec.seed
    p.path_offset = spec.path_offset
    spec._keep = keep
    return p


def _cout(n):
    """A complex output of n entries; _cp(a) passes it as the real pointer
    with a NULL imaginary pointer, for which the C-ABI writes interleaved
    (re, im) pairs (no separate real and imaginary arrays to combine)."""
    return np.empty(n, dtype=np.complex128)


def _cp(a):
    return N.dptr(a.view(np.float64))


def _prec(precision):
    if isinstance(precision, str):
        if precision not in N.PRECISIONS:
            raise ScmcArgumentError(f"precision must be one of {sorted(N.PRECISIONS)}")
        return N.PRECISIONS[precision]
    return int(precision)


def mccov_spec(spec, max_itr, precision="fp64", stats=None):
    """Run MCcov for a ProblemSpec; returns (Ms, Mm) shaped (nv, nv, nl, nl)."""
    L = N.lib()
    nv, nl = spec.nv, spec.nl
    ne = (nv * nl) ** 2
    ms, mm = _cout(ne), _cout(ne)
    st = N.Stats()
    p = c_problem(spec)
    N.check(L.scmc_mccov(C.byref(p), int(max_itr), _prec(precision), _cp(ms), None, _cp(mm), None, C.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    shape = (nv, nv, nl, nl)
    return ms.reshape(shape, order="F"), mm.reshape(shape, order="F")


def mcfield_spec(spec, max_itr, precision="fp64", stats=None):
    L = N.lib()
    nv, nl = spec.nv, spec.nl
    u = _cout(nv * nl)
    st = N.Stats()
    p = c_problem(spec)
    N.check(L.scmc_mcfield(C.byref(p), int(max_itr), _prec(precision), _cp(u), None, C.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    return u.reshape((nv, nl), order="F")


def mcfield_noalloc_spec(spec, max_itr, precision="fp64", stats=None):
    """mcfield_spec through scmc_mcfield_noalloc: the same field, with every
    device buffer in the library's fixed scratch array (no device allocation)."""
    L = N.lib()
    nv, nl = spec.nv, spec.nl
    u = _cout(nv * nl)
    st = N.Stats()
    p = c_problem(spec)
    N.check(L.scmc_mcfield_noalloc(C.byref(p), int(max_itr), _prec(precision), _cp(u), None, C.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    return u.reshape((nv, nl), order="F")


def MCcov(sigt, albedo, box_min, box_max, l, v, is_ff_l, is_ff_v, maxItr, lambda_, doCBS, smpFlg, sct_type,
          ampfunc, ampfunc0=None, *, seed=0, path_offset=0, precision="fp64", stats=None):
    """code/MCcov.m:1-299 on the GPU. Returns (Ms, Mm)."""
    spec = make_problem(sigt, albedo, box_min, box_max, l, v, is_ff_l, is_ff_v, lambda_, doCBS, smpFlg, sct_type,
                        ampfunc, ampfunc0, seed=seed, path_offset=path_offset)
    if not (np.isscalar(maxItr) and maxItr >= 1):
        raise ScmcArgumentError("Invalid photons number")
    return mccov_spec(spec, int(maxItr), precision, stats)


def MCfield(sigt, albedo, box_min, box_max, l, v, is_ff_l, is_ff_v, maxItr, lambda_, doCBS, smpFlg, sct_type,
            ampfunc, ampfunc0=None, lmean0=None, *, seed=0, path_offset=0, precision="fp64", stats=None):
    """code/MCfield.m:1-254 on the GPU. Returns u (nv x nl)."""
    spec = make_problem(sigt, albedo, box_min, box_max, l, v, is_ff_l, is_ff_v, lambda_, doCBS, smpFlg, sct_type,
                        ampfunc, ampfunc0, lmean0=lmean0, seed=seed, path_offset=path_offset)
    if not (np.isscalar(maxItr) and maxItr >= 1):
        raise ScmcArgumentError("Invalid photons number")
    return mcfield_spec(spec, int(maxItr), precision, stats)


def _weights(Wl, nl):
    w = np.asarray(Wl, dtype=complex).reshape(-1)
    if w.size != nl:
        raise ScmcArgumentError(f"Wl must have one weight per light ({nl}), got {w.size}")
    return np.ascontiguousarray(w.real), np.ascontiguousarray(w.imag)


def MCfieldOnWave(sigt, albedo, box_min, box_max, l, v, is_ff_l, is_ff_v, maxItr, lambda_, doCBS, smpFlg, sct_type,
                  ampfunc, ampfunc0=None, lmean0=None, Wl=None, *, seed=0, path_offset=0, precision="fp64",
                  stats=None):
    """code/MCfieldOnWave.m:1-267 on the GPU: the field (nv,) of the
    illumination wave sum_l Wl[l] * light l.  ``ampfunc0`` / ``lmean0`` may be
    None (not given), which the reference's positional form cannot express
    when Wl follows them."""
    spec = make_problem(sigt, albedo, box_min, box_max, l, v, is_ff_l, is_ff_v, lambda_, doCBS, smpFlg, sct_type,
                        ampfunc, ampfunc0, lmean0=lmean0, seed=seed, path_offset=path_offset)
    if not (np.isscalar(maxItr) and maxItr >= 1):
        raise ScmcArgumentError("Invalid photons number")
    if Wl is None:
        raise ScmcArgumentError("Wl (one weight per light) is required")
    wr, wi = _weights(Wl, spec.nl)
    u = _cout(spec.nv)
    st = N.Stats()
    p = c_problem(spec)
    N.check(N.lib().scmc_mcfield_wave(C.byref(p), N.dptr(wr), N.dptr(wi), int(maxItr), _prec(precision), _cp(u), None,
                                      C.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    return u


def c_medium(spec, sigt, box_bin, albedo=None, lrad=None):
    """scmc_medium for a sigt grid over box_bin cells (x fastest, MATLAB
    order); albedo is a scalar (spec.albedo) or a grid of the same shape."""
    dim = spec.dim
    sig = np.asarray(sigt, dtype=np.float64)
    bb = np.asarray(box_bin, dtype=np.float64).reshape(-1)
    if bb.size != dim:
        raise ScmcArgumentError("box_bin must have one size per dimension")
    nb = np.rint((spec.box_max - spec.box_min) / bb).astype(int)
    flat = np.ascontiguousarray(sig.reshape(-1, order="F"))
    if flat.size != int(nb.prod()):
        raise ScmcArgumentError(f"sigt has {flat.size} cells, (box_max-box_min)./box_bin gives {tuple(nb)}")
    m = N.Medium()
    for j in range(3):
        m.nb[j] = int(nb[j]) if j < dim else 1
        m.box_bin[j] = float(bb[j]) if j < dim else 1.0
    m.sigt = N.dptr(flat)
    keep = [flat]
    alb = None if albedo is None else np.asarray(albedo, dtype=np.float64)
    if alb is not None and alb.size > 1:
        af = np.ascontiguousarray(alb.reshape(-1, order="F"))
        if af.size != flat.size:
            raise ScmcArgumentError("albedo grid must match sigt")
        m.albedo = N.dptr(af)
        keep.append(af)
    else:
        m.albedo = C.cast(None, C.POINTER(C.c_double))
    m.has_lrad = 0 if lrad is None else 1
    m.lrad = 0.0 if lrad is None else float(lrad)
    return m, keep


def MCCovHetro(sigt, albedo, box_min, box_max, box_bin, l, v, is_ff_l, is_ff_v, maxItr, lambda_, doCBS, smpFlg,
               sct_type, ampfunc, ampfunc0=None, lrad=None, *, seed=0, path_offset=0, precision="fp64", stats=None):
    """code/MCCovHetro.m:1-276 on the GPU: (Ms, Mm, mean1) for a voxel medium
    (sigt and, optionally, albedo grids over box_bin cells)."""
    sig = np.asarray(sigt, dtype=np.float64)
    alb = np.asarray(albedo, dtype=np.float64)
    a0 = float(alb.reshape(-1)[0]) if alb.size == 1 else 1.0
    spec = make_problem(max(float(sig.max()), 1e-300), a0, box_min, box_max, l, v, is_ff_l, is_ff_v, lambda_, doCBS,
                        smpFlg, sct_type, ampfunc, ampfunc0, seed=seed, path_offset=path_offset)
    if not (np.isscalar(maxItr) and maxItr >= 1):
        raise ScmcArgumentError("Invalid photons number")
    m, keep = c_medium(spec, sig, box_bin, alb if alb.size > 1 else None, lrad)
    nv, nl = spec.nv, spec.nl
    ne, nf = nv * nv * nl * nl, nv * nl
    ms, mm, mean1 = _cout(ne), _cout(ne), _cout(nf)
    st = N.Stats()
    p = c_problem(spec)
    N.check(N.lib().scmc_mccov_hetro(C.byref(p), C.byref(m), int(maxItr), _prec(precision), _cp(ms), None, _cp(mm),
                                     None, _cp(mean1), None, C.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    sh = (nv, nv, nl, nl)
    return ms.reshape(sh, order="F"), mm.reshape(sh, order="F"), mean1.reshape((nv, nl), order="F")


def MCsampleHetro(sigt, albedo, box_min, box_max, box_bin, l, v, is_ff_l, is_ff_v, maxItr, lambda_, doCBS, smpFlg,
                  sct_type, ampfunc, ampfunc0=None, lmean0=None, *, seed=0, path_offset=0, precision="fp64",
                  stats=None):
    """code/MCsampleHetro.m:1-227 on the GPU: the field sample u (Nv x Nl) of
    a voxel medium (sigt and, optionally, albedo grids over box_bin cells).
    The debug pairs (dbgpair, corrM) of the reference are not provided."""
    sig = np.asarray(sigt, dtype=np.float64)
    alb = np.asarray(albedo, dtype=np.float64)
    a0 = float(alb.reshape(-1)[0]) if alb.size == 1 else 1.0
    spec = make_problem(max(float(sig.max()), 1e-300), a0, box_min, box_max, l, v, is_ff_l, is_ff_v, lambda_, doCBS,
                        smpFlg, sct_type, ampfunc, ampfunc0, lmean0=lmean0, seed=seed, path_offset=path_offset)
    if not (np.isscalar(maxItr) and maxItr >= 1):
        raise ScmcArgumentError("Invalid photons number")
    m, keep = c_medium(spec, sig, box_bin, alb if alb.size > 1 else None)
    nv, nl = spec.nv, spec.nl
    u = _cout(nv * nl)
    st = N.Stats()
    p = c_problem(spec)
    N.check(N.lib().scmc_mcfield_hetro(C.byref(p), C.byref(m), int(maxItr), _prec(precision), _cp(u), None,
                                       C.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    return u.reshape((nv, nl), order="F")


def quad_grid(box_min, box_max, v_max, v_stp, lambda_, wave=False):
    """Grid sizes (nv0, nx, nz) of MCfieldFourierQuad.m:80-121 (wave:
    MCfieldFourierQuadOnWave.m:84-131)."""
    L = N.lib()
    bmin = np.ascontiguousarray(np.asarray(box_min, dtype=np.float64).reshape(3))
    bmax = np.ascontiguousarray(np.asarray(box_max, dtype=np.float64).reshape(3))
    nv0, nx, nz = C.c_int32(), C.c_int32(), C.c_int32()
    f = L.scmc_quad_grid_wave if wave else L.scmc_quad_grid
    N.check(f(N.dptr(bmin), N.dptr(bmax), float(v_max), float(v_stp), float(lambda_), C.byref(nv0), C.byref(nx),
              C.byref(nz)))
    return nv0.value, nx.value, nz.value


def MCfieldFourierQuad(sigt, albedo, box_min, box_max, l, v_max, v_stp, maxItr, lambda_, doCBS, smpFlg, lmean0=None,
                       vsign=-1, *, sct_type=1, ampfunc=0, seed=0, path_offset=0, precision="mixed", stats=None):
    """code/MCfieldFourierQuad.m:1-303 on the GPU. Returns u (nv0 x nv0 x nl).

    ``sct_type``/``ampfunc`` select the walk's phase function; the default
    (isotropic) is the reference's only case."""
    spec = make_problem(sigt, albedo, box_min, box_max, l, np.zeros((3, 0)), 1, 1, lambda_, doCBS, smpFlg,
                        sct_type, ampfunc, None, lmean0=lmean0, seed=seed, path_offset=path_offset,
                        need_views=False)
    if spec.dim != 3:
        raise ScmcArgumentError("MCfieldFourierQuad is 3-D only")
    nv0, _, _ = quad_grid(spec.box_min, spec.box_max, v_max, v_stp, lambda_)
    nl = spec.nl
    # interleaved complex output (the imaginary pointer NULL)
    u = np.empty(nv0 * nv0 * nl, dtype=np.complex128)
    st = N.Stats()
    p = c_problem(spec)
    N.check(N.lib().scmc_mcfield_fourier_quad(C.byref(p), float(v_max), float(v_stp), int(maxItr), float(vsign),
                                              _prec(precision), N.dptr(u.view(np.float64)), None, C.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    return u.reshape((nv0, nv0, nl), order="F")


def MCfieldFourierQuadOnWave(sigt, albedo, box_min, box_max, l, v_max, v_stp, maxItr, lambda_, doCBS, smpFlg,
                             lmean0=None, Wl=None, vsign=-1, *, sct_type=1, ampfunc=0, seed=0, path_offset=0,
                             precision="mixed", stats=None):
    """code/MCfieldFourierQuadOnWave.m:1-337 on the GPU. Returns u (nv0 x nv0),
    the field of the illumination wave sum_l Wl[l] * light l."""
    spec = make_problem(sigt, albedo, box_min, box_max, l, np.zeros((3, 0)), 1, 1, lambda_, doCBS, smpFlg,
                        sct_type, ampfunc, None, lmean0=lmean0, seed=seed, path_offset=path_offset,
                        need_views=False)
    if spec.dim != 3:
        raise ScmcArgumentError("MCfieldFourierQuadOnWave is 3-D only")
    if Wl is None:
        raise ScmcArgumentError("Wl (one weight per light) is required")
    wr, wi = _weights(Wl, spec.nl)
    nv0, _, _ = quad_grid(spec.box_min, spec.box_max, v_max, v_stp, lambda_, wave=True)
    u = np.empty(nv0 * nv0, dtype=np.complex128)   # interleaved complex output
    st = N.Stats()
    p = c_problem(spec)
    N.check(N.lib().scmc_mcfield_fourier_quad_wave(C.byref(p), N.dptr(wr), N.dptr(wi), float(v_max), float(v_stp),
                                                   int(maxItr), float(vsign), _prec(precision),
                                                   N.dptr(u.view(np.float64)), None, C.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    return u.reshape((nv0, nv0), order="F")


# ------------------------------------------------------------ path sets

def trace(spec, n_paths):
    """Path set sampled on the GPU for paths [path_offset, path_offset+n)."""
    L = N.lib()
    p = c_problem(spec)
    nvert = C.c_int64()
    N.check(L.scmc_trace(C.byref(p), int(n_paths), C.byref(nvert), None, None, None, None, None, None, None, None,
                         None, None))
    K = nvert.value
    ps = {
        "offset": np.empty(n_paths + 1, dtype=np.int64),
        "x1": np.empty(3 * n_paths), "px": np.empty(n_paths), "w1": np.empty(3 * n_paths),
        "w0p_re": np.empty(n_paths), "w0p_im": np.empty(n_paths),
        "x": np.empty(3 * K), "ow": np.empty(3 * K), "weight": np.empty(K), "xi": np.empty(K),
    }
    N.check(L.scmc_trace(C.byref(p), int(n_paths), C.byref(nvert), N.i64ptr(ps["offset"]), N.dptr(ps["x1"]),
                         N.dptr(ps["px"]), N.dptr(ps["w1"]), N.dptr(ps["w0p_re"]), N.dptr(ps["w0p_im"]),
                         N.dptr(ps["x"]), N.dptr(ps["ow"]), N.dptr(ps["weight"]), N.dptr(ps["xi"])))
    return ps


def c_pathset(ps):
    c = N.PathSet()
    c.n_paths = ps["offset"].size - 1
    c.n_vertices = int(ps["offset"][-1])
    c.offset = N.i64ptr(ps["offset"])
    for k in ("x1", "px", "w1", "w0p_re", "w0p_im", "x", "ow", "weight", "xi"):
        setattr(c, k, N.dptr(ps[k]))
    return c


def mccov_replay(spec, ps, max_itr, precision="fp64", stats=None):
    L = N.lib()
    nv, nl = spec.nv, spec.nl
    ne = (nv * nl) ** 2
    ms, mm = _cout(ne), _cout(ne)
    p = c_problem(spec)
    cps = c_pathset(ps)
    st = N.Stats()
    N.check(L.scmc_mccov_replay(C.byref(p), C.byref(cps), int(max_itr), _prec(precision), _cp(ms), None, _cp(mm),
                                None, C.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    shape = (nv, nv, nl, nl)
    return ms.reshape(shape, order="F"), mm.reshape(shape, order="F")


def mcfield_replay(spec, ps, max_itr, precision="fp64", stats=None):
    L = N.lib()
    nv, nl = spec.nv, spec.nl
    u = _cout(nv * nl)
    p = c_problem(spec)
    cps = c_pathset(ps)
    st = N.Stats()
    N.check(L.scmc_mcfield_replay(C.byref(p), C.byref(cps), int(max_itr), _prec(precision), _cp(u), None,
                                  C.byref(st)))
    if stats is not None:
        stats.update(st.as_dict())
    return u.reshape((nv, nl), order="F")


class Session:
    """Device-resident accumulation of one MCcov (kind "cov") or MCfield
    (kind "field") problem over path ranges; see scmc_session_* in scmc.h."""

    def __init__(self, spec, kind="cov", precision="fp64"):
        self.spec = spec
        self.kind = kind
        self._p = c_problem(spec)
        h = C.c_int64()
        N.check(N.lib().scmc_session_create(C.byref(self._p), 0 if kind == "cov" else 1, _prec(precision),
                                            C.byref(h)))
        self.handle = h.value

    def run(self, first, n_paths):
        N.check(N.lib().scmc_session_run(self.handle, int(first), int(n_paths)))

    def stats(self):
        st = N.Stats()
        N.check(N.lib().scmc_session_stats(self.handle, C.byref(st)))
        return st.as_dict()

    def result(self, max_itr):
        nv, nl = self.spec.nv, self.spec.nl
        if self.kind == "cov":
            ne = (nv * nl) ** 2
            ms, mm = _cout(ne), _cout(ne)
            N.check(N.lib().scmc_session_result(self.handle, int(max_itr), _cp(ms), None, _cp(mm), None))
            shape = (nv, nv, nl, nl)
            return ms.reshape(shape, order="F"), mm.reshape(shape, order="F")
        u = _cout(nv * nl)
        N.check(N.lib().scmc_session_result(self.handle, int(max_itr), _cp(u), None, None, None))
        return u.reshape((nv, nl), order="F")

    def close(self):
        if self.handle:
            N.lib().scmc_session_destroy(self.handle)
            self.handle = 0

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _stop_server_at_exit():
    # Python's atexit hooks run before the C runtime's exit handlers, among
    # which HIP tears itself down; a server still running then is stopped here
    if N._lib is not None:
        N._lib.scmc_server_stop()


atexit.register(_stop_server_at_exit)


class Server:
    """Continuous-batching covariance server (scmc_server_* in scmc.h)."""

    def __init__(self, max_active=4, hbm_budget_gb=0.0, paths_per_slice=0, device=0, shared_launches=False,
                 resident_loop=False, resident_workgroups=0):
        cfg = N.ServerConfig()
        cfg.device, cfg.max_active = device, max_active
        cfg.hbm_budget_gb, cfg.paths_per_slice = hbm_budget_gb, paths_per_slice
        cfg.shared_launches = 1 if shared_launches else 0
        cfg.resident_loop = 1 if resident_loop else 0
        cfg.resident_workgroups = int(resident_workgroups)
        N.check(N.lib().scmc_server_start(C.byref(cfg)))
        self._specs = {}

    def submit(self, spec, max_itr, precision="mixed", kind="cov"):
        """kind "cov" (MCcov) or "field" (MCfield)."""
        if kind not in ("cov", "field"):
            raise ScmcArgumentError(f'kind must be "cov" or "field", not {kind!r}')
        p = c_problem(spec)
        h = C.c_int64()
        fn = N.lib().scmc_submit_mccov if kind == "cov" else N.lib().scmc_submit_mcfield
        N.check(fn(C.byref(p), int(max_itr), _prec(precision), C.byref(h)))
        self._specs[h.value] = (spec, kind)
        return h.value

    def _finish(self, rid, block):
        spec, kind = self._specs[rid]
        nv, nl = spec.nv, spec.nl
        ne = (nv * nl) ** 2 if kind == "cov" else nv * nl
        a = [_cout(ne) for _ in range(2 if kind == "cov" else 1)]
        st = N.Stats()
        fn = N.lib().scmc_wait if block else N.lib().scmc_poll
        ptrs = [_cp(a[0]), None] + ([_cp(a[1]), None] if kind == "cov" else [None, None])
        rc = fn(int(rid), *ptrs, C.byref(st))
        if rc == N.SCMC_E_BUSY:
            return None
        self._specs.pop(rid, None)
        N.check(rc)
        if kind != "cov":
            return a[0].reshape((nv, nl), order="F"), st.as_dict()
        shape = (nv, nv, nl, nl)
        return a[0].reshape(shape, order="F"), a[1].reshape(shape, order="F"), st.as_dict()

    def poll(self, rid):
        """(Ms, Mm, stats) -- (u, stats) for a field request -- when finished,
        None while running."""
        return self._finish(rid, False)

    def wait(self, rid):
        return self._finish(rid, True)

    def cancel(self, rid):
        N.check(N.lib().scmc_cancel(int(rid)))

    def stop(self):
        N.lib().scmc_server_stop()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.stop()


def request_footprint(nv, nl, precision="mixed"):
    return int(N.lib().scmc_request_footprint(int(nv), int(nl), _prec(precision)))
