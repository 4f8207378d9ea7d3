"""scmc() front end: the reference's dispatcher over the GPU engines.

Restates scmc.m:1-427: struct validation (:59-127), options (:137-235), the
mapping to positional engine arguments (:245-323), covariance (:326-378) and
field (:381-427) runs with the parfor repeat semantics.  MATLAB's global
``rng`` becomes the Philox key: ``rng`` sets the seed; without it each call
draws the next seed of a process-wide counter, so repeated calls differ as
they do in MATLAB.  ``parforIters`` repeats use disjoint path ranges of one
key, so their average equals one run over all paths.
"""
import copy

import numpy as np

from . import engines as E
from .engines import ScmcArgumentError

_next_seed = [0x5eed]

_BOOL_OPTS = {
    "CBS": ("CBS", "CBS must be logical scalar"),
    "uniformFirstScatter": ("uniformFirstScatter", "uniformFirstScatter must be logical scalar"),
    "renderCov": ("render.cov", "renderCov must be logical scalar"),
    "singleScattering": ("render.singleScattering", "Single scattering option must be logical scalar"),
    "multipleScattering": ("render.multipleScattering", "Multiple scattering option must be logical scalar"),
    "renderField": ("render.field", "renderField must be logical scalar"),
}


def _set(cfg, dotted, val):
    d = cfg
    ks = dotted.split(".")
    for k in ks[:-1]:
        d = d[k]
    d[ks[-1]] = val


def scmc(targetArea, views=None, lights=None, scatter=None, iterations=None, *varargin, precision="fp64", **kw):
    """mulRes = scmc(targetArea, views, lights, scatter, iterations, 'opt', val, ...)
    or scmc(Config).  Options may also be passed as keywords."""
    if views is None:
        Config = copy.deepcopy(targetArea)          # scmc.m:236-239
    else:
        opts = list(varargin)
        for k, v in kw.items():
            opts += [k, v]
        if len(opts) % 2:
            raise ScmcArgumentError("Parameters must be in struct of option and value")
        if targetArea.get("type") != "targetArea":
            raise ScmcArgumentError("Invalid target area input")
        if views.get("type") != "source":
            raise ScmcArgumentError("Invalid views input")
        if lights.get("type") != "source":
            raise ScmcArgumentError("Invalid lights input")
        if scatter.get("type") != "scatter":
            raise ScmcArgumentError("Invalid scatter input")
        if not np.isscalar(iterations) or iterations < 1:
            raise ScmcArgumentError("Invalid photons number")
        if scatter.get("function") == "measuredFarField" and targetArea["D"] == 3:
            raise ScmcArgumentError("Measured scattering of far field is used only in 2D")
        views, lights = dict(views), dict(lights)
        tag = "3D" if targetArea["D"] == 3 else "2D"   # scmc.m:97-127
        for src in (views, lights):
            for base in ("directions", "positions"):
                if base + tag in src:
                    src[base] = src[base + tag]
        Config = {
            "targetArea": targetArea, "views": views, "lights": lights, "scatter": scatter,
            "render": {"iterations": iterations, "singleScattering": False, "cov": True,
                       "multipleScattering": True, "field": False},
            "parforIters": 1, "rng": -1, "CBS": True, "uniformFirstScatter": False,
        }
        for i in range(0, len(opts), 2):
            name, val = str(opts[i]), opts[i + 1]
            if name == "parforIters":
                if not np.isscalar(val) or val < 0:
                    raise ScmcArgumentError("invalid parfor iterations number")
                if Config["rng"] != -1:
                    raise ScmcArgumentError("rng can not be used while using parfor")
                Config["parforIters"] = int(val)
            elif name == "rng":
                if not np.isscalar(val) or val < 0:
                    raise ScmcArgumentError("invalid rng number")
                if Config["parforIters"] != 1:
                    raise ScmcArgumentError("rng can not be used while using parfor")
                Config["rng"] = int(val)
            elif name in _BOOL_OPTS:
                key, msg = _BOOL_OPTS[name]
                if not isinstance(val, (bool, np.bool_)):
                    raise ScmcArgumentError(msg)
                _set(Config, key, bool(val))
            else:
                raise ScmcArgumentError("Invalid option value")
    return _run(Config, precision)


def _engine_args(Config):
    """scmc.m:245-323."""
    ta, sc = Config["targetArea"], Config["scatter"]
    sigt = 1.0 / ta["MFP"]
    albedo = sc["albedo"]
    ampfunc0 = None
    if sc["function"] == "isotropic":
        sct_type, ampfunc = 1, 0
    elif sc["function"] == "tabulated":
        sct_type, ampfunc = 2, sc["ampfunc"]
        a0 = sc.get("ampfunc0", float("inf"))
        if not (np.isscalar(a0) and np.isinf(a0)):
            ampfunc0 = a0
    elif sc["function"] == "HG":
        sct_type, ampfunc = 3, sc["g"]
        if not np.isinf(sc["g0"]):
            ampfunc0 = sc["g0"]
    else:
        raise ScmcArgumentError("Invalid scatter input")
    if ta["D"] == 3:
        box_min = [ta["x"][0], ta["y"][0], ta["z"][0]]
        box_max = [ta["x"][1], ta["y"][1], ta["z"][1]]
    else:
        box_min = [ta["x"][0], ta["z"][0]]
        box_max = [ta["x"][1], ta["z"][1]]
    L, V = Config["lights"], Config["views"]
    is_ff_l = 1 if L["farField"] == 1 else 0
    l = L["directions"] if is_ff_l else L["positions"]
    is_ff_v = 1 if V["farField"] == 1 else 0
    v = V["directions"] if is_ff_v else V["positions"]
    smpFlg = 1 if Config["uniformFirstScatter"] else 2
    return dict(sigt=sigt, albedo=albedo, box_min=box_min, box_max=box_max, l=l, v=v, is_ff_l=is_ff_l,
                is_ff_v=is_ff_v, lambda_=ta["wavelength"], doCBS=Config["CBS"], smpFlg=smpFlg,
                sct_type=sct_type, ampfunc=ampfunc, ampfunc0=ampfunc0)


def _seed(Config):
    if Config["rng"] != -1:
        return int(Config["rng"])
    s = _next_seed[0]
    _next_seed[0] += 1
    return s


def _run(Config, precision):
    mulRes = {"Config": Config}
    a = _engine_args(Config)
    maxItr = int(Config["render"]["iterations"])
    P = max(1, int(Config["parforIters"]))
    if Config["render"]["cov"]:
        seed = _seed(Config)
        Ms = Mm = 0
        for it in range(P):          # scmc.m:343-367: mean over repeats
            ms, mm = E.MCcov(a["sigt"], a["albedo"], a["box_min"], a["box_max"], a["l"], a["v"], a["is_ff_l"],
                             a["is_ff_v"], maxItr, a["lambda_"], a["doCBS"], a["smpFlg"], a["sct_type"],
                             a["ampfunc"], a["ampfunc0"], seed=seed, path_offset=it * maxItr, precision=precision)
            Ms = Ms + ms
            Mm = Mm + mm
        Ms, Mm = Ms / P, Mm / P
        if Config["render"]["singleScattering"]:
            mulRes["Csingle"] = Ms
        if Config["render"]["multipleScattering"]:
            mulRes["C"] = Ms + Mm
    if Config["render"]["field"]:
        seed = _seed(Config)
        u = 0
        for it in range(P):          # scmc.m:398-421: sum(u sqrt(maxItr)) / sqrt(maxItr P)
            ui = E.MCfield(a["sigt"], a["albedo"], a["box_min"], a["box_max"], a["l"], a["v"], a["is_ff_l"],
                           a["is_ff_v"], maxItr, a["lambda_"], a["doCBS"], a["smpFlg"], a["sct_type"],
                           a["ampfunc"], a["ampfunc0"], seed=seed, path_offset=it * maxItr, precision=precision)
            u = u + ui * np.sqrt(maxItr)
        mulRes["field"] = u / np.sqrt(maxItr * P) if P > 1 else u / np.sqrt(maxItr)
    return mulRes
