"""Problem-description builders with the reference's struct conventions.

Each returns a dict whose keys are the MATLAB struct fields, so code written
against the reference reads the same:
    boxArea                    interface/targetArea/boxArea.m:1-65
    farFieldSource             interface/source/farFieldSource.m:1-57
    nearFieldSource            interface/source/nearFieldSource.m:1-91
    isotropicScatter           interface/scatter/isotropicScatter.m:1-33
    HGScatter                  interface/scatter/HGScatter.m:1-56
    tabulatedAmplitudeScatter  interface/scatter/tabulatedAmplitudeScatter.m:1-101
"""
import numpy as np

from .engines import ScmcArgumentError


def boxArea(wavelength, MFP, z, x, y=None):
    """interface/targetArea/boxArea.m:24-63."""
    if not (np.isscalar(wavelength) and np.isscalar(MFP)):
        raise ScmcArgumentError("wavelength and MFP are scalars")
    z, x = np.asarray(z).reshape(-1), np.asarray(x).reshape(-1)
    if z.size != 2 or x.size != 2 or (y is not None and np.asarray(y).size != 2):
        raise ScmcArgumentError("x, y and z defined as 2 elements vectors")
    if np.iscomplexobj(z) or np.iscomplexobj(x) or (y is not None and np.iscomplexobj(y)):
        raise ScmcArgumentError("x, y and z are real values")
    z, x = z.astype(float), x.astype(float)
    if z[0] > z[1]:
        raise ScmcArgumentError("z(1) must be smaller than z(2)")
    if x[0] > x[1]:
        raise ScmcArgumentError("x(1) must be smaller than x(2)")
    t = {"type": "targetArea", "wavelength": float(wavelength), "MFP": float(MFP), "z": z, "x": x}
    if y is None:
        t["D"] = 2
    else:
        y = np.asarray(y, dtype=float).reshape(-1)
        if y[0] > y[1]:
            raise ScmcArgumentError("y(1) must be smaller than y(2)")
        t["D"] = 3
        t["y"] = y
    return t


def farFieldSource(in1, *theta_flag):
    """interface/source/farFieldSource.m:20-54.

    farFieldSource(directions) with directions 2xN or 3xN (normalised), or
    farFieldSource(theta, 0) with N angles in radians."""
    if len(theta_flag) > 1:
        raise ScmcArgumentError("too many input arguments")
    s = {"type": "source", "farField": 1}
    if not theta_flag:
        d = np.asarray(in1, dtype=float)
        if d.ndim != 2 or d.shape[0] not in (2, 3):
            raise ScmcArgumentError("Invalid positions elements number")
        s["count"] = d.shape[1]
        s["directions"] = d / np.sqrt(np.sum(np.abs(d) ** 2, axis=0))
    else:
        th = np.asarray(in1, dtype=float)
        if th.ndim > 1 and min(th.shape) != 1:
            raise ScmcArgumentError("theta must be vector of directions")
        th = th.reshape(-1)
        s["count"] = th.size
        s["directions2D"] = np.vstack([np.sin(th), np.cos(th)])
        s["directions3D"] = np.vstack([np.sin(th), np.zeros_like(th), np.cos(th)])
    return s


def nearFieldSource(in1, in2=None, in3=None):
    """interface/source/nearFieldSource.m:19-88."""
    s = {"type": "source", "farField": 0}
    if in2 is None:
        p = np.asarray(in1, dtype=float)
        if p.ndim != 2 or p.shape[0] not in (2, 3):
            raise ScmcArgumentError("Invalid positions elements number")
        s["count"] = p.shape[1]
        s["positions"] = p
    elif in3 is None:
        if not (np.isscalar(in1) and in1 > 0):
            raise ScmcArgumentError("r must be positive scalar")
        th = np.asarray(in2, dtype=float)
        if th.ndim > 1 and min(th.shape) != 1:
            raise ScmcArgumentError("theta must be vector of directions")
        th = th.reshape(-1)
        s["count"] = th.size
        s["positions2D"] = in1 * np.vstack([np.sin(th), np.cos(th)])
        s["positions3D"] = in1 * np.vstack([np.sin(th), np.zeros_like(th), np.cos(th)])
    else:
        X, Y = np.asarray(in1, dtype=float), np.asarray(in2, dtype=float)
        if X.ndim != 2 or X.shape[0] not in (2, 3) or X.shape[1] != 1:
            raise ScmcArgumentError("Invalid X position")
        if Y.ndim != 2 or Y.shape[0] not in (2, 3) or Y.shape[1] != 1:
            raise ScmcArgumentError("Invalid Y position")
        if X.shape[0] != Y.shape[0]:
            raise ScmcArgumentError("Both points must have the same diamentions")
        if not (np.isscalar(in3) and float(in3).is_integer() and in3 > 0):
            raise ScmcArgumentError("Number of input points must be positive integer")
        s["count"] = int(in3)
        s["positions"] = np.vstack([np.linspace(X[j, 0], Y[j, 0], int(in3)) for j in range(X.shape[0])])
    return s


def _albedo(albedo):
    if albedo is None:
        return 1.0
    if not np.isscalar(albedo) or albedo < 0 or albedo > 1:
        raise ScmcArgumentError("Invalid albedo input")
    return float(albedo)


def isotropicScatter(albedo=None):
    """interface/scatter/isotropicScatter.m:17-31."""
    return {"type": "scatter", "function": "isotropic", "albedo": _albedo(albedo)}


def HGScatter(g, g0=None, albedo=None):
    """interface/scatter/HGScatter.m:23-54.  g0 = inf means "no ampfunc0"."""
    s = {"type": "scatter", "function": "HG", "albedo": _albedo(albedo)}
    if g0 is not None:
        if not np.isscalar(g0) or g0 < -1 or g0 > 1:
            raise ScmcArgumentError("Invalid g0 input")
        s["g0"] = float(g0)
    else:
        s["g0"] = float("inf")
    if not np.isscalar(g) or g < -1 or g > 1:
        raise ScmcArgumentError("Invalid g input")
    s["g"] = float(g)
    return s


def invCDF(cdf, dim):
    """coreFunctions/invCDF.m:3-17: nearest grid point, index*pi/N scaling."""
    cdf = np.asarray(cdf, dtype=float).reshape(-1)
    n = cdf.size
    grid = np.arange(n) * (1.0 / (n - 1))
    idx = np.empty(n)
    for j in range(n):
        idx[j] = np.argmin(np.abs(cdf - grid[j])) + 1   # first minimum, 1-based
    return idx / n * (2 * np.pi if dim == 2 else np.pi)


def tabulatedAmplitudeScatter(theta, f, albedo=None):
    """interface/scatter/tabulatedAmplitudeScatter.m:42-99.

    The struct also carries ``ampfunc0 = inf`` so scmc() accepts it (the
    reference builder leaves the field undefined and scmc.m:261 then fails)."""
    theta = np.asarray(theta, dtype=float)
    f = np.asarray(f, dtype=complex)
    if (theta.ndim > 1 and min(theta.shape) != 1) or (f.ndim > 1 and min(f.shape) != 1):
        raise ScmcArgumentError("theta and f must be vectors")
    theta, f = theta.reshape(-1), f.reshape(-1)
    if theta.size != f.size:
        raise ScmcArgumentError("theta and f must have same elements number")
    if np.any(np.diff(theta) < 0):
        raise ScmcArgumentError("theta must be monotonic increasing")
    if theta[0] == 0 and theta[-1] == 2 * np.pi:
        D = 2
    elif theta[0] == 0 and theta[-1] == np.pi:
        D = 3
    else:
        raise ScmcArgumentError("theta must be from 0 to 2pi in 2D or -1 to 1 in 3D")
    s = {"type": "scatter", "function": "tabulated", "albedo": 1.0 if albedo is None else float(albedo), "D": D,
         "ampfunc0": float("inf")}
    a2 = np.abs(f) ** 2
    af = {"theta": theta}
    if D == 2:
        af["cs"] = np.mean(a2) * 2 * np.pi
        af["samplePdf"] = a2 / np.sum(a2)
    else:
        af["cs"] = np.mean(a2 * np.sin(theta)) * 2 * np.pi ** 2
        af["samplePdf"] = a2 * np.sin(theta) / np.sum(a2 * np.sin(theta))
    af["sampleCdf"] = np.cumsum(af["samplePdf"])
    af["sampleIcdf"] = invCDF(af["sampleCdf"], D)
    af["evalPdf"] = a2 / af["cs"]
    af["evalAmp"] = af["evalPdf"] ** 0.5 * np.exp(1j * np.angle(f))
    s["ampfunc"] = af
    return s
