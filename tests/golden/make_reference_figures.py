"""Digitises the reference's published example figures into fixtures.

The reference ships, beside each example script, the PDF its author made by
running it in MATLAB.  Four of them plot numbers this project computes:

* examples/buildFarFieldCovMatrix/farFieldCov.pdf -- diag(abs(C(:,:,1,1)))
  over views 0..359 deg, light at 0 deg, with CBS (blue) and without
  (orange) (farFieldCov.m:111-122);
* examples/buildNearFieldCovMatrix/nearFieldCov.pdf -- the same for sensors
  on a circle of radius 20 (near field) and an HG(-0.5) scatterer
  (nearFieldCov.m, its "Compare diagonals" figure);
* examples/memoryEffectMeasurements/memoryEffectMeasurements.pdf -- the
  intensity memory-effect correlation against the tilt 0..0.02 deg, measured
  by scmc() (blue; the dashed curve is the closed form the script plots);
* examples/multipleAndSingleScattering/multipleAndSingleScattering.pdf --
  single-, multiple- and total-scattering correlation against the rotation
  0..90 deg (blue, orange, yellow).

Each PDF embeds its figure as a 560 x 420 RGB raster (FlateDecode).  The axes
are calibrated from the raster itself: the axis lines and their tick marks,
with the tick values the figure labels.  Each curve is read per data column
as the mean row of the pixels of its colour (MATLAB's default line colours),
the value there for a line two pixels wide; the legend box is masked, and a
column where a curve drawn later hides the curve is left out (None).  The
CBS peaks are also read as the top of the blue line at 180 deg.

Run once in the container that holds the reference; tests use the JSON files.
"""
import json
import os
import re
import sys
import zlib

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))

BLUE, ORANGE, YELLOW = (0, 114, 189), (217, 83, 25), (237, 177, 32)

FIGURES = {
    "farFieldCov_diagonal": dict(
        pdf="buildFarFieldCovMatrix/farFieldCov.pdf", x_step=50.0, y_origin=1.5, y_step=0.5,
        xs=list(range(360)), curves={"with_cbs": BLUE, "without_cbs": ORANGE}, peak_at=180,
        config={"boxArea": [1, 10, [-10, 10], [-10, 10]], "views_deg": "0:1:359", "light_deg": 0,
                "scatter": "tabulatedAmplitudeScatter(scatteringAmplitude.mat)", "paths": "1e3 x 12 parfor repeats"}),
    "nearFieldCov_diagonal": dict(
        pdf="buildNearFieldCovMatrix/nearFieldCov.pdf", x_step=50.0, y_origin=0.05, y_step=0.05,
        xs=list(range(360)), curves={"with_cbs": BLUE, "without_cbs": ORANGE}, peak_at=180,
        config={"boxArea": [1, 10, [-10, 10], [-10, 10]], "views": "nearFieldSource(20, deg2rad(0:1:359))",
                "light_deg": 0, "scatter": "HGScatter(-0.5)", "paths": "1e3 x 12 parfor repeats"}),
    "memoryEffect_curve": dict(
        pdf="memoryEffectMeasurements/memoryEffectMeasurements.pdf", x_step=0.002, y_origin=0.0, y_step=0.1,
        xs=[round(0.001 * k, 3) for k in range(21)], curves={"measured": BLUE},
        config={"boxArea": "[1, 200, [0, 2000], [-5e4, 5e4], [-5e4, 5e4]]", "scatter": "HGScatter(0, 0)",
                "lights_deg": "0.002 + [0, dtheta]", "views_deg": "-1:0.1:-0.1 and the same minus dtheta",
                "CBS": False, "paths": "1e3 x 12 parfor repeats"}),
    "multipleAndSingleScattering_curves": dict(
        pdf="multipleAndSingleScattering/multipleAndSingleScattering.pdf", x_step=10.0, y_origin=0.0, y_step=0.1,
        xs=list(range(0, 91, 3)), curves={"single": BLUE, "multiple": ORANGE, "total": YELLOW},
        config={"boxArea": [1, 20, [-50, 50], [-50, 50], [-50, 50]], "scatter": "isotropicScatter",
                "span_deg": 22, "paths": "1e3 x 12 parfor repeats"}),
}


def rasters(data):
    """Pixels of every 8-bit RGB image XObject in a PDF."""
    objs = {int(n): b for n, b in re.findall(rb"(\d+) 0 obj(.*?)endobj", data, re.S)}
    out = []
    for body in objs.values():
        if b"/Subtype /Image" not in body:
            continue
        w = int(re.search(rb"/Width (\d+)", body).group(1))
        h = int(re.search(rb"/Height (\d+)", body).group(1))
        m = re.search(rb"/Length (\d+) 0 R", body)
        n = int(re.search(rb"(\d+)", objs[int(m.group(1))]).group(1)) if m else int(
            re.search(rb"/Length (\d+)", body).group(1))
        s = body.index(b"stream") + len(b"stream")
        while body[s:s + 1] in (b"\r", b"\n"):
            s += 1
        raw = zlib.decompress(body[s:s + n])
        if len(raw) == w * h * 3:
            out.append(np.frombuffer(raw, np.uint8).reshape(h, w, 3).astype(float))
    return out


def colour_mask(a, rgb, tol=60.0):
    return np.sqrt(((a - np.asarray(rgb, dtype=float)) ** 2).sum(axis=2)) < tol


def tick_scale(ticks, origin):
    """Pixels per tick from tick positions measured from the axis origin:
    ticks are assigned to whole tick counts and fitted through the origin."""
    d = np.diff(ticks)
    step = np.median(d[d > 5])
    k = np.round((ticks - origin) / step)
    ok = (k > 0) & (np.abs(ticks - origin - k * step) < 0.25 * step)
    return float((k[ok] * (ticks[ok] - origin)).sum() / (k[ok] ** 2).sum())


def calibrate(a, spec):
    """x = x0 + (X / x_step) * px_x;  y = y0 - ((Y - y_origin) / y_step) * px_y."""
    dark = a.sum(axis=2) < 200
    y0 = int(np.argmax(dark.sum(axis=1)))      # the x axis line
    x0 = int(np.argmax(dark.sum(axis=0)))      # the y axis line
    xt = np.where(dark[y0 - 3, x0 + 1:])[0] + x0 + 1     # ticks point into the axes
    yt = np.where(dark[:y0 - 1, x0 + 2])[0]
    px_x = tick_scale(xt, x0)
    px_y = tick_scale(-yt[::-1].astype(float), -float(y0))
    return x0, px_x / spec["x_step"], y0, px_y / spec["y_step"]


def legend_mask(a, x0, y0, sy, spec):
    """The legend box: two long dark rows in the upper half of the axes."""
    dark = a.sum(axis=2) < 200
    top_of_axes = int(y0 - 0.9 * (y0 - 0))
    rows = [r for r in range(5, int(y0 - (y0 - 30) / 2)) if dark[r, x0 + 3:].sum() > 60 and r > 33]
    m = np.zeros(a.shape[:2], bool)
    if rows:
        top, bottom = min(rows), max(rows)
        cols = np.where(dark[top])[0]
        cols = cols[cols > x0 + 3]
        m[top:bottom + 1, cols.min():cols.max() + 1] = True
    return m


def digitise(a, spec):
    x0, sx, y0, sy = calibrate(a, spec)
    legend = legend_mask(a, x0, y0, sy, spec)
    val = lambda r: spec["y_origin"] + (y0 - r) / sy
    out = {}
    for name, rgb in spec["curves"].items():
        m = colour_mask(a, rgb) & ~legend
        vals = []
        for x in spec["xs"]:
            c = int(round(x0 + x * sx))
            r = np.where(m[:, c])[0]
            r = r[(r > 20) & (r < y0)]
            vals.append(round(float(val(r.mean())), 4) if len(r) else None)
        out[name] = vals
    if "peak_at" in spec:
        c = int(round(x0 + spec["peak_at"] * sx))
        r = np.where(colour_mask(a, BLUE)[:, c] & ~legend[:, c])[0]
        out["cbs_peak"] = round(float(val(r.min() + 1.0)), 4)     # top of a two-pixel line
    return out, {"x0": float(x0), "px_per_x": float(sx), "y0": float(y0), "px_per_y": float(sy)}


def main():
    for name, spec in FIGURES.items():
        figs = [a for a in rasters(open(os.path.join(REF, "examples", spec["pdf"]), "rb").read())
                if a.shape == (420, 560, 3)]
        figs = [a for a in figs if all(colour_mask(a, c).sum() > 300 for c in spec["curves"].values())]
        assert len(figs) == 1, name
        curves, cal = digitise(figs[0], spec)
        out = {"source": "examples/" + spec["pdf"], "config": spec["config"], "calibration_px": cal,
               "x": spec["xs"], **curves}
        with open(os.path.join(HERE, name + ".json"), "w") as f:
            json.dump(out, f, indent=0)
        print(name, cal, {k: (sum(v is None for v in c) if isinstance(c, list) else c) for k, c in curves.items()})


if __name__ == "__main__":
    main()
