"""Digitises the reference's published diagonal-intensity figure into a fixture.

Source: examples/buildFarFieldCovMatrix/farFieldCov.pdf of the reference
repository, the "Compare diagonals" figure of farFieldCov.m:111-122:
diag(abs(C(:,:,1,1))) against the view angle 0..359 deg for the light at 0 deg,
with CBS (blue) and without (orange), from scmc() with 1e3 paths x 12 parfor
repeats (farFieldCov.m:31-38) on the 2-D target boxArea(1, 10, [-10,10],
[-10,10]) with the tabulated scatterer of scatteringAmplitude.mat.

The PDF embeds the figure as a 560 x 420 RGB raster (FlateDecode).  The axes
are calibrated from the raster itself: the x axis line and its tick marks
(0..400 deg, a tick every 50), the y axis line and its ticks (1.5..5, a tick
every 0.5).  Each curve is read per view column as the mean row of the pixels
of its colour (MATLAB's default first and second line colours), which is the
value at that column for a line two pixels wide; where the orange curve hides
the blue one (the curves are equal there) the blue value is left out.  The
CBS peak is also read as the top of the blue line at 180 deg.

Run once in the container that holds the reference; tests use the JSON.
"""
import json
import os
import re
import sys
import zlib

import numpy as np

REF = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
PDF = os.path.join(REF, "examples", "buildFarFieldCovMatrix", "farFieldCov.pdf")


def rasters(data):
    """(width, height, pixels) of every 8-bit RGB image XObject in the file."""
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
            out.append((w, h, np.frombuffer(raw, np.uint8).reshape(h, w, 3).astype(float)))
    return out


def colour_mask(a, rgb, tol=60.0):
    return np.sqrt(((a - np.asarray(rgb, dtype=float)) ** 2).sum(axis=2)) < tol


BLUE, ORANGE = (0, 114, 189), (217, 83, 25)


def axes(a):
    """Pixel calibration: x = x0 + deg * sx, y = y0 - (I - 1.5) * sy."""
    dark = a.sum(axis=2) < 200
    row = int(np.argmax(dark.sum(axis=1)))      # the x axis line
    col = int(np.argmax(dark.sum(axis=0)))      # the y axis line
    xt = np.where(dark[row - 3, col + 1:])[0] + col + 1     # ticks point into the axes
    yt = np.where(dark[:row - 1, col + 2])[0]
    assert len(xt) == 8 and len(yt) == 7, (xt, yt)
    sx = (xt[-1] - col) / 400.0                 # last tick: 400 deg
    sy = (row - yt[0]) / 3.5                    # top tick: 5.0
    return col, sx, row, sy


def digitise(a):
    x0, sx, y0, sy = axes(a)
    blue, orange = colour_mask(a, BLUE), colour_mask(a, ORANGE)
    # the legend box: the two long dark rows inside the axes above the curves,
    # and the columns its top edge spans
    legend = np.zeros_like(blue)
    dark = a.sum(axis=2) < 200
    inner = dark[:, x0 + 2:]
    rows = [r for r in range(int(y0 - 3.5 * sy) + 2, int(y0) - 2) if inner[r].sum() > 60]
    top, bottom = min(rows), max(rows)
    cols = np.where(dark[top])[0]
    cols = cols[cols > x0 + 2]
    legend[top:bottom + 1, cols.min():cols.max() + 1] = True
    val = lambda r: 1.5 + (y0 - r) / sy
    cbs, ncbs = [], []
    for d in range(360):
        c = int(round(x0 + d * sx))
        out = []
        for m in (blue, orange):
            r = np.where(m[:, c] & ~legend[:, c])[0]
            r = r[(r > y0 - 3.5 * sy) & (r < y0)]
            out.append(round(float(val(r.mean())), 3) if len(r) else None)
        cbs.append(out[0])
        ncbs.append(out[1])
    c = int(round(x0 + 180 * sx))
    r = np.where(blue[:, c] & ~legend[:, c])[0]
    peak = round(float(val(r.min() + 1.0)), 3)      # top of a two-pixel line
    return cbs, ncbs, peak, (float(x0), float(sx), float(y0), float(sy))


def main():
    data = open(PDF, "rb").read()
    figs = [r for r in rasters(data) if r[0] == 560 and r[1] == 420]
    # the diagonal figure is the one that holds both line colours
    fig = [a for _, _, a in figs if colour_mask(a, BLUE).sum() > 500 and colour_mask(a, ORANGE).sum() > 500]
    assert len(fig) == 1
    cbs, ncbs, peak, cal = digitise(fig[0])
    out = {"source": "examples/buildFarFieldCovMatrix/farFieldCov.pdf (figure of farFieldCov.m:111-122)",
           "config": {"boxArea": [1, 10, [-10, 10], [-10, 10]], "views_deg": "0:1:359", "light_deg": 0,
                      "scatter": "tabulatedAmplitudeScatter(scatteringAmplitude.mat)",
                      "paths": "1e3 x 12 parfor repeats"},
           "calibration_px": {"x0": cal[0], "px_per_deg": cal[1], "y0": cal[2], "px_per_unit": cal[3]},
           "deg": list(range(360)), "with_cbs": cbs, "without_cbs": ncbs, "cbs_peak_180": peak}
    with open(os.path.join(HERE, "farFieldCov_diagonal.json"), "w") as f:
        json.dump(out, f, indent=0)
    print("peak", peak, "hidden blue", sum(v is None for v in cbs), "missing orange", sum(v is None for v in ncbs))


if __name__ == "__main__":
    main()
