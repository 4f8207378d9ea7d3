"""Serving with the resident update loop (scmc_server_config.resident_loop).

The mixed updates of every request go to one persistent kernel that takes
their work items from a job table in host memory.  Each item runs the same
arithmetic as in a launch of its own and adds into the same accumulator copy,
and a request's chunks are published in order, so every covariance must equal
the one from separate launches exactly (checked to 1e-12 in the
Cauchy-Schwarz metric, and reported when it is not bit-identical):
  * a mix of requests of different N, CBS on and off, FP64 requests beside;
  * requests far enough apart that each instance leaves (idle) and the next
    publish launches another;
  * more jobs than the job table holds (a new generation of the table);
  * a stream drawn from the serving workload against the FP64 one-shot engine
    (<= 1e-4);
  * cancel and server stop with requests in flight, and field requests served
    beside the loop's covariance requests.
"""
import sys
import time

import numpy as np
import pytest

from conftest import cone_dirs, cov_matrix, cs_scale
from mcspeckle_amd import engines as E

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cs_err(M, R):
    C, Rr = cov_matrix(M), cov_matrix(R)
    return float(np.max(np.abs(C - Rr) / np.maximum(cs_scale(Rr), 1e-300)))


def req(i, nv, nl, od, g, seed, cbs=True):
    rng = np.random.default_rng(300 + i)
    return E.make_problem(od, 0.99, [-5e4, -5e4, 0], [5e4, 5e4, 1], cone_dirs(nl, 15, rng), cone_dirs(nv, 15, rng),
                          1, 1, 1.0, cbs, 2, 3, g, seed=seed)


MIX = [(0, 8, 8, 5.0, 0.5, 2000, "mixed", True), (1, 33, 7, 12.0, 0.9, 1500, "mixed", False),
       (2, 5, 4, 2.0, 0.2, 2000, "fp64", True), (3, 64, 16, 8.0, 0.8, 600, "mixed", True),
       (4, 11, 3, 1.0, 0.95, 3000, "mixed", True), (5, 40, 40, 6.0, 0.7, 300, "mixed", False)]


def run(mix, resident, max_active=4, slice_paths=300, gap=0.0):
    out = {}
    with E.Server(max_active=max_active, paths_per_slice=slice_paths, resident_loop=resident) as S:
        ids = {}
        for i, nv, nl, od, g, n, prec, cbs in mix:
            ids[i] = S.submit(req(i, nv, nl, od, g, seed=i, cbs=cbs), n, prec)
            if gap:
                out[i] = S.wait(ids.pop(i))
                time.sleep(gap)
        say("submitted, resident loop", resident)
        out.update({i: S.wait(r) for i, r in ids.items()})
    return out


def compare(a, b, mix):
    for i, *_ in mix:
        Ms_a, Mm_a, sa = a[i]
        Ms_b, Mm_b, sb = b[i]
        assert sa["paths"] == sb["paths"]
        same = np.array_equal(Ms_a, Ms_b) and np.array_equal(Mm_a, Mm_b)
        e = max(cs_err(Ms_b, Ms_a), cs_err(Mm_b, Mm_a))
        say("  request", i, "bit-identical" if same else "max cs err %.3g" % e)
        assert e <= 1e-12, i


def test_resident_loop_matches_separate_launches():
    ref = run(MIX, False)
    res = run(MIX, True)
    compare(ref, res, MIX)


def test_resident_loop_restarts_after_idle():
    """Requests one at a time with 50 ms between them (the loop leaves after
    2 ms without work): every publish after a gap finds the instance gone or
    leaving and launches the next one."""
    mix = [m for m in MIX if m[6] == "mixed"][:4]
    ref = run(mix, False, max_active=1)
    res = run(mix, True, max_active=1, gap=0.05)
    compare(ref, res, mix)


def test_resident_loop_job_table_wraps():
    """One small request in 2300 slices of 256 paths: two jobs per chunk, so
    the 4096-job table fills and the loop starts a new generation mid-request."""
    mix = [(7, 4, 4, 3.0, 0.5, 256 * 2300, "mixed", True)]
    ref = run(mix, False, max_active=1, slice_paths=256)
    res = run(mix, True, max_active=1, slice_paths=256)
    compare(ref, res, mix)


def test_resident_loop_serving_stream_matches_fp64():
    """Eight requests of the serving workload (Nv, Nl in {4..64}, optical depth
    in [1, 20], HG g in [0, 0.95], paths log-uniform in [1e5, 1e7] scaled by
    0.002) through the resident loop, each against the FP64 one-shot engine."""
    from mcspeckle_amd import workloads as W
    stream = W.request_stream(8, 50.0, 0.002, seed=77)
    with E.Server(max_active=4, resident_loop=True) as S:
        ids = [S.submit(spec, n, "mixed") for _, spec, n in stream]
        say("stream submitted", [(spec.nv, spec.nl, n) for _, spec, n in stream])
        res = [S.wait(r) for r in ids]
    for (_, spec, n), (Ms, Mm, st) in zip(stream, res):
        assert st["paths"] == n
        Ms0, Mm0 = E.mccov_spec(spec, n, "fp64")
        err = cs_err(Ms + Mm, Ms0 + Mm0)
        say("  request N", spec.nv * spec.nl, "paths", n, "err", err)
        assert err <= 1e-4, (spec.nv, spec.nl, err)


def test_resident_loop_cancel_and_stop_with_requests_in_flight():
    """Cancelling a running request and stopping the server with requests in
    flight end cleanly with the resident loop: the cancelled and the
    never-admitted requests report SCMC_E_CANCELLED, a request that finishes
    in between is exact, and the next server starts normally."""
    from mcspeckle_amd import _native as N
    ref = run([MIX[0]], False, max_active=1)
    S = E.Server(max_active=2, paths_per_slice=300, resident_loop=True)
    try:
        big = S.submit(req(9, 64, 32, 10.0, 0.9, seed=9), 200000, "mixed")
        small = S.submit(req(0, 8, 8, 5.0, 0.5, seed=0), 2000, "mixed")
        pending = [S.submit(req(10 + i, 48, 48, 8.0, 0.8, seed=10 + i), 100000, "mixed") for i in range(3)]
        Ms, Mm, st = S.wait(small)
        assert np.array_equal(Ms, ref[0][0]) and np.array_equal(Mm, ref[0][1])
        time.sleep(0.2)
        S.cancel(big)
        with pytest.raises(N.ScmcError) as ei:
            S.wait(big)
        assert ei.value.code == N.SCMC_E_CANCELLED
        say("cancelled the running request; stopping with", len(pending), "more in flight or pending")
    finally:
        S.stop()
    for rid in pending:
        with pytest.raises(N.ScmcError) as ei:
            S.wait(rid)
        assert ei.value.code in (N.SCMC_E_CANCELLED, N.SCMC_E_ARG)
    again = run([MIX[0]], True, max_active=1)
    assert np.array_equal(again[0][0], ref[0][0]) and np.array_equal(again[0][1], ref[0][1])


def test_resident_loop_beside_field_requests():
    """Field requests (not updates, so not loop jobs) served beside
    covariance requests whose updates go to the resident loop: both kinds
    equal their results without the loop."""
    rng = np.random.default_rng(44)
    field = E.make_problem(6.0, 0.97, [-5e4, -5e4, 0], [5e4, 5e4, 1], cone_dirs(2, 10, rng), cone_dirs(900, 12, rng),
                           1, 1, 1.0, True, 2, 3, 0.8, seed=70)
    out = {}
    for resident in (False, True):
        with E.Server(max_active=3, paths_per_slice=400, resident_loop=resident) as S:
            fid = S.submit(field, 3000, "mixed", kind="field")
            cids = {i: S.submit(req(i, nv, nl, od, g, seed=i, cbs=cbs), n, prec)
                    for i, nv, nl, od, g, n, prec, cbs in MIX[:3]}
            out[resident] = (S.wait(fid), {i: S.wait(r) for i, r in cids.items()})
    (ua, _), ca = out[False]
    (ub, _), cb = out[True]
    assert np.array_equal(ua, ub)
    compare(ca, cb, MIX[:3])
