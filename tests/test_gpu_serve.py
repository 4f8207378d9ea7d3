"""Serving: continuous batching of covariance requests on one GPU.

* Every result is the same alone and in the mix: a request's result depends
  only on its own Philox paths and slice schedule.  Tolerance: identical to
  1e-12 (Cauchy-Schwarz scaled) between the mixed-in-a-batch and the alone run
  through the same server, and within the precision's bound (1e-4 mixed,
  1e-12 FP64) against the one-shot engine.
* Admission by footprint: with a budget that fits one request at a time, the
  requests still all complete, one after another.
* Under a budget that makes every session be released after its request
  (its blocks going to the process-wide idle cache), results equal those of
  a server that keeps its sessions, to 1e-12.
* Cancel frees the slot and reports SCMC_E_CANCELLED.
* Shared update launches (the chunks of concurrent requests in one
  matrix-core launch) give the results of separate launches, to 1e-12.
"""
import sys
import time

import numpy as np
import pytest

from conftest import cone_dirs, cov_matrix, cs_scale
from mcspeckle_amd import _native as N
from mcspeckle_amd import engines as E

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def cs_err(M, R):
    C, Rr = cov_matrix(M), cov_matrix(R)
    return float(np.max(np.abs(C - Rr) / np.maximum(cs_scale(Rr), 1e-300)))


def req(i, nv, nl, od, g, seed):
    rng = np.random.default_rng(100 + i)
    return E.make_problem(od, 0.99, [-5e4, -5e4, 0], [5e4, 5e4, 1], cone_dirs(nl, 15, rng), cone_dirs(nv, 15, rng),
                          1, 1, 1.0, True, 2, 3, g, seed=seed)


MIX = [(0, 8, 8, 5.0, 0.5, 2000, "mixed"), (1, 33, 7, 12.0, 0.9, 1500, "mixed"), (2, 5, 4, 2.0, 0.2, 4000, "fp64"),
       (3, 64, 16, 8.0, 0.8, 600, "mixed"), (4, 16, 12, 15.0, 0.0, 1000, "fp64"), (5, 11, 3, 1.0, 0.95, 3000, "mixed")]


def test_results_same_alone_and_in_the_mix():
    alone = {}
    with E.Server(max_active=1, paths_per_slice=500) as S:
        for i, nv, nl, od, g, n, prec in MIX:
            say("alone", i)
            rid = S.submit(req(i, nv, nl, od, g, seed=i), n, prec)
            alone[i] = S.wait(rid)
    with E.Server(max_active=4, paths_per_slice=500) as S:
        ids = {}
        for i, nv, nl, od, g, n, prec in MIX:
            ids[i] = S.submit(req(i, nv, nl, od, g, seed=i), n, prec)
        say("mix submitted")
        mixed = {i: S.wait(r) for i, r in ids.items()}
    for i, nv, nl, od, g, n, prec in MIX:
        Ms_a, Mm_a, _ = alone[i]
        Ms_m, Mm_m, _ = mixed[i]
        assert cs_err(Ms_m + Mm_m, Ms_a + Mm_a) <= 1e-12, i
        Ms1, Mm1 = E.mccov_spec(req(i, nv, nl, od, g, seed=i), n, prec)
        tol = 1e-4 if prec == "mixed" else 1e-12
        assert cs_err(Ms_m + Mm_m, Ms1 + Mm1) <= tol, i


def test_shared_update_launches_match_separate_ones():
    """Requests of different N whose chunks share matrix-core launches get
    the same covariance as with one launch per request: every work item reads
    and writes only its own problem's rows and accumulator copies."""
    # N <= 1024 share launches; the N = 1056 request launches on its own
    mix = [m for m in MIX if m[6] == "mixed"] + [(6, 33, 32, 6.0, 0.7, 400, "mixed")]
    out = {}
    for shared in (False, True):
        with E.Server(max_active=4, paths_per_slice=300, shared_launches=shared) as S:
            ids = {i: S.submit(req(i, nv, nl, od, g, seed=i), n, prec) for i, nv, nl, od, g, n, prec in mix}
            say("submitted, shared launches", shared)
            out[shared] = {i: S.wait(r) for i, r in ids.items()}
    for i, *_ in mix:
        Ms_a, Mm_a, _ = out[False][i]
        Ms_b, Mm_b, _ = out[True][i]
        assert cs_err(Ms_b, Ms_a) <= 1e-12, i
        assert cs_err(Mm_b, Mm_a) <= 1e-12, i


def test_sessions_released_under_a_small_budget_give_the_same_results():
    """A budget far below any request's footprint: requests run one at a time
    (a request is admitted alone when nothing runs), and after each one the
    slots keep more than half the budget, so every session is released and
    its blocks go to the process-wide idle cache, from which the next
    session takes them.  Results equal those of a server with the default
    budget (sessions kept), to 1e-12, mixed and FP64 alike."""
    out = {}
    for budget in (0.0, 1e-3):
        with E.Server(max_active=3, hbm_budget_gb=budget, paths_per_slice=500) as S:
            ids = {i: S.submit(req(i, nv, nl, od, g, seed=i), n, prec) for i, nv, nl, od, g, n, prec in MIX}
            say("submitted, budget", budget)
            out[budget] = {i: S.wait(r) for i, r in ids.items()}
    for i, *_ in MIX:
        Ms_a, Mm_a, _ = out[0.0][i]
        Ms_b, Mm_b, _ = out[1e-3][i]
        assert cs_err(Ms_b, Ms_a) <= 1e-12, i
        assert cs_err(Mm_b, Mm_a) <= 1e-12, i


def test_admission_by_footprint_serialises_big_requests():
    fp = E.request_footprint(64, 64, "mixed")
    assert fp > 1e9
    with E.Server(max_active=4, hbm_budget_gb=fp * 1.5 / 1e9, paths_per_slice=200) as S:
        ids = [S.submit(req(i, 64, 64, 10.0, 0.9, seed=i), 400, "mixed") for i in range(3)]
        say("big submitted")
        res = [S.wait(r) for r in ids]
    for Ms, Mm, st in res:
        C = cov_matrix(Ms + Mm)
        assert np.all(np.isfinite(C)) and np.real(np.diag(C)).min() > 0
        assert st["paths"] == 400


def test_cancel_frees_the_slot():
    with E.Server(max_active=1, paths_per_slice=100) as S:
        big = S.submit(req(0, 64, 32, 10.0, 0.9, seed=1), 200000, "fp64")
        small = S.submit(req(1, 4, 4, 3.0, 0.5, seed=2), 500, "mixed")
        time.sleep(0.5)
        S.cancel(big)
        with pytest.raises(N.ScmcError) as ei:
            S.wait(big)
        assert ei.value.code == N.SCMC_E_CANCELLED
        Ms, Mm, st = S.wait(small)
        assert np.all(np.isfinite(Ms + Mm))


def test_server_left_running_at_exit_ends_cleanly():
    """A Python process that exits with the server still running (no stop, a
    request in flight) ends normally: the binding's atexit hook stops the
    server (cancelling the request) before the C runtime's exit handlers."""
    import os
    import subprocess
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    code = (
        "import sys; sys.path.insert(0, %r); sys.path.insert(0, %r)\n"
        "import numpy as np\n"
        "from conftest import cone_dirs\n"
        "from mcspeckle_amd import engines as E\n"
        "rng = np.random.default_rng(3)\n"
        "spec = E.make_problem(5.0, 0.99, [-5e4, -5e4, 0], [5e4, 5e4, 1], cone_dirs(8, 15, rng), cone_dirs(8, 15, rng),"
        " 1, 1, 1.0, True, 2, 3, 0.8, seed=1)\n"
        "S = E.Server(max_active=2)\n"
        "S.submit(spec, 200000, 'mixed')\n"
        "print('submitted', flush=True)\n"
    ) % (root, os.path.join(root, "tests"))
    say("server exit without stop")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=180)
    assert "submitted" in r.stdout, r.stderr[-2000:]
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])


def test_stop_wakes_a_waiter_on_a_pending_request():
    """scmc_server_stop while another thread waits on a request that was never
    admitted (more requests than slots): the wait returns SCMC_E_CANCELLED and
    the server is freed only after the waiter has left it."""
    import ctypes as C
    import threading
    S = E.Server(max_active=1, paths_per_slice=100)
    try:
        spec_big = req(0, 64, 32, 10.0, 0.9, seed=1)
        S.submit(spec_big, 200000, "fp64")
        rids = [S.submit(req(i, 4, 4, 3.0, 0.5, seed=i), 500, "mixed") for i in (1, 2)]
        ne = 16 * 16
        bufs = [np.empty(ne) for _ in range(4)]
        out = {}

        def waiter():
            out["rc"] = N.lib().scmc_wait(int(rids[-1]), *[N.dptr(b) for b in bufs], None)

        t = threading.Thread(target=waiter)
        t.start()
        time.sleep(0.5)
        assert t.is_alive()            # the last request is still pending
        say("stop with a waiter")
    finally:
        S.stop()
    t.join(timeout=60)
    assert not t.is_alive()
    assert out["rc"] == N.SCMC_E_CANCELLED
    # the next server starts normally
    with E.Server(max_active=1) as S2:
        Ms, Mm, st = S2.wait(S2.submit(req(3, 4, 4, 3.0, 0.5, seed=3), 300, "mixed"))
        assert np.all(np.isfinite(Ms + Mm)) and st["paths"] == 300


def test_serving_stream_requests_match_fp64_one_shot():
    """Twelve requests of the serving workload (Nv, Nl uniform in {4..64},
    optical depth in [1, 20], HG g in [0, 0.95], paths log-uniform in
    [1e5, 1e7] scaled by 0.002, Poisson arrivals at 50 req/s) through a
    server with 4 slots in mixed precision; each result is checked against
    the one-shot FP64 engine on the same paths (<= 1e-4 in the
    Cauchy-Schwarz metric)."""
    from mcspeckle_amd import workloads as W
    stream = W.request_stream(12, 50.0, 0.002, seed=2024)
    with E.Server(max_active=4) as S:
        ids = []
        t0 = time.perf_counter()
        for ta, spec, n in stream:
            d = ta - (time.perf_counter() - t0)
            if d > 0:
                time.sleep(d)
            ids.append(S.submit(spec, n, "mixed"))
        say("stream submitted", [(spec.nv, spec.nl, n) for _, spec, n in stream])
        res = [S.wait(r) for r in ids]
    for (ta, spec, n), (Ms, Mm, st) in zip(stream, res):
        assert st["paths"] == n
        Ms0, Mm0 = E.mccov_spec(spec, n, "fp64")
        err = cs_err(Ms + Mm, Ms0 + Mm0)
        say("  request N", spec.nv * spec.nl, "paths", n, "err", err)
        assert err <= 1e-4, (spec.nv, spec.nl, err)
        assert cs_err(Ms, Ms0) <= 1e-4


def test_field_requests_alone_in_the_mix_and_against_mcfield():
    """Field requests (MCfield) served beside covariance requests: each field
    is the same alone and in the mix (1e-12), and equals the one-shot
    scmc_mcfield on the same paths (FP64 1e-12, mixed 2e-5, max |u - u*| /
    max |u*|: the slices chunk the paths differently)."""
    rng = np.random.default_rng(31)
    fields = [(E.make_problem(6.0, 0.97, [-5e4, -5e4, 0], [5e4, 5e4, 1], cone_dirs(nl, 10, rng),
                              cone_dirs(nv, 12, rng), 1, 1, 1.0, cbs, 2, 3, 0.8, seed=60 + k), n, prec)
              for k, (nv, nl, cbs, n, prec) in enumerate([(900, 1, False, 3000, "mixed"), (256, 3, True, 2000, "fp64"),
                                                          (4096, 1, True, 1500, "mixed")])]
    alone = []
    with E.Server(max_active=1, paths_per_slice=700) as S:
        for spec, n, prec in fields:
            alone.append(S.wait(S.submit(spec, n, prec, kind="field")))
    with E.Server(max_active=4, paths_per_slice=700) as S:
        ids = []
        for k, (spec, n, prec) in enumerate(fields):
            ids.append(S.submit(spec, n, prec, kind="field"))
            i, nv, nl, od, g, nc, pc = MIX[k]
            S.submit(req(i, nv, nl, od, g, seed=i), nc, pc)    # covariance requests in between
        say("field mix submitted")
        mixed = [S.wait(r) for r in ids]
    for (spec, n, prec), (ua, sta), (um, stm) in zip(fields, alone, mixed):
        assert ua.shape == (spec.nv, spec.nl) and stm["paths"] == n
        assert np.abs(um - ua).max() <= 1e-12 * np.abs(ua).max()
        u1 = E.mcfield_spec(spec, n, prec)
        tol = 2e-5 if prec == "mixed" else 1e-12
        err = np.abs(um - u1).max() / np.abs(u1).max()
        say("  field", spec.nv, spec.nl, prec, "err", err)
        assert err <= tol
