"""Scheduling logic of the serving loop on the CPU (a stand-in runner that
sleeps instead of solving): FIFO admission, the slot limit, the memory
budget, overlap of solves, failures that do not stop the server."""
import threading
import time

import pytest

from laml_amd.serving import PMMServer, SolveRequest, latency_summary


class _Spec:
    pass


def _req(rid, nbytes, work):
    r = SolveRequest.__new__(SolveRequest)
    r.spec = _Spec()
    r.spec.work = work
    r.rid = rid
    r.bytes = nbytes
    r.t_arrive = None
    return r


def test_budget_slots_and_fifo():
    lock = threading.Lock()
    running = []
    order = []
    peak = {"n": 0, "bytes": 0}

    def runner(req):
        with lock:
            running.append(req)
            order.append(req.rid)
            peak["n"] = max(peak["n"], len(running))
            peak["bytes"] = max(peak["bytes"], sum(r.bytes for r in running))
        time.sleep(req.spec.work)
        with lock:
            running.remove(req)
        return (float(req.rid), None, 0.0, 0.0, 1)

    with PMMServer(slots=3, hbm_budget=100, runner=runner) as srv:
        futs = [srv.submit(_req(i, b, 0.02)) for i, b in enumerate([60, 30, 30, 50, 10, 90, 20, 20, 40])]
        res = [f.result(timeout=30) for f in futs]
    assert [r.nll for r in res] == [float(i) for i in range(9)]
    assert order == list(range(9))            # FIFO admission
    assert peak["n"] <= 3 and peak["n"] >= 2   # solves overlap, within the slot limit
    assert peak["bytes"] <= 100 and srv.peak_in_use <= 100
    assert srv.in_use == 0 and srv.active == 0
    s = latency_summary(res)
    assert s["p99_ms"] >= s["p50_ms"] > 0


def test_request_larger_than_budget_is_rejected():
    with PMMServer(slots=1, hbm_budget=10, runner=lambda r: (0.0, None, 0.0, 0.0, 0)) as srv:
        with pytest.raises(ValueError):
            srv.submit(_req(0, 11, 0.0))


def test_failed_request_does_not_stop_the_server():
    def runner(req):
        if req.rid == 1:
            raise RuntimeError("boom")
        return (1.0, None, 0.0, 0.0, 1)

    with PMMServer(slots=2, hbm_budget=10, runner=runner) as srv:
        futs = [srv.submit(_req(i, 5, 0.0)) for i in range(4)]
        with pytest.raises(RuntimeError):
            futs[1].result(timeout=10)
        assert [futs[i].result(timeout=10).nll for i in (0, 2, 3)] == [1.0, 1.0, 1.0]


def test_late_arrivals_are_admitted_while_others_run():
    started = {}

    def runner(req):
        started[req.rid] = time.perf_counter()
        time.sleep(req.spec.work)
        return (0.0, None, 0.0, 0.0, 1)

    with PMMServer(slots=2, hbm_budget=100, runner=runner) as srv:
        f0 = srv.submit(_req(0, 10, 0.3))
        time.sleep(0.05)
        f1 = srv.submit(_req(1, 10, 0.01))   # arrives while request 0 runs
        r1 = f1.result(timeout=10)
        assert not f0.done()                  # it did not wait for request 0
        f0.result(timeout=10)
    assert started[1] < started[0] + 0.3
