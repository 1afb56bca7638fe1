"""Scheduling logic of the serving loop on the CPU (a stand-in runner that
sleeps instead of solving): FIFO admission, the slot limit, the memory
budget, overlap of solves, failures that do not stop the server, and the
batching of small requests."""
import threading
import time

import pytest

from laml_amd.serving import PMMServer, SolveRequest, latency_summary


class _Spec:
    pass


def _req(rid, nbytes, work, max_iter=1000, tol=1e-8, width=64):
    r = SolveRequest.__new__(SolveRequest)
    r.spec = _Spec()
    r.spec.work = work
    r.rid = rid
    r.bytes = nbytes
    r.max_iter = max_iter
    r.tol = tol
    r.width = width
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


def _batching_server(slots, budget, batch_bytes, batch_max, gate):
    """A server whose single and batch runners wait on `gate` and record
    the rids of each solve."""
    solves = []
    lock = threading.Lock()

    def runner(req):
        gate.wait(10)
        with lock:
            solves.append([req.rid])
        return (float(req.rid), None, 0.0, 0.0, 1)

    def batch_runner(reqs):
        gate.wait(10)
        with lock:
            solves.append([r.rid for r in reqs])
        return [(float(r.rid), None, 0.0, 0.0, 1) for r in reqs]

    srv = PMMServer(slots=slots, hbm_budget=budget, runner=runner, batch_runner=batch_runner,
                    batch_bytes=batch_bytes, batch_max=batch_max)
    return srv, solves


def test_small_requests_are_batched_in_queue_order():
    gate = threading.Event()
    srv, solves = _batching_server(slots=1, budget=1000, batch_bytes=100, batch_max=3, gate=gate)
    with srv:
        first = srv.submit(_req(0, 500, 0.0))      # holds the slot until the gate opens
        time.sleep(0.05)
        # small: bytes <= 25; 4 and 7 are large, 5 has another EM budget
        sizes = {1: 10, 2: 20, 3: 25, 4: 30, 5: 10, 6: 5, 7: 200, 8: 5, 9: 20, 10: 20}
        futs = {}
        for rid in range(1, 11):
            kw = {"max_iter": 50} if rid == 5 else {}
            futs[rid] = srv.submit(_req(rid, sizes[rid], 0.0, **kw))
        gate.set()
        res = {rid: f.result(timeout=10) for rid, f in futs.items()}
        first.result(timeout=10)
    assert all(res[r].nll == float(r) for r in res)
    # head 1 takes 2 and 3 (batch_max 3); 4 is large and runs alone; head 5
    # (max_iter 50) finds no partner; head 6 takes 8 and 9 (7 is large,
    # 10 would pass batch_max); then 7, then 10
    assert solves == [[0], [1, 2, 3], [4], [5], [6, 8, 9], [7], [10]]
    assert srv.batches == [1, 3, 1, 1, 3, 1, 1]
    assert srv.in_use == 0 and srv.active == 0


def test_batch_respects_byte_cap_and_budget():
    gate = threading.Event()
    srv, solves = _batching_server(slots=2, budget=60, batch_bytes=40, batch_max=16, gate=gate)
    with srv:
        futs = [srv.submit(_req(i, 10, 0.0)) for i in range(8)]
        gate.set()
        for f in futs:
            f.result(timeout=10)
    assert sorted(r for b in solves for r in b) == list(range(8))
    assert all(len(b) * 10 <= 40 for b in solves)
    assert srv.peak_in_use <= 60


def test_failed_batch_fails_its_requests_only():
    def batch_runner(reqs):
        if any(r.rid == 1 for r in reqs):
            raise RuntimeError("boom")
        return [(1.0, None, 0.0, 0.0, 1) for _ in reqs]

    gate = threading.Event()

    def runner(req):
        gate.wait(10)
        return (1.0, None, 0.0, 0.0, 1)

    with PMMServer(slots=1, hbm_budget=100, runner=runner, batch_runner=batch_runner, batch_bytes=40,
                   batch_max=2) as srv:
        f0 = srv.submit(_req(0, 50, 0.0))
        time.sleep(0.05)
        futs = [srv.submit(_req(i, 10, 0.0)) for i in range(1, 5)]
        gate.set()
        f0.result(timeout=10)
        for f in futs[:2]:
            with pytest.raises(RuntimeError):
                f.result(timeout=10)
        assert [f.result(timeout=10).nll for f in futs[2:]] == [1.0, 1.0]


def test_batches_keep_one_row_width():
    """Requests whose rows pad to different widths are not solved together
    (a shared forest would pad every row to the widest)."""
    gate = threading.Event()
    srv, solves = _batching_server(slots=1, budget=1000, batch_bytes=100, batch_max=8, gate=gate)
    with srv:
        first = srv.submit(_req(0, 500, 0.0))
        time.sleep(0.05)
        widths = {1: 64, 2: 32, 3: 64, 4: 32, 5: 8}
        futs = [srv.submit(_req(r, 10, 0.0, width=w)) for r, w in widths.items()]
        gate.set()
        for f in futs:
            f.result(timeout=10)
        first.result(timeout=10)
    assert solves == [[0], [1, 3], [2, 4], [5]]
