"""Request-serving loop for PMM likelihood / EM solves on one GPU.

Requests (a problem: trees, character matrix, priors, start parameters, an
EM budget) arrive at any time.  The server keeps a FIFO queue and a fixed
number of execution slots, each a worker thread with its own HIP stream.  A
request is admitted when a slot is free and its estimated device footprint
fits in what is left of the HBM budget; its forest (inside/outside buffers,
counts, parameters) is built on the device and stays resident until its
solve ends, then the memory goes back to the budget and the next request is
admitted at once.  Solves of different sizes therefore overlap on the device
(kernels of several streams run concurrently), and a request that arrives
while others run starts as soon as a slot and memory free up.

Each request runs the same code path as a solve on its own (its own forest,
the same launches, on its own stream), so its result does not depend on what
else is being served; the tests check that bit for bit.
"""
import collections
import threading
import time
from concurrent.futures import Future

import numpy as np
import torch

from .forest import Forest


def segment_width(max_k):
    return 8 if max_k <= 8 else 16 if max_k <= 16 else 32 if max_k <= 32 else 64


def estimate_forest_bytes(specs):
    """Device bytes of a forest of these problems (rows padded to the segment
    width; per node metadata, partials and parameters; per site tables)."""
    K = [p.sites.K for p in specs]
    W = segment_width(max(K))
    chunks = -(-max(K) // W)
    total = 0
    nodes = 0
    for p in specs:
        kp = -(-p.sites.K // W) * W
        for t in p.trees:
            n_leaf = t.n_leaves
            n_int = t.n_nodes - n_leaf
            # leaf codes 2 B; internal: beta 16 + bE 4 + labels 2 + alpha 32 + aE 4 + aG 2
            total += n_leaf * kp * 2 + n_int * kp * 60
            nodes += t.n_nodes
        total += p.sites.K * (8 + 4 + 8 + 4 + 4) + 8 * int(p.sites.qoff[-1])
    total += nodes * (4 * 8 + 8 * 3 + 32 + 48 * chunks + 8 * 8)
    return int(total) + (1 << 20)


class SolveRequest:
    """One solve: a problem spec (laml_amd.forest.ProblemSpec) with start
    branch lengths per tree (post-order arrays), nu, phi and the EM budget."""

    _ids = iter(range(1 << 62))

    def __init__(self, spec, lengths, nu, phi, max_iter=1000, tol=1e-8, rid=None, nbytes=None):
        self.spec = spec
        self.lengths = [np.asarray(x, dtype=np.float64) for x in lengths]
        self.nu = float(nu)
        self.phi = float(phi)
        self.max_iter = int(max_iter)
        self.tol = float(tol)
        self.rid = next(SolveRequest._ids) if rid is None else rid
        self.bytes = estimate_forest_bytes([spec]) if nbytes is None else int(nbytes)
        self.t_arrive = None


class SolveResult:
    def __init__(self, rid, nll, lengths, nu, phi, iters, t_arrive, t_admit, t_done):
        self.rid = rid
        self.nll = nll
        self.lengths = lengths
        self.nu = nu
        self.phi = phi
        self.iters = iters
        self.t_arrive = t_arrive
        self.t_admit = t_admit
        self.t_done = t_done

    @property
    def latency(self):
        return self.t_done - self.t_arrive

    @property
    def wait(self):
        return self.t_admit - self.t_arrive


def solve_alone(req, device=0):
    """Run one request on the calling thread's current stream (the plain
    per-request path, and what every served request computes)."""
    F = Forest([req.spec], device=device)
    try:
        d = torch.tensor(F.lengths_to_level(req.lengths), dtype=torch.float64, device=F.device)
        nu = torch.tensor([req.nu], dtype=torch.float64, device=F.device)
        phi = torch.tensor([req.phi], dtype=torch.float64, device=F.device)
        nll, iters = F.em(d, nu, phi, req.max_iter, req.tol)
        out = (float(nll[0].item()), F.level_to_tree_lengths(d.cpu().numpy()), float(nu.item()),
               float(phi.item()), int(iters[0].item()))
    finally:
        F.close()
    return out


class PMMServer:
    """Continuous-batching server: ``slots`` concurrent solves, admission by
    estimated device footprint within ``hbm_budget`` bytes.

    ``runner(req)`` computes a request's result tuple (nll, lengths, nu, phi,
    iters); the default runs ``solve_alone`` on the slot's stream."""

    def __init__(self, device=0, slots=4, hbm_budget=None, runner=None, use_streams=True):
        self.device = device
        self.slots = int(slots)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(device)
            hbm_budget = int(free * 0.8)
        self.budget = int(hbm_budget)
        self.in_use = 0
        self.peak_in_use = 0
        self.active = 0
        self.peak_active = 0
        self._runner = runner
        self._use_streams = use_streams and runner is None
        self._queue = collections.deque()
        self._cv = threading.Condition()
        self._closing = False
        self._workers = [threading.Thread(target=self._work, name="pmm-slot-%d" % i, daemon=True)
                         for i in range(self.slots)]
        for w in self._workers:
            w.start()

    def submit(self, req):
        """Queue a request; returns a Future resolving to a SolveResult."""
        if req.bytes > self.budget:
            raise ValueError("request %r needs ~%d bytes of device memory, more than the server's budget %d"
                             % (req.rid, req.bytes, self.budget))
        fut = Future()
        req.t_arrive = time.perf_counter()
        with self._cv:
            if self._closing:
                raise RuntimeError("server is closed")
            self._queue.append((req, fut))
            self._cv.notify_all()
        return fut

    def _admit(self):
        # FIFO: the head request runs as soon as a slot and its memory are free
        while True:
            if self._closing and not self._queue:
                return None
            if self._queue and self.in_use + self._queue[0][0].bytes <= self.budget:
                req, fut = self._queue.popleft()
                self.in_use += req.bytes
                self.active += 1
                self.peak_in_use = max(self.peak_in_use, self.in_use)
                self.peak_active = max(self.peak_active, self.active)
                return req, fut
            self._cv.wait()

    def _work(self):
        stream = None
        if self._use_streams:
            torch.cuda.set_device(self.device)
            stream = torch.cuda.Stream(device=self.device)
        while True:
            with self._cv:
                item = self._admit()
            if item is None:
                return
            req, fut = item
            t_admit = time.perf_counter()
            try:
                if stream is not None:
                    with torch.cuda.stream(stream):
                        out = solve_alone(req, self.device)
                else:
                    out = self._runner(req)
                res = SolveResult(req.rid, out[0], out[1], out[2], out[3], out[4], req.t_arrive, t_admit,
                                  time.perf_counter())
                fut.set_result(res)
            except BaseException as e:  # the request fails, the server goes on
                fut.set_exception(e)
            finally:
                with self._cv:
                    self.in_use -= req.bytes
                    self.active -= 1
                    self._cv.notify_all()

    def close(self):
        """Finish every queued request, then stop the workers."""
        with self._cv:
            self._closing = True
            self._cv.notify_all()
        for w in self._workers:
            w.join()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def latency_summary(results):
    lat = np.asarray([r.latency for r in results])
    span = max(r.t_done for r in results) - min(r.t_arrive for r in results)
    return {"p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
            "mean_ms": float(lat.mean() * 1e3), "throughput_req_per_s": len(results) / span}
