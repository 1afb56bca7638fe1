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

With batching on (``batch_bytes`` > 0), a slot whose next request is small
(footprint at most ``batch_bytes`` / 4) also takes the other small queued
requests with the same EM budget and the same row width (sites padded to 8,
16, 32 or 64; a shared forest pads every row to its widest, so at one width
its footprint is the sum of the requests'), in queue order, up to
``batch_bytes`` and ``batch_max`` requests, and solves them as one
device-resident forest: one
problem per request, one launch sequence for all of them, each problem
stopping at its own convergence test.  Small requests then share launches
instead of each paying a full launch sequence; each result equals the
request solved alone to within the engine's tolerance (the launch shapes of
a shared forest change the order of floating-point sums over sites).
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


# Device memory a forest takes beyond estimate_forest_bytes, whatever its
# size: allocation rounding and the caching allocator's segments of the
# forest's small tensors (measured at 2-37 MB); the server reserves it once
# per slot.
FOREST_FIXED_BYTES = 48 << 20


def estimate_forest_bytes(specs):
    """Device bytes of a forest of these problems that grow with it (rows
    padded to the segment width; per node metadata, records, chain plans,
    partials and parameters; per site tables); with FOREST_FIXED_BYTES, an
    upper bound on what its creation and an EM call take."""
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
            # leaf codes 2 B; internal: beta 16 + bE 4 + labels 2 + alpha 32 + aE 4 + aG 2;
            # with 64-site segments every node also has a prior-weight row (8 B)
            q = 8 if W == 64 else 0
            total += n_leaf * kp * (2 + q) + n_int * kp * (60 + q)
            nodes += t.n_nodes
        total += p.sites.K * (8 + 4 + 8 + 4 + 4) + 8 * int(p.sites.qoff[-1])
    # per node: coefficients 40, counts 40, nu solve 24, node record 64,
    # ids and offsets 48, chain plan 48; per 64-site chunk: partial counts 40
    # and dropout sum 8
    total += nodes * (264 + 48 * chunks)
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
        # rows are padded to this many sites; requests share a forest only
        # at the same width, so that its footprint is the sum of theirs
        self.width = segment_width(spec.sites.K)
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


def solve_batch(reqs, device=0):
    """Solve requests with the same EM budget as one forest (a problem per
    request) on the calling thread's current stream; a result tuple per
    request as ``solve_alone`` returns it."""
    if len(reqs) == 1:
        return [solve_alone(reqs[0], device)]
    max_iter, tol = reqs[0].max_iter, reqs[0].tol
    if any(r.max_iter != max_iter or r.tol != tol for r in reqs):
        raise ValueError("a batch shares one EM budget")
    F = Forest([r.spec for r in reqs], device=device)
    try:
        d = torch.tensor(F.lengths_to_level([x for r in reqs for x in r.lengths]), dtype=torch.float64,
                         device=F.device)
        nu = torch.tensor([r.nu for r in reqs], dtype=torch.float64, device=F.device)
        phi = torch.tensor([r.phi for r in reqs], dtype=torch.float64, device=F.device)
        nll, iters = F.em(d, nu, phi, max_iter, tol)
        nll_h, it_h = nll.cpu().numpy(), iters.cpu().numpy()
        nu_h, phi_h = nu.cpu().numpy(), phi.cpu().numpy()
        per_tree = F.level_to_tree_lengths(d.cpu().numpy())
    finally:
        F.close()
    out, t = [], 0
    for i, r in enumerate(reqs):
        nt = len(r.spec.trees)
        out.append((float(nll_h[i]), per_tree[t:t + nt], float(nu_h[i]), float(phi_h[i]), int(it_h[i])))
        t += nt
    return out


class PMMServer:
    """Continuous-batching server: ``slots`` concurrent solves, admission by
    estimated device footprint within ``hbm_budget`` bytes, and (with
    ``batch_bytes`` > 0) small requests solved together as one forest.

    ``runner(req)`` computes a request's result tuple (nll, lengths, nu, phi,
    iters); the default runs ``solve_alone`` on the slot's stream.
    ``batch_runner(reqs)`` computes the tuples of a batch; the default runs
    ``solve_batch`` on the slot's stream."""

    def __init__(self, device=0, slots=4, hbm_budget=None, runner=None, use_streams=True, batch_bytes=0,
                 batch_max=16, batch_runner=None):
        self.device = device
        self.slots = int(slots)
        self.batch_bytes = int(batch_bytes)
        self.batch_max = max(1, int(batch_max))
        self._batch_runner = batch_runner
        self.batches = []  # sizes of the batches solved
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(device)
            hbm_budget = int(free * 0.8) - self.slots * FOREST_FIXED_BYTES
        self.budget = int(hbm_budget)
        self.in_use = 0
        self.peak_in_use = 0
        self.active = 0
        self.peak_active = 0
        self._runner = runner
        self._use_streams = use_streams and runner is None and batch_runner is None
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

    def _small(self, req):
        return self.batch_bytes > 0 and req.bytes * 4 <= self.batch_bytes

    def _admit(self):
        # FIFO: the head request runs as soon as a slot and its memory are
        # free; a small head takes the other small queued requests with its
        # EM budget and row width along (in queue order), within the batch
        # limits
        while True:
            if self._closing and not self._queue:
                return None
            if self._queue and self.in_use + self._queue[0][0].bytes <= self.budget:
                req, fut = self._queue.popleft()
                items = [(req, fut)]
                total = req.bytes
                if self._small(req):
                    keep = collections.deque()
                    width = getattr(req, "width", None)
                    while self._queue:
                        r2, f2 = self._queue.popleft()
                        if (len(items) < self.batch_max and self._small(r2) and r2.max_iter == req.max_iter
                                and r2.tol == req.tol and getattr(r2, "width", None) == width
                                and total + r2.bytes <= self.batch_bytes
                                and self.in_use + total + r2.bytes <= self.budget):
                            items.append((r2, f2))
                            total += r2.bytes
                        else:
                            keep.append((r2, f2))
                    self._queue = keep
                self.in_use += total
                self.active += len(items)
                self.peak_in_use = max(self.peak_in_use, self.in_use)
                self.peak_active = max(self.peak_active, self.active)
                return items
            self._cv.wait()

    def _work(self):
        stream = None
        if self._use_streams:
            torch.cuda.set_device(self.device)
            stream = torch.cuda.Stream(device=self.device)
        while True:
            with self._cv:
                items = self._admit()
            if items is None:
                return
            reqs = [r for r, _ in items]
            t_admit = time.perf_counter()
            try:
                if len(reqs) > 1:
                    if self._batch_runner is not None:
                        outs = self._batch_runner(reqs)
                    elif stream is not None:
                        with torch.cuda.stream(stream):
                            outs = solve_batch(reqs, self.device)
                    else:
                        outs = [self._runner(r) for r in reqs]
                elif stream is not None:
                    with torch.cuda.stream(stream):
                        outs = [solve_alone(reqs[0], self.device)]
                else:
                    outs = [self._runner(reqs[0])]
                t_done = time.perf_counter()
                for (req, fut), out in zip(items, outs):
                    fut.set_result(SolveResult(req.rid, out[0], out[1], out[2], out[3], out[4], req.t_arrive,
                                               t_admit, t_done))
            except BaseException as e:  # the requests fail, the server goes on
                for _, fut in items:
                    if not fut.done():
                        fut.set_exception(e)
            finally:
                with self._cv:
                    self.batches.append(len(items))
                    self.in_use -= sum(r.bytes for r in reqs)
                    self.active -= len(items)
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
