"""Pipelined batches: a set of problems split into parts, each part a forest
stepped on its own HIP stream.

A forest's EM step alternates bandwidth-bound launches (the big bottom levels
of the inside pass and the outside chain rounds) with latency-bound ones (the
small top levels in the tree walk, the first chain rounds, the M-step's
per-problem solves).  With two parts on two streams the GPU runs one part's
latency-bound launches while the other part's bandwidth-bound launches hold
the memory system.  Each part is an ordinary forest with its own launch plan,
so a problem's results are those of its part run on its own.
"""
import numpy as np
import torch

from .forest import Forest


def split_balanced(specs, parts):
    """Contiguous groups of problems with about equal (internal node x site)
    work; never an empty group."""
    parts = max(1, min(int(parts), len(specs)))
    w = np.array([sum(t.n_nodes for t in p.trees) * p.sites.K for p in specs], dtype=np.float64)
    cum = np.cumsum(w)
    cuts = [0]
    for j in range(1, parts):
        c = int(np.searchsorted(cum, cum[-1] * j / parts)) + 1
        cuts.append(min(max(c, cuts[-1] + 1), len(specs) - (parts - j)))
    cuts.append(len(specs))
    return [list(range(cuts[j], cuts[j + 1])) for j in range(parts)]


class PipelinedBatch:
    """Problems `specs` (laml_amd.forest.ProblemSpec) in `parts` forests, each
    with its own stream (default priority)."""

    def __init__(self, specs, parts=2, device=0):
        self.groups = split_balanced(specs, parts)
        self.device = device
        self.streams = [torch.cuda.Stream(device=device) for _ in self.groups]
        self.forests = []
        for g, s in zip(self.groups, self.streams):
            with torch.cuda.stream(s):
                self.forests.append(Forest([specs[i] for i in g], device=device))

    def params(self, lengths, nu, phi):
        """Per part (d, nu, phi) device tensors from per-problem start values:
        lengths[i] is the list of per-tree post-order length arrays of problem
        i, nu[i] and phi[i] its scalars."""
        out = []
        for F, g, s in zip(self.forests, self.groups, self.streams):
            with torch.cuda.stream(s):
                L = [x for i in g for x in lengths[i]]
                d = torch.tensor(F.lengths_to_level(L), dtype=torch.float64, device=self.device)
                n = torch.tensor([nu[i] for i in g], dtype=torch.float64, device=self.device)
                p = torch.tensor([phi[i] for i in g], dtype=torch.float64, device=self.device)
            out.append((d, n, p))
        self.synchronize()
        return out

    def autotune(self, params, **kw):
        """Autotune each part's launch plan (one part at a time)."""
        from .autotune import autotune
        plans = []
        for F, s, prm in zip(self.forests, self.streams, params):
            with torch.cuda.stream(s):
                plans.append(autotune(F, *prm, **kw)[0])
            s.synchronize()
        return plans

    def em(self, params, max_iter, tol):
        """EM on every part, each on its stream (the parts run concurrently);
        returns per part (nll, iters) device tensors, valid after
        synchronize().  With a convergence test (tol > 0) a part's EM loop
        waits on its stream every few steps to read the problems' done
        flags, so each part is driven from a host thread of its own; with a
        fixed number of steps (tol <= 0) the calls only enqueue work and run
        on the calling thread."""
        if tol <= 0 or len(self.forests) == 1:
            out = []
            for F, s, prm in zip(self.forests, self.streams, params):
                with torch.cuda.stream(s):
                    out.append(F.em(*prm, max_iter, tol))
            return out
        import threading

        out = [None] * len(self.forests)
        errors = []

        def run(j):
            try:
                torch.cuda.set_device(self.device)
                with torch.cuda.stream(self.streams[j]):
                    out[j] = self.forests[j].em(*params[j], max_iter, tol)
            except BaseException as e:  # re-raised on the calling thread
                errors.append(e)

        threads = [threading.Thread(target=run, args=(j,)) for j in range(len(self.forests))]
        for t in threads:
            t.start()
        for t in threads:
            t.join()
        if errors:
            raise errors[0]
        return out

    def synchronize(self):
        for s in self.streams:
            s.synchronize()

    def locate(self, i):
        """(part, index within the part) of problem i."""
        for j, g in enumerate(self.groups):
            if i in g:
                return j, g.index(i)
        raise IndexError(i)

    def close(self):
        self.synchronize()
        for F in self.forests:
            F.close()
        self.forests = []
