"""Launch-plan autotuner for a device forest.

A forest's passes can run as level-synchronous launches (per level: the
number of 64-site chunks one wavefront walks, whether the small top levels
share one fused launch, and whether the cherries -- nodes with two leaf
children -- are fused into their parents), as the tree walk (one launch per
pass, a block per (tree, site chunk)), or as a hybrid of the two (level
launches for the big bottom levels, one walk for the small top ones); the
outside levels below the split can run as heavy-path chains.  Which is
fastest depends on the forest: how
many trees, how deep, how many sites, how uneven.  ``autotune`` times the
E-step (inside pass, root, outside pass, counts) under candidate plans on
the forest itself and keeps the fastest; every plan computes the same values
up to the order of floating-point sums over sites.

The search is a coarse pass over pass mode (with a few hybrid splits),
top-level fusion, cherry fusion and outside chains, then a coordinate search
over the chunks per wavefront of the largest levels.
"""
import numpy as np
import torch

from ._lib import PASS_HYBRID, PASS_LEVELS, PASS_TREES


def _time_estep(F, d, nu, phi, reps):
    F.estep(d, nu, phi)
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(reps):
        F.estep(d, nu, phi)
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / reps


class Plan:
    """A launch plan: pass mode, per-level chunks per wavefront, top level,
    cherry fusion, outside chains."""

    def __init__(self, mode, inside_cpw, outside_cpw, top_level, fuse=True, chains=False):
        self.mode = int(mode)
        self.inside_cpw = np.asarray(inside_cpw, dtype=np.int32).copy()
        self.outside_cpw = np.asarray(outside_cpw, dtype=np.int32).copy()
        self.top_level = int(top_level)
        self.fuse = bool(fuse)
        self.chains = bool(chains)

    def copy(self):
        return Plan(self.mode, self.inside_cpw, self.outside_cpw, self.top_level, self.fuse, self.chains)

    def apply(self, F):
        F.set_pass_mode(PASS_LEVELS)
        F.set_shapes(self.inside_cpw, self.outside_cpw, self.top_level)
        F.set_fusion(self.fuse)
        F.set_chains(self.chains)
        F.set_pass_mode(self.mode)

    def describe(self):
        if self.mode == PASS_TREES:
            return {"mode": "trees"}
        return {"mode": "hybrid" if self.mode == PASS_HYBRID else "levels", "top_level": self.top_level,
                "cherry_fusion": self.fuse, "outside_chains": self.chains, "inside_cpw": self.inside_cpw.tolist(),
                "outside_cpw": self.outside_cpw.tolist()}


def current_plan(F):
    a, b, top = F.get_shapes()
    mode = F.pass_mode()
    F.set_pass_mode(PASS_LEVELS)
    fuse = F.fusion()
    F.set_fusion(False)
    chains = F.chains()
    F.set_fusion(fuse)
    F.set_pass_mode(mode)
    return Plan(mode, a, b, top, fuse, chains)


def autotune(F, d, nu, phi, reps=3, max_levels=8, verbose=False):
    """Time candidate plans on forest F with parameters (d, nu, phi) and apply
    the fastest.  Returns (plan, log) where log lists (description, ms)."""
    H = F.num_levels
    base = current_plan(F)
    log = []

    def measure(plan, tag):
        plan.apply(F)
        ms = _time_estep(F, d, nu, phi, reps)
        log.append((tag, ms))
        if verbose:
            print("autotune %-28s %.4f ms" % (tag, ms), flush=True)
        return ms

    best = base
    best_ms = measure(base, "default")
    top0 = base.top_level
    cands = [("levels, fused top", Plan(PASS_LEVELS, base.inside_cpw, base.outside_cpw, top0, False)),
             ("levels", Plan(PASS_LEVELS, base.inside_cpw, base.outside_cpw, H, False)),
             ("tree walk", Plan(PASS_TREES, base.inside_cpw, base.outside_cpw, top0, False))]
    # cherry fusion, where the forest allows it (level mode, no fused top)
    mode = F.pass_mode()
    F.set_pass_mode(PASS_LEVELS)
    F.set_fusion(True)
    can_fuse = F.fusion()
    F.set_fusion(base.fuse)
    F.set_pass_mode(mode)
    if can_fuse:
        cands.append(("cherry fusion", Plan(PASS_LEVELS, base.inside_cpw, base.outside_cpw, H, True)))
    # hybrid: the levels from a split up in one tree-walk launch; splits where
    # the levels get small
    sizes = np.diff(F.level_off)
    splits = set()
    for cap in (64, 256, 1024, 4096):
        h = H
        while h - 1 >= 2 and sizes[h - 1] <= cap:
            h -= 1
        if 2 <= h < H:
            splits.add(int(h))
    for h in sorted(splits):
        if base.mode == PASS_HYBRID and h == base.top_level:
            continue
        cands.append(("hybrid from level %d" % h, Plan(PASS_HYBRID, base.inside_cpw, base.outside_cpw, h, False)))
    # the same level and hybrid plans with the outside levels below the
    # split as heavy-path chains (row kernels only)
    rows = base.inside_cpw[1:].max() > 0  # row kernels in use (more than 32 sites)
    if rows:
        chained = []
        for tag, plan in [("default", base)] + cands:
            if plan.mode == PASS_TREES or plan.fuse:
                continue
            c = plan.copy()
            c.chains = not plan.chains
            chained.append(((tag + ", " if tag != "default" else "") + ("chains" if c.chains else "no chains"), c))
        cands += chained
    for tag, plan in cands:
        ms = measure(plan, tag)
        if ms < best_ms:
            best, best_ms = plan, ms
    if best.mode in (PASS_LEVELS, PASS_HYBRID) and rows:
        sizes = np.diff(F.level_off)
        first = 2 if best.fuse else 1  # the cherry level has no launch of its own when fused
        big = [h for h in np.argsort(-sizes) if first <= h < min(H, best.top_level)][:max_levels]
        max_chunks = int(max(best.inside_cpw.max(), best.outside_cpw.max(), 1))
        options = sorted({1, 2, max(1, max_chunks // 2), max_chunks})
        # chains run the outside levels below the split one chunk per wavefront
        for which in ("inside_cpw",) if best.chains else ("outside_cpw", "inside_cpw"):
            for h in big:
                cur = int(getattr(best, which)[h])
                for c in options:
                    if c == cur:
                        continue
                    trial = best.copy()
                    getattr(trial, which)[h] = c
                    ms = measure(trial, "%s[%d]=%d" % (which, h, c))
                    if ms < best_ms * 0.995:
                        best, best_ms = trial, ms
    best.apply(F)
    return best, log
