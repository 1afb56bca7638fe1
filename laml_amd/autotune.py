"""Launch-plan autotuner for a device forest.

A forest's passes can run as level-synchronous launches (per level: the
number of 64-site chunks one wavefront walks, whether the small top levels
share one fused launch, and whether the cherries -- nodes with two leaf
children -- are fused into their parents), as the tree walk (one launch per
pass, a block per (tree, site chunk)), or as a hybrid of the two (level
launches for the big bottom levels, one walk for the small top ones); the
levels below the split can run as heavy-path chains in either pass.  Which is
fastest depends on the forest: how
many trees, how deep, how many sites, how uneven.  ``autotune`` times the
E-step (inside pass, root, outside pass, counts) under candidate plans on
the forest itself and keeps the fastest; every plan computes the same values
up to the order of floating-point sums over sites.

The search is a coarse pass over pass mode (with a few hybrid splits),
top-level fusion, cherry fusion and chains, then a coordinate search
over the chunks per wavefront of the largest levels.
"""
import numpy as np
import torch

from ._lib import CHAINS_INSIDE, CHAINS_OUTSIDE, PASS_HYBRID, PASS_LEVELS, PASS_TREES


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
    cherry fusion, chain flags (CHAINS_OUTSIDE | CHAINS_INSIDE)."""

    def __init__(self, mode, inside_cpw, outside_cpw, top_level, fuse=True, chains=0, chain_tops=(5, 10)):
        self.mode = int(mode)
        self.chain_tops = tuple(int(x) for x in chain_tops)
        self.inside_cpw = np.asarray(inside_cpw, dtype=np.int32).copy()
        self.outside_cpw = np.asarray(outside_cpw, dtype=np.int32).copy()
        self.top_level = int(top_level)
        self.fuse = bool(fuse)
        self.chains = int(chains)

    def copy(self):
        return Plan(self.mode, self.inside_cpw, self.outside_cpw, self.top_level, self.fuse, self.chains,
                    self.chain_tops)

    def apply(self, F):
        F.set_pass_mode(PASS_LEVELS)
        F.set_shapes(self.inside_cpw, self.outside_cpw, self.top_level)
        F.set_fusion(self.fuse)
        F.set_chains(self.chains, *self.chain_tops)
        F.set_pass_mode(self.mode)

    def describe(self):
        if self.mode == PASS_TREES:
            return {"mode": "trees"}
        return {"mode": "hybrid" if self.mode == PASS_HYBRID else "levels", "top_level": self.top_level,
                "cherry_fusion": self.fuse, "outside_chains": bool(self.chains & CHAINS_OUTSIDE),
                "inside_chains": bool(self.chains & CHAINS_INSIDE), "chain_tops": list(self.chain_tops),
                "inside_cpw": self.inside_cpw.tolist(),
                "outside_cpw": self.outside_cpw.tolist()}


def plan_from_dict(d, F):
    """The plan a ``Plan.describe()`` dictionary names, for forest F (for
    replaying a tuned plan on another run of the same forest)."""
    if d["mode"] == "trees":
        base = current_plan(F)
        return Plan(PASS_TREES, base.inside_cpw, base.outside_cpw, base.top_level, False)
    mode = PASS_HYBRID if d["mode"] == "hybrid" else PASS_LEVELS
    chains = (CHAINS_OUTSIDE if d.get("outside_chains") else 0) | (CHAINS_INSIDE if d.get("inside_chains") else 0)
    return Plan(mode, d["inside_cpw"], d["outside_cpw"], d["top_level"], d.get("cherry_fusion", False), chains,
                d.get("chain_tops", (5, 10)))


def current_plan(F):
    a, b, top = F.get_shapes()
    mode = F.pass_mode()
    F.set_pass_mode(PASS_LEVELS)
    fuse = F.fusion()
    F.set_pass_mode(mode)
    chains, ti, to = F.get_chains()
    return Plan(mode, a, b, top, fuse, chains, (ti, to))


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
    # the same level and hybrid plans with and without outside chains (row
    # kernels only)
    rows = base.inside_cpw[1:].max() > 0  # row kernels in use (more than 32 sites)
    if rows:
        chained = []
        for tag, plan in [("default", base)] + cands:
            if plan.mode == PASS_TREES or plan.fuse:
                continue
            c = plan.copy()
            c.chains = plan.chains ^ CHAINS_OUTSIDE
            chained.append(((tag + ", " if tag != "default" else "") +
                            ("outside chains" if c.chains & CHAINS_OUTSIDE else "no outside chains"), c))
        cands += chained
    for tag, plan in cands:
        ms = measure(plan, tag)
        if ms < best_ms:
            best, best_ms = plan, ms
    # the outside chain-region top, then inside chains on top of the plan.
    # The plans differ by a few per cent, about the spread of one timing, and
    # the incumbent's time is the least of many: it is timed afresh and every
    # plan here keeps the lesser of two timings.
    if rows and best.mode != PASS_TREES and not best.fuse:
        def measure2(plan, tag):
            return min(measure(plan, tag), measure(plan, tag))

        best_ms = measure2(best, "incumbent, timed again")
        trials = []
        if best.chains & CHAINS_OUTSIDE:
            trials += [(best.chains, (best.chain_tops[0], t)) for t in (0, 12, 10, 8, 6) if t != best.chain_tops[1]]
        for tag_tops in trials:
            trial = best.copy()
            trial.chains, trial.chain_tops = tag_tops
            ms = measure2(trial, "chain_tops=%s" % (trial.chain_tops,))
            if ms < best_ms * 0.995:
                best, best_ms = trial, ms
        for t in (4, 6):
            trial = best.copy()
            trial.chains = best.chains | CHAINS_INSIDE
            trial.chain_tops = (t, best.chain_tops[1])
            ms = measure2(trial, "inside chains, top %d" % t)
            if ms < best_ms * 0.995:
                best, best_ms = trial, ms
    if best.mode in (PASS_LEVELS, PASS_HYBRID) and rows:
        # the trials below are single timings: so is the incumbent's
        best_ms = measure(best, "incumbent, timed again")
        sizes = np.diff(F.level_off)
        first = 2 if best.fuse else 1  # the cherry level has no launch of its own when fused
        big = [h for h in np.argsort(-sizes) if first <= h < min(H, best.top_level)][:max_levels]
        max_chunks = int(max(best.inside_cpw.max(), best.outside_cpw.max(), 1))
        options = sorted({1, 2, max(1, max_chunks // 2), max_chunks})
        # a chained pass runs the levels below its chain-region top one
        # chunk per wavefront; the levels above keep their launch shapes
        split = min(H, best.top_level)
        for which, bit, k in (("outside_cpw", CHAINS_OUTSIDE, 1), ("inside_cpw", CHAINS_INSIDE, 0)):
            lo = 1
            if best.chains & bit:
                t = best.chain_tops[k]
                lo = split if t <= 0 else min(t, split)
            for h in [h for h in big if h >= lo]:
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
