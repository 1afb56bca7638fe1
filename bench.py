"""Benchmark: fixed-topology PMM EM on one MI355X.

Workload (default): B = 16 independent problems resident in HBM, each a
seeded 10,000-leaf Yule tree (height 1) x 300 sites, alphabet sizes uniform
in [5, 60], Dirichlet(1) priors, characters simulated with nu = 0.1,
phi = 0.1; EM from a uniform start (all branches 0.05, nu = phi = 0.5).  One
step is one full EM iteration of every problem of the batch (inside pass,
outside pass with posterior counts, device M-step of branch lengths, nu and
phi), all FP64, in shared level-synchronous launches.

One JSON line on stdout: aggregate EM iterations/s as ``value`` (B x steps /
time), the same for a single problem alone, node-site evaluations/s, the HBM
roofline of the dominant kernel family (the outside pass) and a CPU baseline
(the oracle on a bounded sample of the same workload).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E

WORKLOADS = {
    # name: (n leaves, K sites, M lo, M hi, dirichlet alpha, nu, phi, seed)
    "em10k": (10000, 300, 5, 60, 1.0, 0.1, 0.1, 20261206),
    "em1k": (1000, 100, 2, 100, 0.5, 0.05, 0.2, 20261204),
    "em250": (250, 30, 5, 50, 1.0, 0.1, 0.1, 20261203),
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def heavy_children(F):
    """Heavy child of every node in level order: the internal child with more
    leaves below it (the first child on ties), -1 without an internal child
    -- the chains of the outside pass follow these."""
    N, n = F.num_nodes, F.num_leaves
    nl = np.ones(N, dtype=np.int64)
    heavy = np.full(N, -1, dtype=np.int64)
    for h in range(1, F.num_levels):
        v = np.arange(F.level_off[h], F.level_off[h + 1])
        a, b = F.child0[v], F.child1[v]
        nl[v] = nl[a] + nl[b]
        ia, ib = a >= n, b >= n
        heavy[v] = np.where(ia & ib, np.where(nl[b] > nl[a], b, a), np.where(ia, a, np.where(ib, b, -1)))
    return heavy


def algorithmic_bytes(F, fused=False, chain_split=None, in_chain_split=None):
    """Compulsory HBM bytes of one inside pass and one outside pass.

    Per (internal node, site): inside reads the node's label (2 B) and each
    child's state (leaf: int16 code, 2 B; internal: label + beta pair +
    exponent, 22 B) and writes beta + exponent (20 B).  Outside reads the
    node's outside state (alpha quad + exponent + specific state, 38 B), each
    child's inside state again (2 / 22 B) and writes each internal child's
    outside state (38 B).  With fused cherries (nodes whose children are both
    leaves) a cherry has no rows: its parent reads the two leaf codes (4 B)
    in place of its state, and the cherry's own steps move nothing else.
    With outside chains below level chain_split, a heavy child's outside
    state stays in registers: it is neither written by its parent nor read by
    itself (38 + 38 B less per heavy child and site); with inside chains
    below level in_chain_split its parent does not read its inside state
    (22 B less)."""
    N, n = F.num_nodes, F.num_leaves
    K = F.tree_K[F.tree_of].astype(np.int64)
    lo, hi = (int(F.level_off[1]), int(F.level_off[2])) if fused else (n, n)
    inter = np.arange(hi if fused else n, N)
    c = np.stack([F.child0[inter], F.child1[inter]])
    cherry = (c >= lo) & (c < hi)
    rc = np.where(c < n, 2, np.where(cherry, 4, 22))
    wo = np.where((c < n) | cherry, 0, 38)
    ra = np.full(len(inter), 38)
    if chain_split is not None:
        B = int(F.level_off[chain_split])
        heavy = heavy_children(F)
        below = inter < B
        kept = (c == heavy[inter][None, :]) & below[None, :]
        wo = np.where(kept, 0, wo)
        # a node below the split reads its own state only as a chain head
        par = np.full(N, -1, dtype=np.int64)
        par[F.child0[inter]] = inter
        par[F.child1[inter]] = inter
        p = par[inter]
        head = (p < 0) | (p >= B) | (heavy[np.maximum(p, 0)] != inter)
        ra = np.where(below & ~head, 0, ra)
    rci = rc
    if in_chain_split is not None:
        Bi = int(F.level_off[in_chain_split])
        heavy = heavy_children(F)
        rci = np.where((c == heavy[inter][None, :]) & (inter < Bi)[None, :], 0, rc)
    Kv = K[inter]
    inside = int((Kv * (2 + 20 + rci.sum(0))).sum())
    outside = int((Kv * (ra + rc.sum(0) + wo.sum(0))).sum())
    return inside, outside


PROFILE = "profiles/outside_pass_em10k_batch16.md"


def outside_traffic_from_profile(args):
    """HBM bytes of one outside pass of this workload, as measured with
    rocprofv3 FETCH_SIZE / WRITE_SIZE passes (scaled by a plain copy's
    counts), and the vector-ALU issue fraction of the pass (SQ_INSTS_VALU,
    GRBM_GUI_ACTIVE), as recorded in the committed profile summary; None
    where absent."""
    import re
    path = os.path.join(ROOT, PROFILE)
    if not os.path.exists(path):
        return None, None
    txt = open(path).read()
    m = re.search(r"measured traffic total \| ([\d.]+) ", txt)
    v = re.search(r"valu issue fraction of the pass \| ([\d.]+)", txt)
    return (int(float(m.group(1)) * 1e9) if m else None), (float(v.group(1)) if v else None)


def profile_plan():
    """The launch plan the committed profile summary was measured with (the
    roofline leg applies it, so that the profile's counters describe the
    launches it times); None if absent."""
    import re
    path = os.path.join(ROOT, PROFILE)
    if not os.path.exists(path):
        return None
    m = re.search(r"applies it to its roofline leg\): `(\{.*?\})`", open(path).read())
    return json.loads(m.group(1)) if m else None


def cpu_baseline_python(tree, Q, codes, target_s=4.0):
    """Time the scalar Python oracle (a restatement of EM_solver with the
    reference's own per-(node, site) Python loops) on a bounded sample: the
    same tree with the first k sites, one EM iteration (E-step incl. inside
    pass, M-step, NLL).  The E-step part is scaled to all sites; the M-step
    (independent of K) is added once.  Returns (EM iters/s, sample text)."""
    from laml_amd import sim
    from oracle.em_oracle import OracleEM, postorder

    newick = tree.newick()
    msa_full = sim.codes_to_charmtrx(tree, codes)
    K = len(Q)
    ks = 2
    while True:
        msa = {c: r[:ks] for c, r in msa_full.items()}
        o = OracleEM([newick], msa, Q[:ks], nu=0.5, phi=0.5)
        for v in postorder(o.trees[0]):
            v.edge_length = 0.05
        o.az_partition()
        t0 = time.perf_counter()
        o.estep()
        t1 = time.perf_counter()
        o.mstep()
        t2 = time.perf_counter()
        o.lineage_llh()
        t3 = time.perf_counter()
        e_time = (t1 - t0) + (t3 - t2)
        m_time = t2 - t1
        if e_time + m_time > target_s / 3 or ks >= K:
            break
        ks = min(K, ks * 2)
    per_iter = e_time * (K / ks) + m_time
    return 1.0 / per_iter, ("scalar Python oracle, 1 core, first %d of %d sites, E-step %.2f s scaled x%.1f, "
                            "M-step %.2f s" % (ks, K, e_time, K / ks, m_time))


def cpu_baseline(tree, Q, codes, target_s=12.0, max_iters=60):
    """Time the C/OpenMP restatement of the EM oracle (oracle/pmm_ref.c: the
    reference's log-space E-step, exact M-step) on the full first problem of
    the workload, from the same start as the device (all branches 0.05,
    nu = phi = 0.5), with OpenMP over sites on every thread the host gives:
    whole EM iterations (E-step, M-step, NLL) until about target_s of CPU
    work.  The scalar Python oracle is timed beside it on a sample."""
    from oracle import pmm_ref

    K = len(Q)
    p = pmm_ref.RefProblem(tree.c0, tree.c1, tree.parent, codes, Q, None, 1.0 / (K * codes.shape[0]))
    d = np.full(tree.n_nodes, 0.05)
    nu, phi = 0.5, 0.5
    p.nll(d, nu, phi)  # page in, start the thread pool
    t0 = time.perf_counter()
    it = 0
    while True:
        _, counts, R, Rt = p.estep(d, nu, phi)
        d, nu, phi = p.mstep(counts, R, Rt, d, nu, phi)
        nll = p.nll(d, nu, phi)
        it += 1
        el = time.perf_counter() - t0
        if el > target_s or it >= max_iters:
            break
    threads = pmm_ref.threads()
    py_rate, py_sample = cpu_baseline_python(tree, Q, codes)
    return {
        "value": round(it / el, 4),
        "unit": "EM iters/s",
        "cores": threads,
        "kind": "port",
        "sample": "C/OpenMP restatement of EM_solver (oracle/pmm_ref.c), %d threads over sites, problem 0 of the "
                  "workload (%d leaves x %d sites) from the device's start, %d whole EM iterations in %.2f s "
                  "(NLL after them %.6f); one problem at a time, so the rate is also the batch's per-problem rate"
                  % (threads, (tree.n_nodes + 1) // 2, K, it, el, nll),
        "python_oracle": {"value": round(py_rate, 5), "unit": "EM iters/s", "cores": 1, "sample": py_sample},
    }


SCORING = {
    # batched candidate scoring: NNI variants of one tree plus uneven
    # independent trees, all with the same number of sites
    "nni768": dict(base_n=2000, variants=512, others=256, n_lo=100, n_hi=5000, K=200, m_lo=5, m_hi=50,
                   alpha=1.0, nu=0.1, phi=0.15, seed=20261208),
}


SERVING = {
    # Poisson stream of likelihood + EM requests of mixed sizes
    "serve": dict(requests=48, n_lo=200, n_hi=50000, K_lo=20, K_hi=500, nu=(0.02, 0.2), phi=(0.05, 0.3),
                  em_iters=50, tol=1e-8, slots=4, loads=(0.5, 1.0, 2.0), seed=20261209,
                  batch_bytes=512 << 20, batch_max=8),
}


def bench_serving(args):
    """Serving workload: requests arrive as a Poisson stream at a rate set
    relative to the plain path's capacity; the continuous-batching server
    (concurrent slots, admission by device footprint) against the plain path
    (one request at a time, in arrival order)."""
    import threading

    import torch

    from laml_amd import sim
    from laml_amd.serving import PMMServer, SolveRequest, latency_summary, solve_alone

    c = SERVING[args.workload]
    rng = np.random.Generator(np.random.PCG64(c["seed"]))
    reqs = []
    for i in range(c["requests"]):
        n = int(np.exp(rng.uniform(np.log(c["n_lo"]), np.log(c["n_hi"]))))
        K = int(rng.integers(c["K_lo"], c["K_hi"] + 1))
        nu = float(rng.uniform(*c["nu"]))
        phi = float(rng.uniform(*c["phi"]))
        tree, Q, codes, _ = sim.make_problem(n, K, 5, 50, 1.0, nu, phi, c["seed"] + 1 + i)
        reqs.append((sim.problem_spec(tree, Q, codes), tree.n_nodes))
    log("serving workload: %d requests, leaves %d-%d, sites %d-%d, %d EM iterations each" % (
        len(reqs), c["n_lo"], c["n_hi"], c["K_lo"], c["K_hi"], c["em_iters"]))
    torch.cuda.set_device(0)

    def make(i):
        spec, nn = reqs[i]
        return SolveRequest(spec, [np.full(nn, 0.05)], 0.5, 0.5, max_iter=c["em_iters"], tol=c["tol"], rid=i)

    # warm the code paths, then the plain path's service times
    solve_alone(make(0))
    torch.cuda.synchronize()
    service = []
    for i in range(len(reqs)):
        t0 = time.perf_counter()
        solve_alone(make(i))
        service.append(time.perf_counter() - t0)
    capacity = len(reqs) / sum(service)
    log("plain path: mean service %.1f ms, capacity %.2f req/s" % (1e3 * np.mean(service), capacity))

    def replay(slots, rate, seed, batch_bytes=0):
        arr = np.cumsum(np.random.Generator(np.random.PCG64(seed)).exponential(1.0 / rate, len(reqs)))
        results = [None] * len(reqs)
        with PMMServer(slots=slots, batch_bytes=batch_bytes, batch_max=c["batch_max"]) as srv:
            t0 = time.perf_counter()
            futs = []
            for i, ta in enumerate(arr):
                dt = t0 + ta - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                futs.append(srv.submit(make(i)))
            for i, f in enumerate(futs):
                results[i] = f.result()
        return latency_summary(results), srv.peak_active, srv.batches

    runs = []
    for load in c["loads"]:
        rate = load * capacity
        served, peak, batches = replay(c["slots"], rate, c["seed"] + 7, c["batch_bytes"])
        unbatched, _, _ = replay(c["slots"], rate, c["seed"] + 7)
        plain, _, _ = replay(1, rate, c["seed"] + 7)
        log("offered load %.1fx plain capacity (%.2f req/s): served p50 %.0f ms p99 %.0f ms %.2f req/s "
            "(peak %d concurrent, %d solves for %d requests); unbatched p50 %.0f ms p99 %.0f ms; plain p50 %.0f "
            "ms p99 %.0f ms %.2f req/s" % (
                load, rate, served["p50_ms"], served["p99_ms"], served["throughput_req_per_s"], peak,
                len(batches), sum(batches), unbatched["p50_ms"], unbatched["p99_ms"], plain["p50_ms"],
                plain["p99_ms"], plain["throughput_req_per_s"]))
        runs.append({"offered_load": load, "rate_req_per_s": round(rate, 3),
                     "served": {k: round(v, 3) for k, v in served.items()},
                     "served_unbatched": {k: round(v, 3) for k, v in unbatched.items()},
                     "plain": {k: round(v, 3) for k, v in plain.items()}, "peak_concurrent": peak,
                     "device_solves": len(batches)})
    top = runs[-1]
    res = {
        "metric": "p99 solve latency at %.1fx the plain path's capacity" % top["offered_load"],
        "value": top["served"]["p99_ms"],
        "unit": "ms",
        "n_gpus": 1,
        "steps": c["requests"],
        "warmup": 1,
        "ms_per_step": round(1e3 / top["served"]["throughput_req_per_s"], 3),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Yule trees, PMM-simulated characters, Poisson arrivals)",
        "config": {"workload": "%d requests, leaves log-uniform %d-%d, sites uniform %d-%d, nu in [%g,%g], "
                               "phi in [%g,%g], %d EM iterations (tol %g) from a uniform start; %d slots, "
                               "admission by device footprint, requests of at most %d MiB batched up to %d "
                               "MiB and %d requests" % (
                                   c["requests"], c["n_lo"], c["n_hi"], c["K_lo"], c["K_hi"], c["nu"][0],
                                   c["nu"][1], c["phi"][0], c["phi"][1], c["em_iters"], c["tol"], c["slots"],
                                   c["batch_bytes"] >> 22, c["batch_bytes"] >> 20, c["batch_max"]),
                   "plain_capacity_req_per_s": round(capacity, 3)},
        "runs": runs,
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


QUARTET = {
    # one NNI round of local candidate scoring on an optimised tree
    "nni_local": dict(n=2000, K=200, m_lo=5, m_hi=50, alpha=1.0, nu=0.1, phi=0.15, seed=20261210, sample=16),
}


def bench_quartet(args):
    """Local NNI scoring (nu, phi and all but the five branches around each
    move fixed): every candidate of one round by quartet-local EM in one
    launch, against the per-candidate solver (a device forest and EM per
    candidate, as the reference builds a solver per candidate)."""
    import random

    import torch

    from laml_amd import sim
    from laml_amd.quartet import QuartetScorer
    from laml_amd.solvers import DEFAULT_STRATEGY, EM_solver
    from laml_amd.topology_search import Topology_search_batched, _nni_swap, _nni_unswap

    c = QUARTET[args.workload]
    tree, Q, codes, _ = sim.make_problem(c["n"], c["K"], c["m_lo"], c["m_hi"], c["alpha"], c["nu"], c["phi"],
                                         c["seed"])
    msa = sim.codes_to_charmtrx(tree, codes)
    torch.cuda.set_device(0)
    random.seed(c["seed"])
    t0 = time.perf_counter()
    s = EM_solver([tree.newick()], {"charMtrx": msa}, {"Q": Q}, {"nu": 0, "phi": 0})
    s.score_tree(strategy=dict(DEFAULT_STRATEGY, fixed_brlen=None))
    log("starting tree optimised in %.1f s (score %.4f)" % (time.perf_counter() - t0, -s.negative_llh()))
    ts = Topology_search_batched(s.get_tree_newick(), EM_solver, data={"charMtrx": msa}, prior={"Q": Q},
                                 params=s.get_params())
    moves = ts._list_moves()
    # warm-up round on a few moves, then the timed full round
    sc = QuartetScorer(ts, DEFAULT_STRATEGY, EM_solver)
    sc.score(moves[:8])
    sc.close()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sc = QuartetScorer(ts, DEFAULT_STRATEGY, EM_solver)
    t1 = time.perf_counter()
    res = sc.score(moves)
    torch.cuda.synchronize()
    t2 = time.perf_counter()
    iters = sc.iters.copy()
    sc.close()
    q_time = t2 - t0
    q_rate = len(moves) / q_time
    log("quartet-local EM: %d candidates in %.3f s (set-up %.3f s, EM %.3f s), %.0f candidates/s, "
        "mean %.1f EM iterations" % (len(moves), t2 - t0, t1 - t0, t2 - t1, q_rate, iters.mean()))
    # per-candidate solver on a sample
    sample = moves[:: max(1, len(moves) // c["sample"])][:c["sample"]]
    idx = [moves.index(m) for m in sample]
    worst = 0.0
    times = []
    for i, mv in zip(idx, sample):
        u, v, uc, w = mv
        st = ts._local_strategy(u, DEFAULT_STRATEGY)
        _nni_swap(u, v, uc, w)
        nwks = [t.newick() for t in ts.treeList_obj]
        _nni_unswap(u, v, uc, w)
        t0 = time.perf_counter()
        sv = EM_solver(nwks, {"charMtrx": msa}, {"Q": Q}, ts.params)
        nllh, _ = sv.optimize(initials=1, verbose=-1, fixed_phi=st["fixed_phi"], fixed_nu=st["fixed_nu"],
                              fixed_brlen=st["fixed_brlen"], random_seeds=[7])
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t0)
        worst = max(worst, abs(-res[i][0] - nllh) / abs(nllh))
    full_rate = 1.0 / float(np.mean(times))
    log("per-candidate solver: %.1f ms per candidate (%d sampled), %.1f candidates/s; "
        "max relative NLL difference to quartet EM %.2e" % (1e3 * np.mean(times), len(times), full_rate, worst))
    if worst > 1e-9:
        raise SystemExit("quartet-local EM differs from the per-candidate solver")
    res_line = {
        "metric": "NNI candidates scored/s (local branch optimisation)",
        "value": round(q_rate, 1),
        "unit": "candidates/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 1,
        "ms_per_step": round(q_time * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Yule tree, PMM-simulated characters)",
        "config": {"workload": "one NNI round (%d candidates) on an optimised %d-leaf Yule tree x %d sites, "
                               "M in [%d,%d], nu=%g, phi=%g; nu, phi and all but 5 branches fixed per candidate"
                               % (len(moves), c["n"], c["K"], c["m_lo"], c["m_hi"], c["nu"], c["phi"]),
                   "mean_em_iterations": round(float(iters.mean()), 2)},
        "per_candidate_solver": {"candidates_per_s": round(full_rate, 2), "sampled": len(times),
                                 "speedup_of_quartet": round(q_rate / full_rate, 1),
                                 "max_rel_nll_diff": worst},
        "cpu_baseline": None,
    }
    print(json.dumps(res_line), flush=True)


def bench_scoring(args):
    """Batched scoring workload: one forest holding every candidate and
    independent tree, EM from a uniform start, against the plain path of one
    forest (one launch sequence) per tree."""
    import torch

    from laml_amd import sim
    from laml_amd.autotune import autotune
    from laml_amd.forest import Forest

    c = SCORING[args.workload]
    problems = sim.scoring_workload(**c)
    sizes = np.asarray([(t.n_nodes + 1) // 2 for (t, _, _) in problems[c["variants"]:]])
    P = len(problems)
    log("scoring workload %s: %d trees (%d NNI variants of a %d-leaf tree + %d trees of %d-%d leaves), K=%d" % (
        args.workload, P, c["variants"], c["base_n"], c["others"], sizes.min(), sizes.max(), c["K"]))
    torch.cuda.set_device(0)

    def start(F, probs):
        d = torch.tensor(F.lengths_to_level([np.full(t.n_nodes, 0.05) for (t, _, _) in probs]),
                         dtype=torch.float64, device=F.device)
        nu = torch.full((len(probs),), 0.5, dtype=torch.float64, device=F.device)
        phi = torch.full((len(probs),), 0.5, dtype=torch.float64, device=F.device)
        return d, nu, phi

    specs = [sim.problem_spec(t, q, cd) for (t, q, cd) in problems]
    t0 = time.perf_counter()
    F = Forest(specs, device=0)
    torch.cuda.synchronize()
    setup = time.perf_counter() - t0
    d, nu, phi = start(F, problems)
    plan = None
    if not args.no_autotune:
        plan, _ = autotune(F, d, nu, phi)
    F.em(d, nu, phi, args.warmup, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nll, _ = F.em(d, nu, phi, args.steps, 0.0)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    batch_rate = P * args.steps / el
    nll_b = nll.cpu().numpy().copy()
    log("batched: %.3f ms per EM iteration of all %d trees, %.1f tree-iterations/s (setup %.1f s)" % (
        el / args.steps * 1e3, P, batch_rate, setup))
    F.close()
    # plain path: one forest per tree, the same EM schedule, a sample of trees
    sample = list(range(0, c["variants"], max(1, c["variants"] // 16))) + \
        list(range(c["variants"], P, max(1, c["others"] // 16)))
    per_tree = []
    for i in sample:
        Fi = Forest([specs[i]], device=0)
        di, nui, phii = start(Fi, [problems[i]])
        Fi.em(di, nui, phii, args.warmup, 0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nli, _ = Fi.em(di, nui, phii, args.steps, 0.0)
        torch.cuda.synchronize()
        per_tree.append((time.perf_counter() - t0) / args.steps)
        got = float(nli[0].item())
        if abs(got - nll_b[i]) > 1e-9 * abs(got):
            raise SystemExit("tree %d: batched result %r differs from alone %r" % (i, nll_b[i], got))
        Fi.close()
    # sampled mean time per tree-iteration, weighted to the workload's mix
    nv = sum(1 for i in sample if i < c["variants"])
    mean_v = float(np.mean(per_tree[:nv]))
    mean_o = float(np.mean(per_tree[nv:]))
    plain_s = c["variants"] * mean_v + c["others"] * mean_o
    plain_rate = P / plain_s
    log("one forest per tree: %.3f ms per variant iteration, %.3f ms per independent-tree iteration "
        "(%d sampled); %.1f tree-iterations/s" % (mean_v * 1e3, mean_o * 1e3, len(sample), plain_rate))
    res = {
        "metric": "tree EM iterations/s, batched candidate scoring",
        "value": round(batch_rate, 2),
        "unit": "tree EM iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Yule trees, NNI variants, PMM-simulated characters)",
        "config": {"workload": "%d NNI variants of a %d-leaf Yule tree + %d Yule trees with log-uniform %d-%d "
                               "leaves, %d sites each, M in [%d,%d], nu=%g, phi=%g, one forest, EM from a "
                               "uniform start" % (c["variants"], c["base_n"], c["others"], c["n_lo"], c["n_hi"],
                                                  c["K"], c["m_lo"], c["m_hi"], c["nu"], c["phi"]),
                   "trees": P, "nodes": int(sum(t.n_nodes for (t, _, _) in problems)),
                   "launch_plan": plan.describe().get("mode") if plan else "default"},
        "one_forest_per_tree": {"tree_iters_per_s": round(plain_rate, 2), "sampled_trees": len(sample),
                                "speedup_of_batch": round(batch_rate / plain_rate, 2)},
        "cpu_baseline": None,
    }
    print(json.dumps(res), flush=True)


def serving_leg(requests=24, load=2.0, slots=4, cpu_budget_s=6.0):
    """The latency half of the metric, bounded for the default run: a
    Poisson stream of `requests` likelihood + EM solves in the serving
    workload's ranges (leaves log-uniform 200-50,000, sites 20-500, nu in
    [0.02,0.2], phi in [0.05,0.3], 50 EM iterations, tol 1e-8) offered at
    `load` x the plain path's capacity, served by the continuous-batching
    server (`slots` concurrent solves, admission by device footprint, small
    requests sharing a forest), by the same server without batching and by
    the plain per-request path (one solve at a time) on the same arrivals."""
    import torch

    from laml_amd import sim
    from laml_amd.serving import PMMServer, SolveRequest, latency_summary, solve_alone

    c = SERVING["serve"]
    rng = np.random.Generator(np.random.PCG64(c["seed"]))
    probs = []
    for i in range(requests):
        n = int(np.exp(rng.uniform(np.log(c["n_lo"]), np.log(c["n_hi"]))))
        K = int(rng.integers(c["K_lo"], c["K_hi"] + 1))
        nu = float(rng.uniform(*c["nu"]))
        phi = float(rng.uniform(*c["phi"]))
        probs.append(sim.make_problem(n, K, 5, 50, 1.0, nu, phi, c["seed"] + 1 + i))
    specs = [sim.problem_spec(t, q, cd) for (t, q, cd, _) in probs]

    def make(i):
        return SolveRequest(specs[i], [np.full(probs[i][0].n_nodes, 0.05)], 0.5, 0.5, max_iter=c["em_iters"],
                            tol=c["tol"], rid=i)

    torch.cuda.set_device(0)
    solve_alone(make(0))
    torch.cuda.synchronize()
    service = []
    for i in range(requests):
        t0 = time.perf_counter()
        solve_alone(make(i))
        service.append(time.perf_counter() - t0)
    capacity = requests / sum(service)
    rate = load * capacity
    arrivals = np.cumsum(np.random.Generator(np.random.PCG64(c["seed"] + 7)).exponential(1.0 / rate, requests))

    def replay(n_slots, batch_bytes=0):
        with PMMServer(slots=n_slots, batch_bytes=batch_bytes, batch_max=c["batch_max"]) as srv:
            t0 = time.perf_counter()
            futs = []
            for i, ta in enumerate(arrivals):
                dt = t0 + ta - time.perf_counter()
                if dt > 0:
                    time.sleep(dt)
                futs.append(srv.submit(make(i)))
            res = [f.result() for f in futs]
        return latency_summary(res), srv.peak_active, srv.batches

    served, peak, batches = replay(slots, c["batch_bytes"])
    unbatched, _, _ = replay(slots)
    plain, _, _ = replay(1)
    log("serving leg: %d requests at %.1fx plain capacity (%.2f req/s): p50 %.0f ms p99 %.0f ms %.2f req/s "
        "(peak %d concurrent, %d device solves); unbatched p50 %.0f ms p99 %.0f ms %.2f req/s; plain path p50 "
        "%.0f ms p99 %.0f ms %.2f req/s" % (
            requests, load, rate, served["p50_ms"], served["p99_ms"], served["throughput_req_per_s"], peak,
            len(batches), unbatched["p50_ms"], unbatched["p99_ms"], unbatched["throughput_req_per_s"],
            plain["p50_ms"], plain["p99_ms"], plain["throughput_req_per_s"]))
    # CPU baseline: the C/OpenMP oracle solving the same requests in arrival
    # order (same start, same EM schedule) until its time budget is spent
    from oracle import pmm_ref

    done, t_cpu = 0, 0.0
    for i in sorted(range(requests), key=lambda j: probs[j][0].n_nodes * len(probs[j][1])):
        t, q, cd, _ = probs[i]
        rp = pmm_ref.RefProblem(t.c0, t.c1, t.parent, cd, q, None, 1.0 / (len(q) * cd.shape[0]))
        t0 = time.perf_counter()
        rp.em(np.full(t.n_nodes, 0.05), 0.5, 0.5, max_iter=c["em_iters"], tol=c["tol"])
        t_cpu += time.perf_counter() - t0
        done += 1
        if t_cpu > cpu_budget_s:
            break
    threads = pmm_ref.threads()
    return {
        "requests": requests, "offered_load": load, "rate_req_per_s": round(rate, 3),
        "p50_ms": round(served["p50_ms"], 2), "p99_ms": round(served["p99_ms"], 2),
        "throughput_req_per_s": round(served["throughput_req_per_s"], 3), "slots": slots,
        "peak_concurrent": peak, "device_solves": len(batches),
        "batching": "requests of at most %d MiB share a forest, up to %d MiB and %d requests" % (
            c["batch_bytes"] >> 22, c["batch_bytes"] >> 20, c["batch_max"]),
        "unbatched": {"p50_ms": round(unbatched["p50_ms"], 2), "p99_ms": round(unbatched["p99_ms"], 2),
                      "throughput_req_per_s": round(unbatched["throughput_req_per_s"], 3)},
        "plain_path": {"p50_ms": round(plain["p50_ms"], 2), "p99_ms": round(plain["p99_ms"], 2),
                       "throughput_req_per_s": round(plain["throughput_req_per_s"], 3),
                       "capacity_req_per_s": round(capacity, 3)},
        "workload": "leaves log-uniform %d-%d, sites uniform %d-%d, nu in [%g,%g], phi in [%g,%g], %d EM "
                    "iterations (tol %g) from a uniform start, seed %d" % (
                        c["n_lo"], c["n_hi"], c["K_lo"], c["K_hi"], c["nu"][0], c["nu"][1], c["phi"][0],
                        c["phi"][1], c["em_iters"], c["tol"], c["seed"]),
        "cpu_baseline": {"value": round(done / t_cpu, 3), "unit": "req/s", "cores": threads, "kind": "port",
                         "sample": "C/OpenMP restatement of EM_solver (oracle/pmm_ref.c), %d threads, the %d "
                                   "smallest of the %d requests solved one after another (%.2f s)"
                                   % (threads, done, requests, t_cpu)},
    }


def scoring_leg(steps=10, warmup=2, sample=16):
    """Batched scoring, bounded for the default run: the 768-tree workload
    (512 NNI variants of a 2000-leaf tree + 256 trees of 100-5000 leaves,
    200 sites) as one forest with its default launch plan, tree EM
    iterations/s against one forest per tree on a sample."""
    import torch

    from laml_amd import sim
    from laml_amd.forest import Forest, SiteTable

    c = SCORING["nni768"]
    problems = sim.scoring_workload(**c)
    tables = {}
    specs = []
    for (t, q, cd) in problems:
        if id(q) not in tables:
            tables[id(q)] = SiteTable(q)
        specs.append(sim.problem_spec(t, q, cd, sites=tables[id(q)]))
    P = len(problems)
    torch.cuda.set_device(0)

    def start(F, probs):
        d = torch.tensor(F.lengths_to_level([np.full(t.n_nodes, 0.05) for (t, _, _) in probs]),
                         dtype=torch.float64, device=F.device)
        nu = torch.full((len(probs),), 0.5, dtype=torch.float64, device=F.device)
        phi = torch.full((len(probs),), 0.5, dtype=torch.float64, device=F.device)
        return d, nu, phi

    F = Forest(specs, device=0)
    d, nu, phi = start(F, problems)
    F.em(d, nu, phi, warmup, 0.0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    nll, _ = F.em(d, nu, phi, steps, 0.0)
    torch.cuda.synchronize()
    el = time.perf_counter() - t0
    nll_b = nll.cpu().numpy().copy()
    F.close()
    batch_rate = P * steps / el
    nv = c["variants"]
    idx = list(range(0, nv, nv // (sample // 2)))[:sample // 2] + \
        list(range(nv, P, (P - nv) // (sample // 2)))[:sample // 2]
    per_tree, worst = [], 0.0
    for i in idx:
        Fi = Forest([specs[i]], device=0)
        di, nui, phii = start(Fi, [problems[i]])
        Fi.em(di, nui, phii, warmup, 0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nli, _ = Fi.em(di, nui, phii, steps, 0.0)
        torch.cuda.synchronize()
        per_tree.append((time.perf_counter() - t0) / steps)
        got = float(nli[0].item())
        worst = max(worst, abs(got - nll_b[i]) / abs(got))
        Fi.close()
    if worst > 1e-9:
        raise SystemExit("batched scoring differs from one forest per tree (relative %g)" % worst)
    k = len(idx) // 2
    plain_s = nv * float(np.mean(per_tree[:k])) + (P - nv) * float(np.mean(per_tree[k:]))
    plain_rate = P / plain_s
    log("scoring leg: %d trees, %.3f ms per EM iteration of the batch, %.0f tree-iterations/s; one forest "
        "per tree %.0f tree-iterations/s (%d sampled); x%.1f" % (P, el / steps * 1e3, batch_rate, plain_rate,
                                                               len(idx), batch_rate / plain_rate))
    # CPU baseline: the C/OpenMP oracle on sampled trees, the same EM steps
    from oracle import pmm_ref

    t_cpu, it_cpu = 0.0, 0
    for i in idx[:4] + idx[k:k + 4]:
        t, q, cd = problems[i]
        rp = pmm_ref.RefProblem(t.c0, t.c1, t.parent, cd, q, None, 1.0 / (len(q) * cd.shape[0]))
        t0 = time.perf_counter()
        rp.em(np.full(t.n_nodes, 0.05), 0.5, 0.5, max_iter=3, tol=0.0, fixed_iters=3)
        t_cpu += time.perf_counter() - t0
        it_cpu += 3
    threads = pmm_ref.threads()
    return {
        "trees": P, "steps": steps, "ms_per_step": round(el / steps * 1e3, 4),
        "value": round(batch_rate, 2), "unit": "tree EM iters/s",
        "one_forest_per_tree": {"tree_iters_per_s": round(plain_rate, 2), "sampled_trees": len(idx),
                                "speedup_of_batch": round(batch_rate / plain_rate, 2),
                                "max_rel_nll_diff": worst},
        "launch_plan": "default",
        "workload": "%d NNI variants of a %d-leaf Yule tree + %d Yule trees with log-uniform %d-%d leaves, %d "
                    "sites, M in [%d,%d], nu=%g, phi=%g, EM from a uniform start" % (
                        c["variants"], c["base_n"], c["others"], c["n_lo"], c["n_hi"], c["K"], c["m_lo"], c["m_hi"],
                        c["nu"], c["phi"]),
        "cpu_baseline": {"value": round(it_cpu / t_cpu, 3), "unit": "tree EM iters/s", "cores": threads,
                         "kind": "port",
                         "sample": "C/OpenMP restatement of EM_solver (oracle/pmm_ref.c), %d threads, 3 EM "
                                   "iterations on each of 8 sampled trees (4 NNI variants, 4 independent "
                                   "trees) in %.2f s" % (threads, t_cpu)},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="em10k", choices=sorted(WORKLOADS) + sorted(SCORING) + sorted(SERVING) + sorted(QUARTET))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--roofline-reps", type=int, default=10)
    ap.add_argument("--batch", type=int, default=16, help="independent problems per step")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default launch plan of the batch forest (no autotuning before timing)")
    ap.add_argument("--streams", type=int, default=2,
                    help="split the batch into this many forests, each stepped on its own HIP stream "
                         "(laml_amd.streams.PipelinedBatch); 1: one forest")
    ap.add_argument("--save-plan", default=None,
                    help="write the batch forest's autotuned launch plan to this JSON file")
    ap.add_argument("--plan", default=None,
                    help="apply the launch plan in this JSON file (written by --save-plan) instead of autotuning; "
                         "'profile': the plan the committed profile summary of the workload was measured with")
    ap.add_argument("--no-extra-legs", action="store_true",
                    help="skip the serving-latency and batched-scoring legs of the default run")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the measurement, run 3 plain device copies of this many MiB (a known-size "
                         "reference for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the PMM path on a single MI355X "
              "(multi-GPU sharding by tree or by site is designed, not built); --gpus %d is not supported"
              % args.gpus, file=sys.stderr)
        sys.exit(2)

    if args.workload in SCORING:
        return bench_scoring(args)
    if args.workload in SERVING:
        return bench_serving(args)
    if args.workload in QUARTET:
        return bench_quartet(args)

    import torch

    from laml_amd import sim
    from laml_amd.forest import Forest

    n, K, mlo, mhi, alpha, nu_true, phi_true, seed = WORKLOADS[args.workload]
    B = max(1, args.batch)
    log("generating workload %s: %d problems of n=%d K=%d" % (args.workload, B, n, K))
    probs = [sim.make_problem(n, K, mlo, mhi, alpha, nu_true, phi_true, seed + 1000 * b) for b in range(B)]
    tree, Q, codes, truth = probs[0]
    torch.cuda.set_device(0)

    plan_desc = {}

    def run(problems, steps, warmup, tune=False):
        specs = [sim.problem_spec(t, q, c) for (t, q, c, _) in problems]
        t0 = time.perf_counter()
        F = Forest(specs, device=0)
        torch.cuda.synchronize()
        setup = time.perf_counter() - t0
        dev = F.device
        d0 = F.lengths_to_level([np.full(t.n_nodes, 0.05) for (t, _, _, _) in problems])
        d = torch.tensor(d0, dtype=torch.float64, device=dev)
        nu = torch.full((len(problems),), 0.5, dtype=torch.float64, device=dev)
        phi = torch.full((len(problems),), 0.5, dtype=torch.float64, device=dev)
        if tune and args.plan:
            from laml_amd.autotune import plan_from_dict
            plan = plan_from_dict(profile_plan() if args.plan == "profile" else json.load(open(args.plan)), F)
            plan.apply(F)
            plan_desc.update(plan.describe())
            log("launch plan from %s: %s" % (args.plan, plan.describe().get("mode")))
        elif tune:
            from laml_amd.autotune import autotune
            t1 = time.perf_counter()
            plan, tlog = autotune(F, d, nu, phi)
            plan_desc.update(plan.describe())
            log("autotuned launch plan in %.2f s: %s (E-step %.3f ms -> %.3f ms)" % (
                time.perf_counter() - t1, plan.describe().get("mode"), tlog[0][1], min(ms for _, ms in tlog)))
            if args.save_plan:
                with open(args.save_plan, "w") as fh:
                    json.dump(plan.describe(), fh)
        if warmup > 0:
            F.em(d, nu, phi, warmup, 0.0)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        nll, _ = F.em(d, nu, phi, steps, 0.0)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        return F, d, nu, phi, nll, el, setup

    # single problem alone (the plain per-tree GPU path)
    F1, d1, nu1, phi1, nll1, el1, _ = run(probs[:1], args.steps, args.warmup)
    single_per_s = args.steps / el1
    single_nll = float(nll1[0].item())
    log("single problem: %.3f ms/iter, %.1f EM it/s, nllh %.6f" % (el1 / args.steps * 1e3, single_per_s, single_nll))
    F1.close()
    del F1, d1, nu1, phi1
    # the batch: one forest (also the roofline leg's forest), then, with
    # --streams S > 1, the measured run on S forests stepped on S streams
    roof_plan = profile_plan() if (B == 16 and args.workload == "em10k" and not args.plan) else None
    F, d, nu, phi, nll, el, setup_s = run(probs, args.steps if args.streams <= 1 else 0,
                                          args.warmup if args.streams <= 1 else 0,
                                          tune=not args.no_autotune and (args.streams <= 1 or roof_plan is None))
    execution = {"forests": 1, "streams": 1}
    if args.streams > 1:
        from laml_amd.streams import PipelinedBatch
        t0 = time.perf_counter()
        pb = PipelinedBatch([sim.problem_spec(t, q, c) for (t, q, c, _) in probs], parts=args.streams, device=0)
        prm = pb.params([[np.full(t.n_nodes, 0.05)] for (t, _, _, _) in probs], [0.5] * B, [0.5] * B)
        setup_s = time.perf_counter() - t0
        plans = pb.autotune(prm) if not args.no_autotune else None
        # warm-up steps, then the timed steps from where they end (as for
        # the single problem)
        if args.warmup > 0:
            pb.em(prm, args.warmup, 0.0)
        pb.synchronize()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = pb.em(prm, args.steps, 0.0)
        pb.synchronize()
        el = time.perf_counter() - t0
        j, i0 = pb.locate(0)
        nll = res[j][0][i0:i0 + 1]
        execution = {"forests": len(pb.forests), "streams": len(pb.streams),
                     "problems_per_forest": [len(g) for g in pb.groups],
                     "plans": [{k: v for k, v in p.describe().items() if "cpw" not in k} for p in plans] if plans
                     else "default"}
        pb.close()
    em_per_s = B * args.steps / el
    final_nll = float(nll[0].item())
    log("batch of %d: %.3f ms/iter, %.1f aggregate EM it/s (x%.1f the single-problem rate); setup %.2f s; %s" % (
        B, el / args.steps * 1e3, em_per_s, em_per_s / single_per_s, setup_s,
        "one forest" if args.streams <= 1 else "%d forests on %d streams" % (execution["forests"], execution["streams"])))
    if abs(final_nll - single_nll) > 1e-9 * abs(single_nll):
        raise SystemExit("batched result differs from the single-problem result: %r vs %r" % (final_nll, single_nll))

    # the roofline leg runs the plan of the committed profile of this workload
    # when there is one (its traffic and issue counters then describe these
    # launches)
    if roof_plan is not None:
        from laml_amd.autotune import plan_from_dict
        plan_from_dict(roof_plan, F).apply(F)
        log("roofline leg: launch plan of %s" % PROFILE)

    # roofline of the dominant kernel family: the outside pass (all level
    # launches of k_outside, the fused top-level kernel, k_root, k_counts,
    # k_rtilde), timed with events on the launch stream over resident inside
    # state
    # chain regions: the levels below each pass's region top and below the split
    split = min(max(1, int(F.get_shapes()[2])), F.num_levels)
    flags = F.chains()
    _, top_in, top_out = F.get_chains()
    region = lambda t: split if t <= 0 else min(t, split)
    inside_b, outside_b = algorithmic_bytes(F, fused=F.fusion(), chain_split=region(top_out) if flags & 1 else None,
                                            in_chain_split=region(top_in) if flags & 2 else None)
    F.nll(d, nu, phi)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    # untimed passes first: the first ten passes of a fresh forest run about
    # 4 % slower than the steady state
    for _ in range(max(10, args.roofline_reps)):
        F.outside(phi)
    e0.record(stream)
    for _ in range(args.roofline_reps):
        F.outside(phi)
    e1.record(stream)
    torch.cuda.synchronize()
    out_ms = e0.elapsed_time(e1) / args.roofline_reps
    for _ in range(3):
        F.nll(d, nu, phi)
    e0.record(stream)
    for _ in range(args.roofline_reps):
        F.nll(d, nu, phi)
    e1.record(stream)
    torch.cuda.synchronize()
    in_ms = e0.elapsed_time(e1) / args.roofline_reps
    achieved = outside_b / (out_ms * 1e-3) / 1e9
    log("outside pass %.3f ms (%.1f GB/s), inside pass %.3f ms (%.1f GB/s)" % (
        out_ms, achieved, in_ms, inside_b / (in_ms * 1e-3) / 1e9))

    node_site = 2 * F.num_nodes * K  # over the whole batch
    traffic, valu = outside_traffic_from_profile(args) if B == 16 and args.workload == "em10k" else (None, None)
    res = {
        "metric": "EM iters/s at 10k cells x 300 sites",
        "value": round(em_per_s, 3),
        "unit": "EM iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "single_problem": {"em_iters_per_s": round(single_per_s, 3), "ms_per_iter": round(el1 / args.steps * 1e3, 4),
                           "batch_speedup": round(em_per_s / single_per_s, 2)},
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Yule tree + PMM-simulated characters)",
        "config": {"workload": "fixed-topology PMM EM, batch of %d independent problems, each a %d-leaf Yule "
                               "tree x %d sites, M in [%d,%d], Dirichlet(%g), nu=%g, phi=%g, seeds %d+1000*b"
                               % (B, n, K, mlo, mhi, alpha, nu_true, phi_true, seed),
                   "batch": B, "leaves": n, "sites": K, "nodes": F.num_nodes, "levels": F.num_levels,
                   "launch_plan": (plan_desc or "default") if args.streams <= 1 else "per forest, see execution",
                   "execution": execution},
        "node_site_evals_per_s": round(node_site * em_per_s, 1),
        "final_nllh_problem0": final_nll,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": traffic,
                     "traffic_source": PROFILE if traffic is not None else None,
                     "valu_issue_frac": valu,
                     "traffic_gbs": round(traffic / (out_ms * 1e-3) / 1e9, 1) if traffic else None,
                     "note": "achieved counts the compulsory bytes of the algorithm the plan runs (outside chains "
                             "keep heavy children's outside states in registers, about half the bytes of level "
                             "launches); the chain steps read each child's prior weight from a row built at "
                             "forest creation (8 B per child and site, streamed) instead of gathering prior-table "
                             "entries, a line per lane: those bytes are in traffic, not in achieved; the chain "
                             "steps are bound by instruction issue (FP64 VALU) and memory latency together, not by "
                             "HBM bandwidth (traffic_gbs)",
                     "launch_plan": "as in " + PROFILE if roof_plan is not None else "as the batch",
                     "kernel": "outside pass (laml_outside: k_tree_walk top levels + root in the hybrid plan, "
                               "k_outside_chain rounds below the split when the plan has chains, else "
                               "k_outside_rows levels, k_counts, k_rtilde)",
                     "algorithmic_bytes": outside_b, "ms": round(out_ms, 4),
                     "inside_pass": {"algorithmic_bytes": inside_b, "ms": round(in_ms, 4)}},
    }
    if not args.no_cpu_baseline:
        log("timing the CPU oracle (C/OpenMP restatement) on problem 0")
        cb = cpu_baseline(tree, Q, codes)
        res["cpu_baseline"] = cb
    else:
        res["cpu_baseline"] = None
    F.close()
    if args.workload == "em10k" and not args.no_extra_legs:
        res["serving"] = serving_leg()
        res["batched_scoring"] = scoring_leg()
    if args.calib_copy_mib > 0:
        src = torch.ones(args.calib_copy_mib * (1 << 20) // 8, dtype=torch.float64, device=0)
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        log("calibration copies: 3 x %d MiB" % args.calib_copy_mib)
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
