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


def algorithmic_bytes(F):
    """Compulsory HBM bytes of one inside pass and one outside pass.

    Per (internal node, site): inside reads the node's label (2 B) and each
    child's state (leaf: int16 code, 2 B; internal: label + beta pair +
    exponent, 22 B) and writes beta + exponent (20 B).  Outside reads the
    node's outside state (alpha quad + exponent + specific state, 38 B), each
    child's inside state again (2 / 22 B) and writes each internal child's
    outside state (38 B)."""
    N, n = F.num_nodes, F.num_leaves
    K = F.tree_K[F.tree_of].astype(np.int64)
    inter = np.arange(n, N)
    c = np.stack([F.child0[inter], F.child1[inter]])
    rc = np.where(c < n, 2, 22)
    wo = np.where(c < n, 0, 38)
    Kv = K[inter]
    inside = int((Kv * (2 + 20 + rc.sum(0))).sum())
    outside = int((Kv * (38 + rc.sum(0) + wo.sum(0))).sum())
    return inside, outside


def cpu_baseline(tree, Q, codes, nu, phi, target_s=12.0):
    """Time the CPU oracle (scalar Python restatement of EM_solver) on a
    bounded sample: the same tree with the first k sites, one EM iteration
    (E-step incl. inside pass, M-step, NLL).  The E-step part is scaled to all
    sites; the M-step (independent of K) is added once."""
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
    return {
        "value": 1.0 / per_iter,
        "unit": "EM iters/s",
        "cores": 1,
        "kind": "port",
        "sample": "oracle (scalar FP64 Python restatement of EM_solver) on the same tree with the first %d of %d "
                  "sites, one EM iteration: E-step+NLL %.2f s scaled x%.1f to all sites, M-step %.2f s added once"
                  % (ks, K, e_time, K / ks, m_time),
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="em10k", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--roofline-reps", type=int, default=10)
    ap.add_argument("--batch", type=int, default=16, help="independent problems per step")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="after the measurement, run 3 plain device copies of this many MiB (a known-size "
                         "reference for hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs the PMM path on a single MI355X "
              "(see DESIGN.md for how it would shard); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    from laml_amd import sim
    from laml_amd.forest import Forest

    n, K, mlo, mhi, alpha, nu_true, phi_true, seed = WORKLOADS[args.workload]
    B = max(1, args.batch)
    log("generating workload %s: %d problems of n=%d K=%d" % (args.workload, B, n, K))
    probs = [sim.make_problem(n, K, mlo, mhi, alpha, nu_true, phi_true, seed + 1000 * b) for b in range(B)]
    tree, Q, codes, truth = probs[0]
    torch.cuda.set_device(0)

    def run(problems, steps, warmup):
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
    # the batch
    F, d, nu, phi, nll, el, setup_s = run(probs, args.steps, args.warmup)
    em_per_s = B * args.steps / el
    final_nll = float(nll[0].item())
    log("batch of %d: %.3f ms/iter, %.1f aggregate EM it/s (x%.1f the single-problem rate); setup %.2f s" % (
        B, el / args.steps * 1e3, em_per_s, em_per_s / single_per_s, setup_s))
    if abs(final_nll - single_nll) > 1e-9 * abs(single_nll):
        raise SystemExit("batched result differs from the single-problem result: %r vs %r" % (final_nll, single_nll))

    # roofline of the dominant kernel family: the outside pass (all level
    # launches of k_outside, the fused top-level kernel, k_root, k_counts,
    # k_rtilde), timed with events on the launch stream over resident inside
    # state
    inside_b, outside_b = algorithmic_bytes(F)
    F.nll(d, nu, phi)
    stream = torch.cuda.current_stream()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    F.outside(phi)
    e0.record(stream)
    for _ in range(args.roofline_reps):
        F.outside(phi)
    e1.record(stream)
    torch.cuda.synchronize()
    out_ms = e0.elapsed_time(e1) / args.roofline_reps
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
                   "batch": B, "leaves": n, "sites": K, "nodes": F.num_nodes, "levels": F.num_levels},
        "node_site_evals_per_s": round(node_site * em_per_s, 1),
        "final_nllh_problem0": final_nll,
        "roofline": {"bound": "hbm", "achieved": round(achieved, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 5), "traffic": None,
                     "kernel": "outside pass (k_outside levels + k_root + k_rtilde)",
                     "algorithmic_bytes": outside_b, "ms": round(out_ms, 4),
                     "inside_pass": {"algorithmic_bytes": inside_b, "ms": round(in_ms, 4)}},
    }
    if not args.no_cpu_baseline:
        log("timing the CPU oracle on a bounded sample")
        cb = cpu_baseline(tree, Q, codes, nu_true, phi_true)
        res["cpu_baseline"] = cb
    else:
        res["cpu_baseline"] = None
    F.close()
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
