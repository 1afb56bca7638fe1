"""Shared fixtures/inputs for the test-suite."""
import json
import os

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load_kats():
    with open(os.path.join(GOLDEN, "reference_kats.json")) as f:
        return json.load(f)


def _conv(x):
    return x if x == "?" else int(x)


def kat_inputs(case):
    Q = [{int(k): v for k, v in q.items()} for q in case["Q"]]
    msa = {c: [_conv(x) for x in row] for c, row in case["charMtrx"].items()}
    return case["tree"], msa, Q, case["nu"], case["phi"]


def golden_path(*parts):
    return os.path.join(GOLDEN, *parts)


def load_test4():
    from oracle.em_oracle import read_charmtrx, read_priors_csv

    T = open(golden_path("test4.tre")).read()
    msa = read_charmtrx(golden_path("test4_charMtrx.txt"), masked_symbol="?")
    k = len(next(iter(msa.values())))
    Q = read_priors_csv(golden_path("test4_prior.csv"), k)
    return T, msa, Q


def load_test1():
    """test_data/test_EM/test1 with the uniform observed-state priors of
    unit_tests_EMSolver.py:505-520."""
    from oracle.em_oracle import read_charmtrx

    T = open(golden_path("test1.tre")).read()
    msa = read_charmtrx(golden_path("test1_charMtrx.txt"), masked_symbol="-")
    Q = []
    for i in range(60):
        M = set(msa[x][i] for x in msa if msa[x][i] not in [0, "?"])
        q = {x: 1 / len(M) for x in M}
        q[0] = 0
        Q.append(q)
    return T, msa, Q


def load_example1():
    """examples/example1 (250 cells x 30 sites, the shape of the 250x30 config).
    The starting tree carries no branch lengths; seeded lengths in
    [0.01, 0.3] are attached so the likelihood is defined."""
    import random

    from laml_amd.tree import read_tree_newick
    from oracle.em_oracle import read_charmtrx

    t = read_tree_newick(open(golden_path("example1", "starting.tree")).read())
    rng = random.Random(250)
    for v in t.traverse_postorder():
        v.edge_length = round(rng.uniform(0.01, 0.3), 6)
    T = t.newick()
    msa = read_charmtrx(golden_path("example1", "character_matrix.csv"))
    k = len(next(iter(msa.values())))
    Q = [{} for _ in range(k)]
    with open(golden_path("example1", "priors.csv")) as f:
        for line in f:
            line = line.strip()
            if not line:
                continue
            a, b, c = line.split(",")
            Q[int(a)][int(b)] = float(c)
    return T, msa, Q
