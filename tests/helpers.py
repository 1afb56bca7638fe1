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


def load_topology_kats():
    """Fixture of the reference's topology-search tests: (Q, msa dict by name,
    brute-force optima, cases)."""
    with open(os.path.join(GOLDEN, "topology_search_kats.json")) as f:
        k = json.load(f)
    Q = [{int(a): b for a, b in q.items()} for q in k["Q"]]
    return Q, k["msa"], k["brute_force"], k["cases"]


def rooted_topologies(labels):
    """Every rooted binary topology on the labels, as Newick strings, by
    inserting each further leaf on every edge (root edge included):
    (2n-3)!! trees."""
    def insert(leaf, t):
        out = [(t, leaf)]
        if isinstance(t, tuple):
            left, right = t
            out += [(x, right) for x in insert(leaf, left)] + [(left, x) for x in insert(leaf, right)]
        return out

    def nwk(t):
        return t if isinstance(t, str) else "(" + nwk(t[0]) + "," + nwk(t[1]) + ")"

    topos = [(labels[0], labels[1])]
    for leaf in labels[2:]:
        topos = [x for t in topos for x in insert(leaf, t)]
    return [nwk(t) + ";" for t in topos]


class OracleSolver:
    """The oracle behind the solver surface that Topology_search drives
    (constructor, score_tree, get_tree_newick, get_params), so that the
    search's host logic can run on the CPU.  score_tree follows
    ML_solver.optimize(initials=1, random_seeds=None): one seed is drawn from
    the global generator."""
    solver_name = "EM_solver"

    def __init__(self, treeList, data, prior, params={'nu': 0, 'phi': 0}):
        self.treeList = list(treeList)
        self.charMtrx = data['charMtrx']
        self.Q = prior['Q']
        self.params = dict(params)
        self.trees = list(treeList)

    def score_tree(self, strategy):
        import random
        from laml_amd.tree import read_tree_newick
        from oracle.em_oracle import OracleEM, to_newick

        seed = int(random.random() * 10000)
        o = OracleEM(self.treeList, self.charMtrx, self.Q)
        res = o.optimize(initials=1, random_seeds=seed, fixed_phi=strategy['fixed_phi'],
                         fixed_nu=strategy['fixed_nu'], ultra_constr=strategy['ultra_constr'],
                         fixed_brlen=strategy.get('fixed_brlen'))
        nllh, _, nu, phi, _, status = res
        out = []
        for r in o.trees:
            t = read_tree_newick(to_newick(r))
            t.collapse_short_branches(o.dmin * 0.01)
            out.append(t.newick())
        self.trees = out
        self.params = {'nu': nu, 'phi': phi}
        return -nllh, status

    def get_tree_newick(self):
        return list(self.trees)

    def get_params(self):
        return dict(self.params)
