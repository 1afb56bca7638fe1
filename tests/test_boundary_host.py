"""Solvers cross a fork on the host (Topology_search_parallel.py:25-26 sends
them to pool workers): a solver built in the parent pickles, and a forked
child that never touches the GPU unpickles it with its trees, data and
parameters intact."""
import multiprocessing as mp
import pickle

from laml_amd import sim
from laml_amd.solvers import EM_solver


def _child(blob, q):
    s = pickle.loads(blob)
    q.put((s.get_tree_newick(), s.get_params(), s.numsites, s._forest is None))


def test_solver_crosses_a_fork():
    tree, Q, codes, _ = sim.make_problem(30, 8, 3, 6, 1.0, 0.1, 0.1, 2)
    s = EM_solver([tree.newick()], {"charMtrx": sim.codes_to_charmtrx(tree, codes)}, {"Q": Q},
                  {"nu": 0.2, "phi": 0.05})
    ctx = mp.get_context("fork")
    q = ctx.Queue()
    p = ctx.Process(target=_child, args=(pickle.dumps(s), q))
    p.start()
    got = q.get(timeout=60)
    p.join(timeout=60)
    assert p.exitcode == 0
    assert got == (s.get_tree_newick(), s.get_params(), s.numsites, True)
