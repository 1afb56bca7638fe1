"""Writes ml_kats.json and the n64 data files from the reference's own tests.

Run in a container that holds the reference (read as text, nothing of it is
imported or executed):

    python tests/golden/make_ml_kats.py /root/reference

  - ``ml_nllh``: every case of laml_unit_tests/unit_tests_MLSolver.py (tree,
    character matrix, priors, nu, phi, expected negative log-likelihood),
    taken from the test source with ``ast.literal_eval``;
  - ``ml_poly_nllh``: case_tests/old/test6, the likelihood of a 64-leaf tree
    with a four-way polytomy at phi = 0, nu = 0.32 (sample_output.txt);
  - ``ml_optimize``: case_tests/old/test1 and test2, the optimum of branch
    lengths, nu and phi on the resolved 64-leaf tree from one random start
    (sample_output.txt);
  - the polytomous tree is copied next to this script (the character
    matrices and priors of these cases are already here).
"""
import ast
import json
import os
import re
import shutil
import sys


def unit_cases(path, rel):
    src = open(path).read()
    out = []
    for cls in [n for n in ast.parse(src).body if isinstance(n, ast.ClassDef)]:
        for fn in cls.body:
            if not isinstance(fn, ast.FunctionDef):
                continue
            env, params = {}, None
            for st in fn.body:
                if not (isinstance(st, ast.Assign) and isinstance(st.targets[0], ast.Name)):
                    continue
                try:
                    env[st.targets[0].id] = ast.literal_eval(st.value)
                except ValueError:
                    v = st.value
                    if isinstance(v, ast.Call) and getattr(v.func, "id", "") == "ML_solver":
                        params = ast.literal_eval(v.args[3])
            out.append({"name": "%s.%s" % (cls.name, fn.name),
                        "source": "%s:%d-%d" % (rel, fn.lineno, fn.end_lineno),
                        "tree": env["T"], "charMtrx": env["msa"],
                        "Q": [{str(k): v for k, v in q.items()} for q in env["Q"]],
                        "nu": params["nu"], "phi": params["phi"], "nllh": env["true_nllh"]})
    return out


def sample_nllh(path):
    text = open(path).read()
    m = re.search(r"Optimal negative-llh: \(?([0-9.]+)", text)
    return float(m.group(1))


def main(ref):
    here = os.path.dirname(os.path.abspath(__file__))
    rel = "laml_unit_tests/unit_tests_MLSolver.py"
    kats = {"ml_nllh": unit_cases(os.path.join(ref, rel), rel)}
    old = os.path.join(ref, "case_tests", "old")
    # the n64 character matrices and priors of these cases are byte-identical
    # to test3_charMtrx.txt (d0s32), test4_charMtrx.txt (d0) and
    # test4_prior.csv (both) already here; only the polytomous tree is new
    shutil.copyfile(os.path.join(old, "test6", "n64.tre"), os.path.join(here, "n64_poly.tre"))
    kats["ml_poly_nllh"] = {
        "source": "case_tests/old/test6/sample_output.txt:3",
        "tree_file": "n64_poly.tre", "charMtrx_file": "test3_charMtrx.txt",
        "prior_file": "test4_prior.csv", "phi": 0.0, "nu": 0.32,
        "nllh": sample_nllh(os.path.join(old, "test6", "sample_output.txt"))}
    # test1/test2 use the resolved n64 tree, which is test4.tre here
    kats["ml_optimize"] = [
        {"source": "case_tests/old/test1/sample_output.txt", "tree_file": "test4.tre",
         "charMtrx_file": "test4_charMtrx.txt", "prior_file": "test4_prior.csv",
         "nllh": sample_nllh(os.path.join(old, "test1", "sample_output.txt"))},
        {"source": "case_tests/old/test2/sample_output.txt", "tree_file": "test4.tre",
         "charMtrx_file": "test3_charMtrx.txt", "prior_file": "test4_prior.csv",
         "nllh": sample_nllh(os.path.join(old, "test2", "sample_output.txt"))}]
    with open(os.path.join(here, "ml_kats.json"), "w") as f:
        json.dump(kats, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
