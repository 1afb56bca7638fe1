"""Input and output formats of LAML's path (laml_libs/sequence_lib.py,
laml_libs/runner.py:323-435): character matrices, mutation priors,
duplicate-sequence handling, and the tree / parameter writers.

Readers return the reference's in-memory forms: a character matrix is a dict
cell -> list of ints with the missing symbol '?'; priors are a list (one per
site, in file order) of dicts state -> probability.
"""
import json
import pickle
from collections import defaultdict

from .tree import Node, read_tree_newick

DMIN = 0.005  # laml_libs/core_constants.py:6

# sequence_lib.py:8
RECOGNIZED_MISSING = {'-', '?', '-1'}


def check_missing(seen_missing, x):
    """sequence_lib.py:125-137: a symbol is missing if it was seen as missing,
    is one of '-', '?', '-1', is alphabetic, or is a negative integer."""
    if x in seen_missing or x in RECOGNIZED_MISSING:
        return True
    if x.isalpha():
        return True
    try:
        return int(x) < 0
    except ValueError:
        return False


def read_charMtrx(lines, delimiter=",", masked_symbol=None, replace_mchar='?', convert_to_int=True,
                  stop_key=None):
    """sequence_lib.py:139-175.  ``lines``: an iterable of text lines (an open
    file).  Returns (matrix, site_names)."""
    it = iter(lines)
    site_names = next(it).strip().split(delimiter)
    if site_names[0] == 'cell_name' or "cell" in site_names[0]:
        site_names = site_names[1:]
    seen_missing = {masked_symbol} if masked_symbol is not None else set()
    D = {}
    for line in it:
        if stop_key is not None and line.startswith(stop_key):
            break
        parts = line.strip().split(delimiter)
        seq = []
        for x in parts[1:]:
            if check_missing(seen_missing, x):
                seen_missing.add(x)
                seq.append(replace_mchar if replace_mchar is not None else x)
            else:
                seq.append(int(x) if convert_to_int else x)
        D[parts[0]] = seq
    return D, site_names


def read_fasta(lines):
    """sequence_lib.py:111-123: '>' names, '|'-separated integer sequences,
    '_' lines separate matrices; returns a list of dicts."""
    S, D, name = [], {}, None
    for line in lines:
        if line.startswith(">"):
            name = line.strip()[1:]
        elif line.startswith("_"):
            S.append(D)
            D = {}
        else:
            D[name] = [int(x) for x in line.strip().split("|")]
    return S


def read_sequences(inFile, filetype="charMtrx", delimiter=",", masked_symbol=None, suppress_warnings=False,
                   replace_mchar='?'):
    """sequence_lib.py:102-109."""
    with open(inFile) as fin:
        if filetype == "fasta":
            return read_fasta(fin)
        if filetype == "charMtrx":
            return read_charMtrx(fin, delimiter=delimiter, masked_symbol=masked_symbol, replace_mchar=replace_mchar)
    raise ValueError("unknown sequence file type %r" % (filetype,))


def write_sequences(char_mtrx, nsites, outFile, delimiter=","):
    """sequence_lib.py:86-100: header 'cell,r0,r1,...', one row per cell."""
    with open(outFile, "w") as fout:
        fout.write("cell" + "".join(delimiter + "r" + str(i) for i in range(nsites)) + "\n")
        for cell, row in char_mtrx.items():
            fout.write(cell + "".join(delimiter + str(x) for x in row) + "\n")


def read_Q(inFile):
    """sequence_lib.py:177-188: 'char,state,prob' lines -> {char: {state: prob}}."""
    Q = {}
    with open(inFile) as fin:
        for line in fin:
            char, state, prob = line.strip().split(',')
            Q.setdefault(int(char), {})[int(state)] = float(prob)
    return Q


def _priors_from_dict(priors, msa, site_names):
    """The dictionary branch of read_priors (sequence_lib.py:256-323):
    sites in sorted key order, state 0 given prior 0; with site names that
    do not match the keys, a mapping by position (equal counts) or by the
    digits of the names (with an offset when keys are missing), and uniform
    priors over the observed states for a site without a key."""
    keys = sorted(priors.keys())
    if site_names is not None:
        try:
            names_as_keys = sorted(int(x[1:]) for x in site_names)
        except ValueError:
            names_as_keys = None
        if keys != names_as_keys:
            mapping = {}
            if len(site_names) == len(keys):
                for i, name in enumerate(site_names):
                    mapping[name] = keys[i]
            else:
                digits = [int(''.join(ch for ch in name if ch.isdigit())) for name in site_names]
                all_present = all(dg in priors for dg in digits)
                offset = 0 if all_present else min(digits) - min(keys)
                for name, dg in zip(site_names, digits):
                    mapping[name] = dg - offset
                Q = []
                for i, name in enumerate(site_names):
                    pn = mapping[name]
                    if pn in priors:
                        q = {int(x): priors[pn][x] for x in priors[pn]}
                    else:
                        M = set(msa[c][i] for c in msa if msa[c][i] not in (0, "?"))
                        q = {x: 1 / len(M) for x in M} if M else {"1": 1.0}
                        q[0] = 0
                    Q.append(q)
                return Q
    Q = []
    for i in keys:
        q = {int(x): priors[i][x] for x in priors[i]}
        q[0] = 0
        Q.append(q)
    return Q


def read_priors(pfile, msa, site_names=None):
    """sequence_lib.py:233-363.  CSV ('site,state,prob' lines, optional
    header, site names with a one-letter prefix allowed), JSON (site ->
    {state: prob}), or a pickled dict (a file the user supplies)."""
    ext = pfile.strip().split(".")[-1]
    if ext in ("pkl", "pickle"):
        with open(pfile, "rb") as f:
            priors = pickle.load(f)
        return _priors_from_dict(priors, msa, site_names)
    if ext == "json":
        with open(pfile) as f:
            raw = json.load(f)
        priors = {int(k): {int(s): float(p) for s, p in v.items()} for k, v in raw.items()}
        return _priors_from_dict(priors, msa, site_names)
    if ext != "csv":
        raise ValueError("unknown priors file type %r" % (ext,))
    Q, Qi, seen = [], {}, set()
    with open(pfile) as fin:
        lines = fin.readlines()
    tokens = lines[0].split(',')
    if not tokens[1].isnumeric() and not tokens[2].isnumeric():
        lines = lines[1:]
    charname_is_str = not lines[0].split(',')[0].isnumeric()
    for line in lines:
        site_idx, char_state, prob = line.strip().split(',')
        site_idx = int(site_idx[1:]) if charname_is_str else int(site_idx)
        if site_idx not in seen:
            if seen:
                Q.append(Qi)
                Qi = {}
            seen.add(site_idx)
        Qi[int(char_state)] = float(prob)
    Q.append(Qi)
    return Q


def alphabet_size(mtx):
    """sequence_lib.py:202-213: (max, min, mean) number of distinct symbols
    per entry of ``mtx``."""
    sizes = [len(set(mtx[k])) for k in mtx]
    return max(sizes), min(sizes), sum(sizes) / len(sizes)


def dedup(char_mtrx, input_tree_strings):
    """sequence_lib.py:10-52: cells with identical sequences are represented
    by the first of them; the others are pruned from the input trees.
    Returns (has_dup, dup_map representative -> [others], char_mtrx, trees)."""
    seq_to_taxa = defaultdict(list)
    for taxon, seq in char_mtrx.items():
        seq_to_taxa[tuple(seq)].append(taxon)
    if len(seq_to_taxa) == len(char_mtrx):
        return False, {}, char_mtrx, input_tree_strings
    dup_map = {}
    for group in seq_to_taxa.values():
        if len(group) > 1:
            dup_map[group[0]] = group[1:]
    out = []
    for s in input_tree_strings:
        tree = read_tree_newick(s)
        for rep in dup_map:
            tree = tree.extract_tree_without(dup_map[rep], suppress_unifurcations=True)
        out.append(tree.newick())
    return True, dup_map, char_mtrx, out


def add_dup(input_tree_strings, dup_map, dmin=DMIN):
    """sequence_lib.py:54-84: each representative becomes a node
    'node_dummy_<rep>' (taking its branch) with the representative and its
    duplicates below at length dmin."""
    out = []
    for s in input_tree_strings:
        tree = read_tree_newick(s)
        l2n = tree.label_to_node(selection='leaves')
        for rep, dups in dup_map.items():
            node = l2n.get(rep)
            if node is None:
                continue
            p = node.parent
            dummy = Node(label="node_dummy_%s" % rep, edge_length=node.edge_length)
            p.remove_child(node)
            p.add_child(dummy)
            node.edge_length = dmin
            dummy.add_child(node)
            for taxon in dups:
                dummy.add_child(Node(label=taxon, edge_length=dmin))
        out.append(tree.newick())
    return out


def _stem_wrapped(tree):
    """runner.py:341-345: the tree text with the stem as an extra pair of
    parentheses ('[&R] ' kept in front when present)."""
    text = tree.newick()
    parts = text.split()
    if len(parts) > 1:
        return "".join([parts[0], "(", parts[1][:-1], ");"])
    return "".join(["(", parts[0][:-1], ");"])


def scaled_trees_text(opt_trees, timescale=1.0):
    """runner.py:330-345: every branch divided by tree height / timescale.
    Returns (text, mutation rate of the last tree)."""
    lines, rate = [], None
    for tstr in opt_trees:
        tree = read_tree_newick(tstr)
        scale = tree.height() / float(timescale)
        for node in tree.traverse_preorder():
            node.edge_length = node.edge_length / scale
        rate = scale
        lines.append(_stem_wrapped(tree))
    return "\n".join(lines) + "\n", rate


def collapsed_trees_text(opt_trees, timescale=1.0, dmin=DMIN):
    """runner.py:347-360: branches up to 1.2 dmin collapsed, then scaled;
    zero or missing lengths become dmin / scale."""
    lines = []
    for tstr in opt_trees:
        tree = read_tree_newick(tstr)
        tree.collapse_short_branches(dmin + 0.2 * dmin)
        scale = tree.height() / float(timescale)
        for node in tree.traverse_preorder():
            node.edge_length = node.edge_length / scale if node.edge_length else dmin / scale
        lines.append(_stem_wrapped(tree))
    return "\n".join(lines) + "\n"


def write_outputs(prefix, opt_trees, annotations, phi, nu, nllh, timescale=1.0, dup_map=None):
    """The output files of runner.py:323-435: <prefix>_trees.nwk,
    <prefix>_trees.collapsed.nwk, <prefix>_annotations.txt, <prefix>_params.txt."""
    from .annotate import params_text
    if dup_map:
        opt_trees = add_dup(opt_trees, dup_map)
    text, rate = scaled_trees_text(opt_trees, timescale)
    paths = {k: "%s_%s" % (prefix, k) for k in ("trees.nwk", "trees.collapsed.nwk", "annotations.txt", "params.txt")}
    with open(paths["trees.nwk"], "w") as f:
        f.write(text)
    with open(paths["trees.collapsed.nwk"], "w") as f:
        f.write(collapsed_trees_text(opt_trees, timescale))
    with open(paths["annotations.txt"], "w") as f:
        f.write(annotations)
    with open(paths["params.txt"], "w") as f:
        f.write(params_text(phi, nu, nllh, rate))
    return paths
