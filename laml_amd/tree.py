"""Rooted trees with the treeswift behaviour the LAML solvers rely on.

The reference builds every tree with treeswift (>= 1.1.39, not a dependency
here): ``read_tree_newick``, ``Tree.newick``, post/pre-order traversal with
children in Newick order, ``suppress_unifurcations`` (ML_solver.py:29-30),
``resolve_polytomies`` (EM_solver.py:35) and ``collapse_short_branches``
(ML_solver.py:317).  This module restates those operations so that traversal
order, branch order and the written Newick text match.
"""
from collections import deque

__all__ = ["Node", "Tree", "read_tree_newick"]


class Node:
    __slots__ = ("label", "edge_length", "children", "parent", "__dict__")

    def __init__(self, label=None, edge_length=None):
        self.label = label
        self.edge_length = edge_length
        self.children = []
        self.parent = None

    # -- structure ---------------------------------------------------------
    def is_leaf(self):
        return len(self.children) == 0

    def is_root(self):
        return self.parent is None

    def add_child(self, child):
        self.children.append(child)
        child.parent = self

    def remove_child(self, child):
        self.children.remove(child)
        child.parent = None

    def child_nodes(self):
        return list(self.children)

    def get_parent(self):
        return self.parent

    def set_parent(self, parent):
        self.parent = parent

    def get_edge_length(self):
        return self.edge_length

    def set_edge_length(self, length):
        self.edge_length = None if length is None else float(length)

    def get_label(self):
        return self.label

    def set_label(self, label):
        self.label = label

    # -- traversal ---------------------------------------------------------
    def traverse_postorder(self):
        stack = [(self, False)]
        while stack:
            v, done = stack.pop()
            if done:
                yield v
            else:
                stack.append((v, True))
                for c in reversed(v.children):
                    stack.append((c, False))

    def traverse_preorder(self):
        """treeswift's preorder: a node, then the subtrees of its children
        from the last child to the first (a stack that pops the last child
        pushed).  The reference's outputs are in this order: the node lines
        of its annotation files (runner.py:421-428), e.g.
        case_tests/test1/example3_annotations.txt lists I0, I1, c, I2, b, a
        for ((a,b),c)."""
        stack = [self]
        while stack:
            v = stack.pop()
            yield v
            stack.extend(v.children)

    def traverse_leaves(self):
        """The leaves in preorder (treeswift's traverse_leaves)."""
        for v in self.traverse_preorder():
            if not v.children:
                yield v

    def newick(self):
        out = {}
        for v in self.traverse_postorder():
            if not v.children:
                out[v] = "" if v.label is None else str(v.label)
                continue
            parts = ["("]
            for c in v.children:
                parts.append(out.pop(c))
                if c.edge_length is not None:
                    parts.append(":" + _fmt_len(c.edge_length))
                parts.append(",")
            parts[-1] = ")"
            if v.label is not None:
                parts.append(str(v.label))
            out[v] = "".join(parts)
        return out[self]

    def __repr__(self):
        return "Node(label=%r, edge_length=%r, n_children=%d)" % (self.label, self.edge_length, len(self.children))


def _fmt_len(x):
    if isinstance(x, int):
        return str(x)
    if isinstance(x, float) and x.is_integer():
        return str(int(x))
    return str(x)


class Tree:
    def __init__(self, root=None, is_rooted=False):
        self.root = Node() if root is None else root
        self.is_rooted = is_rooted

    def traverse_postorder(self):
        return self.root.traverse_postorder()

    def traverse_preorder(self):
        return self.root.traverse_preorder()

    def traverse_leaves(self):
        return self.root.traverse_leaves()

    def num_nodes(self, leaves=True, internal=True):
        n = 0
        for v in self.traverse_preorder():
            if (leaves and v.is_leaf()) or (internal and not v.is_leaf()):
                n += 1
        return n

    def label_to_node(self, selection="leaves"):
        nodes = self.traverse_leaves() if selection == "leaves" else self.traverse_preorder()
        return {v.label: v for v in nodes}

    def newick(self):
        e = self.root.edge_length
        suffix = ";" if e is None else ":" + _fmt_len(e) + ";"
        s = self.root.newick() + suffix
        return ("[&R] " + s) if self.is_rooted else s

    def suppress_unifurcations(self):
        """Splice out every node with exactly one child, adding its branch to
        the child's and passing its label down when the child has none."""
        q = deque([self.root])
        while q:
            node = q.popleft()
            if len(node.children) != 1:
                q.extend(node.children)
                continue
            child = node.children.pop()
            if node.parent is None:
                self.root = child
                child.parent = None
            else:
                par = node.parent
                par.remove_child(node)
                par.add_child(child)
            if node.edge_length is not None:
                if child.edge_length is None:
                    child.edge_length = 0.0
                child.edge_length += node.edge_length
            if child.label is None and node.label is not None:
                child.label = node.label
            q.append(child)

    def resolve_polytomies(self):
        """Arbitrarily resolve polytomies: pop the last two children into a new
        zero-length child until every node is at most binary.  Returns the
        created nodes."""
        created = []
        q = deque([self.root])
        while q:
            node = q.popleft()
            while len(node.children) > 2:
                c1 = node.children.pop()
                c2 = node.children.pop()
                nn = Node(edge_length=0.0)
                node.add_child(nn)
                nn.add_child(c1)
                nn.add_child(c2)
                created.append(nn)
            q.extend(node.children)
        return created

    def collapse_short_branches(self, threshold):
        """Contract internal non-root branches of length <= threshold (None
        counts as 0), in postorder: the children of a contracted node are
        appended to its parent's children and keep their own lengths (the
        contracted branch's length is dropped).  This is what the reference's
        writer produced: examples/out_example2_api/*_trees.collapsed.nwk has
        the contracted nodes' children at the end of their new parent's list,
        and its scale (height / timescale) is the height with the contracted
        lengths removed."""
        if threshold < 0:
            raise ValueError("threshold must be non-negative")
        for node in list(self.traverse_postorder()):
            if node.edge_length is None:
                node.edge_length = 0.0
            if node.edge_length <= threshold and not node.is_leaf() and not node.is_root():
                par = node.parent
                par.children.remove(node)
                node.parent = None
                for c in node.children:
                    c.parent = par
                par.children.extend(node.children)
                node.children = []

    def extract_tree_without(self, labels, suppress_unifurcations=True):
        """A copy of the tree without the leaves named in ``labels`` (internal
        nodes left without children are removed too), unifurcations
        suppressed (what dedup needs from treeswift's method of that name)."""
        drop = set(labels)
        t = read_tree_newick(self.newick())
        t.is_rooted = self.is_rooted
        for v in list(t.traverse_postorder()):
            if v.is_root():
                continue
            if (v.is_leaf() and v.label in drop) or (not v.children and v.label is None):
                par = v.parent
                par.children.remove(v)
                v.parent = None
        if suppress_unifurcations:
            t.suppress_unifurcations()
        return t

    def height(self):
        """Largest distance from the top of the root's branch to a leaf
        (treeswift's weighted height: the root's own edge length counts).
        The reference's writer divides by it: its example3 run logs a
        pre-scaling height of 1.1061115018134888 for a tree whose stem is
        0.005 and whose leaves sit 1.1011115 below the root."""
        best = 0.0
        dist = {}
        for v in self.traverse_preorder():
            dist[v] = (0.0 if v.parent is None else dist[v.parent]) + (v.edge_length or 0.0)
            if v.is_leaf():
                best = max(best, dist[v])
        return best


def read_tree_newick(text):
    """Parse one Newick string (labels kept as str, lengths as float)."""
    s = text.strip()
    rooted = False
    if s.startswith("[&R]"):
        rooted = True
        s = s[4:].strip()
    elif s.startswith("[&U]"):
        s = s[4:].strip()
    if s.endswith(";"):
        s = s[:-1]
    root = Node()
    cur = root
    stack = []
    i, n = 0, len(s)
    delims = ",():;"
    while i < n:
        ch = s[i]
        if ch == "(":
            c = Node()
            cur.add_child(c)
            stack.append(cur)
            cur = c
            i += 1
        elif ch == ",":
            if not stack:
                raise ValueError("malformed Newick: ',' at top level")
            c = Node()
            stack[-1].add_child(c)
            cur = c
            i += 1
        elif ch == ")":
            if not stack:
                raise ValueError("malformed Newick: unbalanced ')'")
            cur = stack.pop()
            i += 1
        elif ch == ":":
            j = i + 1
            while j < n and s[j] not in delims:
                j += 1
            cur.edge_length = float(s[i + 1:j])
            i = j
        elif ch.isspace():
            i += 1
        else:
            if ch == "'":
                j = s.index("'", i + 1)
                cur.label = s[i + 1:j]
                i = j + 1
                continue
            j = i
            while j < n and s[j] not in delims:
                j += 1
            cur.label = s[i:j].strip()
            i = j
    if stack:
        raise ValueError("malformed Newick: unbalanced '('")
    return Tree(root, is_rooted=rooted)
