"""ctypes front of ``libpmm_ref.so``, the C/OpenMP restatement of the EM
oracle (``pmm_ref.c``).  TEST INFRASTRUCTURE ONLY: tests hold it against the
Python oracle, and bench.py times it as the CPU baseline.

Build with ``make -C oracle`` (``__graft_entry__.build()`` does it)."""
import ctypes
import os

import numpy as np

from .em_oracle import DMAX, DMIN, postorder

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_i32p = ctypes.POINTER(ctypes.c_int32)
_i16p = ctypes.POINTER(ctypes.c_int16)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_f64p = ctypes.POINTER(ctypes.c_double)


class _Problem(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int32), ("K", ctypes.c_int32), ("c0", _i32p), ("c1", _i32p), ("parent", _i32p),
                ("root_of", _i32p), ("leaf_row", _i32p), ("codes", _i16p), ("qoff", _i32p),
                ("q", _f64p), ("free_", _u8p), ("phimin", ctypes.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libpmm_ref.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/libpmm_ref.so is not built (make -C oracle)")
        L = ctypes.CDLL(path)
        P = ctypes.POINTER(_Problem)
        L.pmm_ref_threads.restype = ctypes.c_int
        L.pmm_ref_nll.argtypes = [P, _f64p, ctypes.c_double, ctypes.c_double]
        L.pmm_ref_nll.restype = ctypes.c_double
        L.pmm_ref_estep.argtypes = [P, _f64p, ctypes.c_double, ctypes.c_double, _f64p, _f64p, _f64p]
        L.pmm_ref_estep.restype = ctypes.c_double
        L.pmm_ref_mstep.argtypes = [P, _f64p, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int, _f64p, _f64p, _f64p]
        L.pmm_ref_mstep.restype = None
        L.pmm_ref_em.argtypes = [P, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_double, ctypes.c_int, _f64p, _f64p, _f64p, _f64p]
        L.pmm_ref_em.restype = ctypes.c_int
        _LIB = L
    return _LIB


def _ptr(a, t):
    return a.ctypes.data_as(t)


class RefProblem:
    """One problem (binary trees over one character matrix) for the C oracle.

    c0, c1, parent: int arrays over nodes in post-order (children before
    parents; several trees concatenated); codes: [leaves, K] raw entries
    (0, an edit state, -1 for '?', -2 for an explicit -1) with leaf rows in
    the order leaves appear in the node order; Q: per-site dict state ->
    normalised prior; free: branches the M-step optimises."""

    def __init__(self, c0, c1, parent, codes, Q, free=None, phimin=0.0):
        self.c0 = np.ascontiguousarray(c0, dtype=np.int32)
        self.c1 = np.ascontiguousarray(c1, dtype=np.int32)
        self.parent = np.ascontiguousarray(parent, dtype=np.int32)
        n = len(self.c0)
        self.n = n
        root_of = np.empty(n, dtype=np.int32)
        for v in range(n - 1, -1, -1):
            root_of[v] = v if self.parent[v] < 0 else root_of[self.parent[v]]
        self.root_of = root_of
        leaf = self.c0 < 0
        self.leaf_row = np.full(n, -1, dtype=np.int32)
        self.leaf_row[leaf] = np.arange(int(leaf.sum()), dtype=np.int32)
        K = len(Q)
        self.K = K
        states = [sorted(s for s in q if s != 0) for q in Q]
        local = [{s: j + 1 for j, s in enumerate(st)} for st in states]
        raw = np.asarray(codes)
        enc = np.empty(raw.shape, dtype=np.int16)
        for k in range(K):
            col = raw[:, k]
            enc[:, k] = [x if x <= 0 else local[k][int(x)] for x in col]
        self.codes = np.ascontiguousarray(enc)
        self.qoff = np.zeros(K + 1, dtype=np.int32)
        self.qoff[1:] = np.cumsum([len(s) for s in states])
        self.q = np.asarray([Q[k][s] for k in range(K) for s in states[k]] + [0.0], dtype=np.float64)
        self.free = np.ascontiguousarray(np.ones(n) if free is None else free, dtype=np.uint8)
        self.s = _Problem(n, K, _ptr(self.c0, _i32p), _ptr(self.c1, _i32p), _ptr(self.parent, _i32p),
                          _ptr(self.root_of, _i32p), _ptr(self.leaf_row, _i32p), _ptr(self.codes, _i16p),
                          _ptr(self.qoff, _i32p), _ptr(self.q, _f64p), _ptr(self.free, _u8p), float(phimin))

    @classmethod
    def from_oracle(cls, o):
        """The problem an ``OracleEM`` holds (its trees in post-order, its
        character matrix and normalised priors, its fixed/polytomy marks);
        returns (problem, branch lengths)."""
        nodes = [v for r in o.trees for v in postorder(r)]
        idx = {id(v): i for i, v in enumerate(nodes)}
        c0 = [idx[id(v.children[0])] if v.children else -1 for v in nodes]
        c1 = [idx[id(v.children[1])] if v.children else -1 for v in nodes]
        parent = [idx[id(v.parent)] if v.parent is not None else -1 for v in nodes]
        rows = []
        for v in nodes:
            if not v.children:
                rows.append([-1 if x == "?" else (-2 if x == -1 else x) for x in o.charMtrx[v.label]])
        free = [not v.a["poly"] and not v.a["fixed"] for v in nodes]
        phimin = 1.0 / (o.numsites * len(o.charMtrx))
        p = cls(c0, c1, parent, np.asarray(rows), o.Q, free, phimin)
        return p, np.asarray([v.edge_length for v in nodes], dtype=np.float64)

    def nll(self, d, nu, phi):
        d = np.ascontiguousarray(d, dtype=np.float64)
        return lib().pmm_ref_nll(ctypes.byref(self.s), _ptr(d, _f64p), float(nu), float(phi))

    def estep(self, d, nu, phi):
        """(nll, counts [n, 5], R, Rt)."""
        d = np.ascontiguousarray(d, dtype=np.float64)
        counts = np.zeros((self.n, 5))
        R = ctypes.c_double()
        Rt = ctypes.c_double()
        nll = lib().pmm_ref_estep(ctypes.byref(self.s), _ptr(d, _f64p), float(nu), float(phi),
                                  _ptr(counts, _f64p), ctypes.byref(R), ctypes.byref(Rt))
        return nll, counts, R.value, Rt.value

    def mstep(self, counts, R, Rt, d, nu, phi, opt_nu=True, opt_phi=True, dmin=DMIN, dmax=DMAX):
        """Returns (d, nu, phi) after the M-step."""
        d = np.array(d, dtype=np.float64)
        counts = np.ascontiguousarray(counts, dtype=np.float64)
        x = ctypes.c_double(nu)
        y = ctypes.c_double(phi)
        lib().pmm_ref_mstep(ctypes.byref(self.s), _ptr(counts, _f64p), float(R), float(Rt), dmin, dmax,
                            int(opt_nu), int(opt_phi), _ptr(d, _f64p), ctypes.byref(x), ctypes.byref(y))
        return d, x.value, y.value

    def em(self, d, nu, phi, opt_nu=True, opt_phi=True, max_iter=1000, tol=1e-8, fixed_iters=0,
           dmin=DMIN, dmax=DMAX):
        """EM_optimization; returns (nll, iterations, d, nu, phi)."""
        d = np.array(d, dtype=np.float64)
        x = ctypes.c_double(nu)
        y = ctypes.c_double(phi)
        out = ctypes.c_double()
        it = lib().pmm_ref_em(ctypes.byref(self.s), dmin, dmax, int(opt_nu), int(opt_phi), int(max_iter),
                              float(tol), int(fixed_iters), _ptr(d, _f64p), ctypes.byref(x), ctypes.byref(y),
                              ctypes.byref(out))
        return out.value, it, d, x.value, y.value


def threads():
    return lib().pmm_ref_threads()
