"""ctypes binding of the C ABI in include/laml_pmm.h.

The shared library is built in-tree by ``laml_amd.build``.  Nothing here falls
back to a CPU implementation: a missing or unloadable library raises.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblaml_pmm.so")

LAML_OK = 0
LAML_OPT_PHI = 1
LAML_OPT_NU = 2
LAML_MSTEP_FAILED_D = 1
LAML_MSTEP_FAILED_NU = 2
PASS_LEVELS = 0
PASS_TREES = 1
PASS_HYBRID = 2
CHAINS_OUTSIDE = 1
CHAINS_INSIDE = 2

CODE_UNEDITED = 0
CODE_MISSING = -1
CODE_SILENCED = -2

_i32p = ctypes.POINTER(ctypes.c_int32)
_u8p = ctypes.POINTER(ctypes.c_uint8)
_u32p = ctypes.POINTER(ctypes.c_uint32)
_i16p = ctypes.POINTER(ctypes.c_int16)
_f64p = ctypes.POINTER(ctypes.c_double)


class ForestDesc(ctypes.Structure):
    _fields_ = [
        ("num_problems", ctypes.c_int32),
        ("num_trees", ctypes.c_int32),
        ("num_nodes", ctypes.c_int32),
        ("num_leaves", ctypes.c_int32),
        ("num_levels", ctypes.c_int32),
        ("level_off", _i32p),
        ("child0", _i32p),
        ("child1", _i32p),
        ("parent", _i32p),
        ("tree_of", _i32p),
        ("free_branch", _u8p),
        ("tree_root", _i32p),
        ("tree_problem", _i32p),
        ("tree_nsites", _i32p),
        ("site_qoff", _i32p),
        ("prior", _f64p),
        ("leaf_codes", _i16p),
        ("problem_nobs", _f64p),
        ("problem_phimin", _f64p),
        ("problem_flags", _u32p),
        ("dmin", ctypes.c_double),
        ("dmax", ctypes.c_double),
        ("site_qsum", _f64p),
    ]


_i8p = ctypes.POINTER(ctypes.c_int8)


class UltraDesc(ctypes.Structure):
    _fields_ = [
        ("num_classes", ctypes.c_int32),
        ("num_levels", ctypes.c_int32),
        ("level_problem_off", _i32p),
        ("up_cls", _i32p),
        ("dn_cls", _i32p),
        ("up_off", _f64p),
        ("dn_off", _f64p),
        ("cls_parent", _i32p),
        ("cls_pedge", _i32p),
        ("inc_off", _i32p),
        ("inc_branch", _i32p),
        ("inc_sign", _i8p),
        ("problem_has_free", _u8p),
    ]


# name -> (restype, argtypes); every symbol include/laml_pmm.h declares
SIGNATURES = {
    "laml_forest_create": (ctypes.c_int, [ctypes.POINTER(ForestDesc), ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_void_p)]),
    "laml_forest_validate": (ctypes.c_int, [ctypes.POINTER(ForestDesc)]),
    "laml_forest_destroy": (None, [ctypes.c_void_p]),
    "laml_forest_set_checkpoint": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "laml_forest_inside_rows": (ctypes.c_int64, [ctypes.c_void_p]),
    "laml_forest_set_flags": (ctypes.c_int, [ctypes.c_void_p, _u32p]),
    "laml_forest_labels": (ctypes.c_int, [ctypes.c_void_p, _i16p, ctypes.c_int64]),
    "laml_nll": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 5 + [ctypes.c_void_p]),
    "laml_estep": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 6 + [ctypes.c_void_p]),
    "laml_outside": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 3 + [ctypes.c_void_p]),
    "laml_grad": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 6 + [ctypes.c_void_p]),
    "laml_mstep": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.c_void_p] * 5 + [ctypes.c_void_p]),
    "laml_em": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_int32, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p,
                               ctypes.c_void_p, ctypes.c_void_p]),
    "laml_forest_set_ultrametric": (ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(UltraDesc)]),
    "laml_forest_set_heights": (ctypes.c_int, [ctypes.c_void_p, _f64p]),
    "laml_forest_set_pass_mode": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "laml_forest_pass_mode": (ctypes.c_int32, [ctypes.c_void_p]),
    "laml_forest_set_fusion": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32]),
    "laml_forest_fusion": (ctypes.c_int32, [ctypes.c_void_p]),
    "laml_forest_set_chains": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32, ctypes.c_int32, ctypes.c_int32]),
    "laml_forest_get_chains": (ctypes.c_int, [ctypes.c_void_p] + [ctypes.POINTER(ctypes.c_int32)] * 3),
    "laml_forest_chains": (ctypes.c_int32, [ctypes.c_void_p]),
    "laml_posteriors": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _f64p, _f64p, ctypes.c_int64,
                                         ctypes.c_void_p]),
    "laml_posteriors_noalloc": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, _f64p, _f64p, ctypes.c_int64,
                                                 ctypes.c_void_p]),
    "laml_quartet_em": (ctypes.c_int, [ctypes.c_void_p, ctypes.c_int32] + [ctypes.c_void_p] * 6 +
                        [ctypes.c_int32, ctypes.c_double] + [ctypes.c_void_p] * 4),
    "laml_forest_get_shapes": (ctypes.c_int, [ctypes.c_void_p, _i32p, _i32p, _i32p]),
    "laml_forest_set_shapes": (ctypes.c_int, [ctypes.c_void_p, _i32p, _i32p, ctypes.c_int32]),
    "laml_forest_ultra_status": (ctypes.c_int, [ctypes.c_void_p, _i32p]),
    "laml_forest_mstep_status": (ctypes.c_int, [ctypes.c_void_p, _i32p]),
    "laml_last_error": (ctypes.c_char_p, []),
    "laml_version": (ctypes.c_char_p, []),
}

_lib = None


class LamlError(RuntimeError):
    pass


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise LamlError("liblaml_pmm.so is not built; run `python -m laml_amd.build` "
                        "(the GPU path has no CPU fallback)")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != LAML_OK:
        msg = load().laml_last_error().decode(errors="replace")
        raise LamlError("laml error %d: %s" % (rc, msg))


def ptr(t):
    """Device (or host) address of a torch tensor / numpy array as void*."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return ctypes.c_void_p(t.data_ptr())
    return ctypes.c_void_p(t.ctypes.data)
