"""ctypes loading of libmgldg.so and the C ABI signatures (include/mgldg.h)."""
import ctypes as C
import os

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libmgldg.so")

SIZEFN = C.CFUNCTYPE(None, C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_void_p)


class MGParams(C.Structure):
    _fields_ = [("relaxation", C.c_int), ("omega", C.c_double), ("coarse", C.c_int), ("npre", C.c_int),
                ("npost", C.c_int), ("restriction", C.c_int), ("precision", C.c_int), ("coarse_dense_max", C.c_int),
                ("pool_off", C.c_int), ("row_tiles", C.c_int)]


class SolveParams(C.Structure):
    _fields_ = [("tol", C.c_double), ("maxit", C.c_int), ("criterion", C.c_int), ("flexible", C.c_int),
                ("precision", C.c_int), ("use_graph", C.c_int)]


class SolveStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("status", C.c_int), ("relres_true", C.c_double),
                ("relres_prec", C.c_double), ("seconds", C.c_double)]


class MGLDGError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mgldg error {code}: {msg}")
        self.code = code


# every symbol declared in include/mgldg.h with its ctypes signature
V, I, D, L = C.c_void_p, C.c_int, C.c_double, C.c_long
PD, PI, PL = C.POINTER(C.c_double), C.POINTER(C.c_int), C.POINTER(C.c_long)
SIGNATURES = {
    "mgldg_ctx_create": (I, [I, V, C.POINTER(V)]),
    "mgldg_ctx_destroy": (I, [V]),
    "mgldg_last_error": (C.c_char_p, [V]),
    "mgldg_ctx_synchronize": (I, [V]),
    "mgldg_quadtree_create": (I, [V, I, PD, PD, SIZEFN, V, I, I, I, C.POINTER(V)]),
    "mgldg_quadtree_info": (I, [V, PI]),
    "mgldg_quadtree_layer": (I, [V, I, PI, PI]),
    "mgldg_quadtree_destroy": (I, [V]),
    "mgldg_ldg_create": (I, [V, I, PI, D, D, C.POINTER(V)]),
    "mgldg_wireframe_create": (I, [V, I, I, PD, I, PI, C.POINTER(V)]),
    "mgldg_wireframe_read": (I, [V, I, C.c_char_p, C.POINTER(V)]),
    "mgldg_wireframe_get": (I, [V, PI, PD, PI]),
    "mgldg_wireframe_destroy": (I, [V]),
    "mgldg_ldg_create_simplex": (I, [V, I, I, PI, PI, D, D, C.POINTER(V)]),
    "mgldg_problem_sizes": (I, [V, PI]),
    "mgldg_problem_nodes": (I, [V, PD]),
    "mgldg_problem_bc_qpoints": (I, [V, PD, PI]),
    "mgldg_rhs": (I, [V, PD, PD, I, PD]),
    "mgldg_problem_export": (I, [V, I, PI, PI, PD, PL]),
    "mgldg_problem_destroy": (I, [V]),
    "mgldg_mg_create": (I, [V, C.POINTER(MGParams), C.POINTER(V)]),
    "mgldg_mg_create_p": (I, [V, C.POINTER(MGParams), I, PI, I, C.POINTER(V)]),
    "mgldg_mg_info": (I, [V, PL, I]),
    "mgldg_mg_level": (I, [V, I, PI, PI, PL]),
    "mgldg_mg_batch_kernels": (I, [V, I]),
    "mgldg_mg_batch_plan": (I, [V, PI, C.POINTER(C.c_float)]),
    "mgldg_agglomerate": (I, [V, I, PI, PI, PI]),
    "mgldg_mgridgen": (I, [I, PI, PD, PD, PI, PD, I, I, PI, PI, PI, PI]),
    "mgldg_mg_create_agglomerated": (I, [V, C.POINTER(MGParams), I, PI, C.POINTER(V)]),
    "mgldg_mg_export": (I, [V, I, I, PI, PI, PD, PL, PI, PI]),
    "mgldg_mg_colors": (I, [V, I, PI, PI]),
    "mgldg_mg_level_pool": (I, [V, I, PI, PI]),
    "mgldg_mg_level_tiles": (I, [V, I, PI, PI, PI]),
    "mgldg_mg_solve_cost": (I, [V, I, I, PI, PD, PI, PD, PI]),
    "mgldg_mg_destroy": (I, [V]),
    "mgldg_solve_device": (I, [V, V, V, I, C.POINTER(SolveParams), C.POINTER(SolveStats)]),
    "mgldg_solve": (I, [V, PD, PD, I, C.POINTER(SolveParams), C.POINTER(SolveStats)]),
    "mgldg_adjoint_device": (I, [V, V, V, V, V, I, C.POINTER(SolveParams), C.POINTER(SolveStats)]),
    "mgldg_rhs_device": (I, [V, V, V, I, V]),
    "mgldg_adjoint_maps_device": (I, [V, V, V, V, I]),
    "mgldg_spmv_device": (I, [V, I, I, V, V, I]),
    "mgldg_smooth_device": (I, [V, I, I, I, I, V, V, I]),
    "mgldg_vcycle_device": (I, [V, I, V, V, I]),
    "mgldg_vcycles_device": (I, [V, I, V, V, I, D, PD, PI]),
    "mgldg_mean_zero_device": (I, [V, L, V, I]),
    "mgldg_mean_zero_device_noalloc": (I, [V, L, V, I]),
    "mgldg_bsr_write_mm": (I, [C.c_char_p, I, I, I, I, PI, PI, PD]),
    "mgldg_problem_write_mm": (I, [V, I, C.c_char_p]),
    "mgldg_mg_write_mm": (I, [V, I, I, C.c_char_p]),
    "mgldg_fun_write": (I, [C.c_char_p, I, I, I, I, PD, PD]),
    "mgldg_problem_write_fun": (I, [V, PD, C.c_char_p]),
    "mgldg_problem_dump": (I, [V, PD, C.c_char_p]),
    "mgldg_mesh_write": (I, [C.c_char_p, I, I, PD, I, PI, I]),
    "mgldg_wireframe_write": (I, [V, C.c_char_p, I]),
    "mgldg_mg_vcycle_matrix": (I, [V, I, PD]),
    "mgldg_dense_write": (I, [C.c_char_p, I, I, PD]),
    "mgldg_profile_relax": (I, [V, I, I, I, PD, PD, PI]),
    "mgldg_profile_relax_part": (I, [V, I, I, I, I, PD, PD, PI]),
    "mgldg_profile_spmv": (I, [V, I, I, I, PD, PD]),
    "mgldg_setup_times": (I, [V, PD]),
    "mgldg_problem_setup_times": (I, [V, PD]),
}

_lib = None


def lib():
    """Load libmgldg.so (built in-tree by __graft_entry__.build()).  Raises if absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MGLDGError(-5, f"{LIB_PATH} is not built; run __graft_entry__.build()")
    # torch ships its own libamdhip64 (same soname, different file name).  Load
    # torch first so that libmgldg binds to the runtime torch already uses and
    # the process holds a single HIP runtime.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L_ = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L_, name)
        f.restype = res
        f.argtypes = args
    _lib = L_
    return L_


def check(ctx, status):
    if status != 0:
        h = None if ctx is None else ctx.h
        msg = lib().mgldg_last_error(h)
        raise MGLDGError(status, msg.decode() if msg else "")
    return status
