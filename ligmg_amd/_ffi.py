"""ctypes declarations of the C ABI (include/ligmg_amd.h, include/ligmg_graphs.h).

The structures below mirror the C structs field for field; keep them in the
same order as the headers.
"""
import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_SOLVER = os.path.join(_HERE, "libligmg_amd.so")
LIB_GRAPHS = os.path.join(_HERE, "libligmg_graphs.so")

i32, i64, u64, f64 = C.c_int32, C.c_int64, C.c_uint64, C.c_double
P = C.POINTER


class ligmg_options(C.Structure):
    _fields_ = [
        ("verbose", i32), ("always_elim", i32), ("top_elim", i32),
        ("constant_correction", i32), ("strength_metric", i32),
        ("max_levels", i32), ("min_size", i64), ("elim_degree", i32),
        ("tol", f64), ("thresh", f64), ("agg_iters", i32), ("iters", i32),
        ("pre_smooth", i32), ("post_smooth", i32), ("num_tvs", i32),
        ("tvs_smooth", i32), ("smoother_type", i32), ("kcycle", i32),
        ("outer_ksp_type", i32), ("kcycle_ksp_type", i32),
        ("kcycle_iters", i32), ("votes_to_seed", i64), ("max_agg_size", i64),
        ("agg_iter_factor", f64), ("agg_init_factor", f64),
        ("cheby_lower", f64), ("cheby_upper", f64), ("aff_scaling", i32),
        ("elim_factor", f64), ("elim_iters", i32),
        ("neighbor_of_neighbor", i32), ("seed", u64),
        ("cheby_eig_iters", i32), ("precision", i32), ("use_graphs", i32),
        ("spmv_slice", i32),
        ("ordering", i32),
        ("solve_layout", i32),
        ("tail_nnz", i32),
        ("rejection_ratio", f64),
        ("rejection_to_seed", i32),
        ("spmm_tune", i32),
    ]


class ligmg_info(C.Structure):
    _fields_ = [
        ("iters", i32), ("status", i32), ("history_len", i32),
        ("base_level", i32), ("rel_residual", f64),
        ("full_rel_residual", f64), ("time", f64), ("work", f64),
        ("cycle_complexity", f64), ("cf", f64), ("ecf", f64), ("wda", f64),
        ("setup_time", f64),
    ]


class ligmg_graph(C.Structure):
    _fields_ = [
        ("n", i64), ("nnz", i64), ("rowptr", P(i32)), ("colind", P(i32)),
        ("vals", P(f64)), ("ncomp", i32), ("comp_ptr", P(i64)),
    ]


# enums (include/ligmg_amd.h)
OK, ERR_INVALID, ERR_UNSUPPORTED, ERR_DEVICE, ERR_NOMEM, ERR_NUMERIC, ERR_IO = range(7)
KSP_CG, KSP_FCG, KSP_PIPEFCG, KSP_FGMRES = 0, 1, 2, 3
SMOOTHER_JACOBI, SMOOTHER_CHEBYSHEV = 0, 1
AFF_NONE, AFF_ROW, AFF_COLUMN, AFF_ROWCOLUMN, AFF_SYM, AFF_MAX = range(6)
STRENGTH_AFFINITY, STRENGTH_ALGEBRAIC_DISTANCE, STRENGTH_NONE = 0, 1, 2
PREC_FP64, PREC_MIXED = 0, 1
ORDER_NATURAL, ORDER_DEGREE = 0, 1
MEM_HOST, MEM_DEVICE = 0, 1
LEVEL_ELIM, LEVEL_AGG, LEVEL_COARSE = 0, 1, 2
NOALLOC_CC_MAX_N = 4194304  # LIGMG_NOALLOC_CC_MAX_N
LEVEL_NAMES = {LEVEL_ELIM: "elim", LEVEL_AGG: "agg", LEVEL_COARSE: "coarse"}
CONVERGED, DIVERGED_ITS, DIVERGED_BREAKDOWN, DIVERGED_NAN = range(4)
W_UNIT, W_UNIFORM, W_LOGNORMAL, W_GAUSS_DIST = range(4)

# AggState.state values (aggregation.hpp:20)
ST_SEED, ST_SEEDNEIGHBOR, ST_DECIDED, ST_DECIDEDROOT, ST_UNDECIDED = range(5)

HANDLE = C.c_void_p

SOLVER_PROTOS = {
    "ligmg_options_default": (None, [P(ligmg_options)]),
    "ligmg_create": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                               P(ligmg_options), i32, P(i64), P(HANDLE)]),
    "ligmg_solve": (C.c_int, [HANDLE, C.c_void_p, C.c_void_p, C.c_int,
                              P(ligmg_info), P(f64), i32]),
    "ligmg_solve_from": (C.c_int, [HANDLE, i32, C.c_void_p, C.c_void_p, C.c_int,
                                   P(ligmg_info), P(f64), i32]),
    "ligmg_solve_batch": (C.c_int, [HANDLE, i32, C.c_void_p, C.c_void_p, C.c_int,
                                    P(ligmg_info), P(i32)]),
    "ligmg_edge_grad": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                  C.c_void_p]),
    "ligmg_assemble_laplacian": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                           C.c_void_p]),
    "ligmg_destroy": (C.c_int, [HANDLE]),
    "ligmg_ordering": (C.c_int, [HANDLE, C.c_void_p]),
    "ligmg_last_error": (C.c_char_p, []),
    "ligmg_num_levels": (C.c_int, [HANDLE, P(i32)]),
    "ligmg_tail_start": (C.c_int, [HANDLE, P(i32)]),
    "ligmg_trim_memory": (C.c_int, []),
    "ligmg_dump_hierarchy": (C.c_int, [HANDLE, C.c_char_p]),
    "ligmg_load_hierarchy": (C.c_int, [C.c_char_p, P(ligmg_options), P(HANDLE)]),
    "ligmg_dump_csv": (C.c_int, [HANDLE, C.c_char_p, P(ligmg_info), C.c_char_p, i64]),
    "ligmg_setup_time": (C.c_int, [HANDLE, P(f64)]),
    "ligmg_level_info": (C.c_int, [HANDLE, i32, P(i32), P(i64), P(i64), P(i64)]),
    "ligmg_level_matrix": (C.c_int, [HANDLE, i32, P(i32), P(i32), P(f64)]),
    "ligmg_level_map": (C.c_int, [HANDLE, i32, P(i32)]),
    "ligmg_level_states": (C.c_int, [HANDLE, i32, P(i32), P(i32)]),
    "ligmg_level_strength": (C.c_int, [HANDLE, i32, P(f64)]),
    "ligmg_level_cheby": (C.c_int, [HANDLE, i32, P(f64), P(f64)]),
    "ligmg_level_spmv": (C.c_int, [HANDLE, i32, i32, P(f64), P(f64)]),
    "ligmg_apply_cycle": (C.c_int, [HANDLE, i32, i32, P(f64), P(f64)]),
    "ligmg_bench_spmv": (C.c_int, [HANDLE, i32, i32, i32, P(f64), P(i64)]),
    "ligmg_bench_smoother": (C.c_int, [HANDLE, i32, i32, i32, P(f64), P(i64)]),
    "ligmg_components": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, P(i64)]),
    "ligmg_components_noalloc": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_void_p, P(i64)]),
    "ligmg_bench_coarse": (C.c_int, [HANDLE, i32, i32, P(f64), P(f64)]),
    "ligmg_bench_spmm": (C.c_int, [HANDLE, i32, i32, i32, i32, P(f64), P(i64)]),
    "ligmg_check_units": (C.c_int, [i64, P(i32), P(i32), P(i64)]),
    "ligmg_coarse_apply_block": (C.c_int, [HANDLE, i32, P(f64), P(f64)]),
    "ligmg_bench_pinv": (C.c_int, [HANDLE, P(f64), P(f64), P(f64)]),
    "ligmg_dense_pinv": (C.c_int, [i64, P(i32), P(i32), P(f64), P(f64)]),
    "ligmg_pc_create": (C.c_int, [P(C.c_void_p)]),
    "ligmg_pc_begin": (C.c_int, [C.c_void_p, i64]),
    "ligmg_pc_add_row": (C.c_int, [C.c_void_p, i64, i64, P(i64), P(f64)]),
    "ligmg_pc_setup": (C.c_int, [C.c_void_p]),
    "ligmg_pc_apply": (C.c_int, [C.c_void_p, P(f64), P(f64), P(ligmg_info)]),
    "ligmg_pc_destroy": (C.c_int, [C.c_void_p]),
    "ligmg_largest_component": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, P(i64),
                                          P(i64)]),
    "ligmg_extract": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, i64, C.c_int,
                                C.c_void_p, C.c_void_p, C.c_void_p]),
    "ligmg_laplacian_nnz": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_int, P(i64)]),
    "ligmg_adjacency_to_laplacian": (C.c_int, [i64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                               C.c_void_p, C.c_void_p, C.c_void_p]),
    "ligmg_aggregate_with_strength": (C.c_int, [i64, P(i32), P(i32), P(f64),
                                                P(ligmg_options), P(i32), P(i32),
                                                P(i32), P(i64)]),
}

GRAPH_PROTOS = {
    "ligmg_graph_free": (None, [P(ligmg_graph)]),
    "ligmg_graph_read_mtx": (C.c_int, [C.c_char_p, P(ligmg_graph)]),
    "ligmg_gen_grid2d": (C.c_int, [i64, i64, C.c_int, f64, f64, u64, P(ligmg_graph)]),
    "ligmg_gen_grid3d": (C.c_int, [i64, i64, i64, C.c_int, f64, f64, u64, P(ligmg_graph)]),
    "ligmg_gen_rmat": (C.c_int, [C.c_int, C.c_int, f64, f64, f64, C.c_int, f64, f64,
                                 f64, C.c_int, u64, C.c_int, P(ligmg_graph)]),
    "ligmg_gen_ba": (C.c_int, [i64, C.c_int, C.c_int, f64, f64, u64, P(ligmg_graph)]),
    "ligmg_gen_rgg": (C.c_int, [i64, f64, C.c_int, f64, f64, u64, P(ligmg_graph)]),
    "ligmg_graph_union": (C.c_int, [P(ligmg_graph), C.c_int, P(ligmg_graph)]),
    "ligmg_graph_from_edges": (C.c_int, [i64, i64, P(i32), P(i32), P(f64), P(ligmg_graph)]),
    "ligmg_gen_rhs": (C.c_int, [i64, C.c_int, u64, C.c_int, P(i64), P(f64)]),
    "ligmg_rng_u01": (f64, [u64, u64, u64]),
    "ligmg_rng_normal": (f64, [u64, u64, u64]),
}


def bind(lib, protos):
    for name, (res, args) in protos.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


_graphs = None
_solver = None


def graphs_lib():
    global _graphs
    if _graphs is None:
        if not os.path.exists(LIB_GRAPHS):
            raise RuntimeError(f"{LIB_GRAPHS} is not built; run __graft_entry__.build()")
        _graphs = bind(C.CDLL(LIB_GRAPHS), GRAPH_PROTOS)
    return _graphs


def solver_lib():
    """The HIP solver library.  There is no CPU fallback: a missing or
    unloadable library is an error."""
    global _solver
    if _solver is None:
        try:
            # torch bundles its own libamdhip64.so.7; loading it first makes
            # both libraries share one HIP runtime in this process.
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_SOLVER):
            raise RuntimeError(f"{LIB_SOLVER} is not built; run __graft_entry__.build()")
        _solver = bind(C.CDLL(LIB_SOLVER), SOLVER_PROTOS)
    return _solver
