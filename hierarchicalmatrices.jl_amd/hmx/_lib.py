"""ctypes binding of libhmx.so (the C ABI declared in include/hmx.h).

The library is loaded from this package directory. There is no fallback: if
libhmx.so is missing or fails to load, importing the package raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhmx.so")

HM_OK = 0
HM_ERR_INVALID = 1
HM_ERR_NOT_DESCENDING = 2
HM_ERR_DEGENERATE = 3
HM_ERR_OOM = 4
HM_ERR_HIP = 5
HM_ERR_UNSUPPORTED = 6
HM_DENSE_NOALLOC_MAX = 8192  # items per hm_dense_gemm_noalloc call

HM_USER = 4  # kernel id of a host function f (hm_plan_set_leaf_data)

HM_PLAN_STORED = 0
HM_PLAN_MATRIX_FREE = 1
HM_PLAN_FAR_F32 = 2
HM_PLAN_TREE_ONLY = 4
HM_PLAN_HYBRID = 8
HM_MV_MATRIX_FREE = 1
HM_MV_CHEB_BASIS = 2
HM_MV_TUNED = 4
HM_MV_FAR_F32 = 8
HM_MV_FAR_BF16X3 = 16
HM_MV_MFMA = 32
HM_MV_NO_MFMA = 64
HM_MV_FAR_BF16X6 = 128

INFO_KEYS = [
    "nrows", "ncols", "rank", "blocksize", "kernel", "n_d", "n_k1", "n_k2", "n_dense", "n_far",
    "n_rowclusters", "n_colclusters", "depth", "dense_entries", "basis_rows_u", "basis_rows_v",
    "bytes_stored", "bytes_ref_layout", "flags", "n_rowchunks", "n_colchunks", "nonfinite",
    "n_instances", "from_blocks", "float32", "nested",
]

_vp = ctypes.c_void_p
_i64 = ctypes.c_int64
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int64)


class HMError(RuntimeError):
    """A libhmx call returned a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"libhmx error {code}: {msg}")
        self.code = code


def _load():
    # One HIP runtime per process: torch ships its own libamdhip64 (soname
    # libamdhip64.so.7, requested by torch as "libamdhip64.so"). Loading torch
    # first lets libhmx's NEEDED libamdhip64.so.7 resolve to that same copy;
    # loading libhmx first would pull in /opt/rocm's copy and torch would then
    # load a second runtime beside it.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"libhmx.so not found at {LIB_PATH}; build it with `python -c 'import "
            f"__graft_entry__ as g; g.build()'` (hipcc, gfx950)")
    L = ctypes.CDLL(LIB_PATH)
    L.hm_last_error.restype = ctypes.c_char_p
    L.hm_last_error.argtypes = []
    L.hm_version.restype = ctypes.c_int
    L.hm_blockrank.restype = ctypes.c_int
    L.hm_blockrank.argtypes = [ctypes.c_int]
    for name in ("hm_chebyshevpoints", "hm_chebyshevbarycentricweights"):
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
    L.hm_indsplit.restype = ctypes.c_int
    L.hm_indsplit.argtypes = [_dp, _i64, _i64, _i64, ctypes.c_double, ctypes.c_double, _ip]
    L.hm_plan_from_points.restype = ctypes.c_int
    L.hm_plan_from_points.argtypes = [ctypes.c_int, _dp, _i64, _dp, _i64, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_int, ctypes.c_int, ctypes.c_uint, _vp,
                                      ctypes.POINTER(_vp)]
    L.hm_plan_from_blocks.restype = ctypes.c_int
    L.hm_plan_from_blocks.argtypes = [_i64, _i64, _i64, ctypes.POINTER(ctypes.c_int32), _ip, _ip,
                                      _ip, _ip, _ip, ctypes.POINTER(_dp), ctypes.POINTER(_dp),
                                      ctypes.POINTER(_dp), ctypes.POINTER(_dp), _vp,
                                      ctypes.POINTER(_vp)]
    L.hm_plan_from_blocks_f32.restype = ctypes.c_int
    L.hm_plan_from_blocks_f32.argtypes = [_i64, _i64, _i64, ctypes.POINTER(ctypes.c_int32), _ip,
                                          _ip, _ip, _ip, _ip, ctypes.POINTER(_fp),
                                          ctypes.POINTER(_fp), ctypes.POINTER(_fp),
                                          ctypes.POINTER(_fp), _vp, ctypes.POINTER(_vp)]
    L.hm_plan_set_leaf_data.restype = ctypes.c_int
    L.hm_plan_set_leaf_data.argtypes = [_vp, _dp, _i64, _dp, _i64, _vp]
    L.hm_plan_set_leaf_data_f32.restype = ctypes.c_int
    L.hm_plan_set_leaf_data_f32.argtypes = [_vp, _fp, _i64, _fp, _i64, _vp]
    L.hm_bary_evaluate.restype = ctypes.c_int
    L.hm_bary_evaluate.argtypes = [ctypes.c_int, _vp, _vp, _vp, _vp, _vp, _i64, _vp, _vp, _vp, _vp]
    L.hm_bary_evaluate_f32.restype = ctypes.c_int
    L.hm_bary_evaluate_f32.argtypes = L.hm_bary_evaluate.argtypes
    L.hm_bary_basis.restype = ctypes.c_int
    L.hm_bary_basis.argtypes = [ctypes.c_int, _vp, _vp, _i64, _vp, _vp, _i64, _vp]
    L.hm_bary_basis_f32.restype = ctypes.c_int
    L.hm_bary_basis_f32.argtypes = L.hm_bary_basis.argtypes
    L.hm_plan_batched.restype = ctypes.c_int
    L.hm_plan_batched.argtypes = [ctypes.c_int, _i64, ctypes.POINTER(_dp), _ip,
                                  ctypes.POINTER(_dp), _ip, _dp, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_uint, _vp, ctypes.POINTER(_vp)]
    L.hm_plan_batch_offsets.restype = ctypes.c_int
    L.hm_plan_batch_offsets.argtypes = [_vp, _ip, _ip]
    L.hm_plan_destroy.restype = ctypes.c_int
    L.hm_plan_destroy.argtypes = [_vp]
    L.hm_plan_info.restype = ctypes.c_int
    L.hm_plan_info.argtypes = [_vp, _ip, ctypes.c_int]
    L.hm_plan_blocks.restype = ctypes.c_int
    L.hm_plan_blocks.argtypes = [_vp, _ip, _dp, _i64, _ip]
    for name in ("hm_matvec", "hm_matvec_t"):
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64, ctypes.c_int,
                       ctypes.c_uint, _vp]
    L.hm_matvec_grad_x.restype = ctypes.c_int
    L.hm_matvec_grad_x.argtypes = [_vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64,
                                   ctypes.c_int, _vp]
    L.hm_matvec_grad.restype = ctypes.c_int
    L.hm_matvec_grad.argtypes = [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, ctypes.c_int,
                                 _vp]
    L.hm_matvec_t_grad.restype = ctypes.c_int
    L.hm_matvec_t_grad.argtypes = [_vp, _vp, _vp, _vp, _vp, _i64, _i64, _i64, _i64, ctypes.c_int,
                                   _vp]
    L.hm_graph_matvec.restype = ctypes.c_int
    L.hm_graph_matvec.argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64,
                                  ctypes.c_int, ctypes.c_uint, ctypes.POINTER(_vp)]
    L.hm_graph_launch.restype = ctypes.c_int
    L.hm_graph_launch.argtypes = [_vp, _vp]
    L.hm_graph_destroy.restype = ctypes.c_int
    L.hm_graph_destroy.argtypes = [_vp]
    L.hm_plan_tune.restype = ctypes.c_int
    L.hm_plan_tune.argtypes = [_vp, _vp, _vp, _i64, _vp, _ip]
    L.hm_plan_profile.restype = ctypes.c_int
    L.hm_plan_profile.argtypes = [_vp, ctypes.c_int]
    L.hm_plan_profile_read.restype = ctypes.c_int
    L.hm_plan_profile_read.argtypes = [_vp, _dp, _ip]
    L.hm_entries.restype = ctypes.c_int
    L.hm_entries.argtypes = [_vp, _i64, _vp, _vp, _vp, _vp]
    L.hm_plan_update_points.restype = ctypes.c_int
    L.hm_plan_update_points.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.hm_plan_update_points_f32.restype = ctypes.c_int
    L.hm_plan_update_points_f32.argtypes = [_vp, _vp, _vp, _vp, _vp]
    L.hm_plan_add_col.restype = ctypes.c_int
    L.hm_plan_add_col.argtypes = [_vp, _vp, _i64, _i64, _vp]
    L.hm_plan_scale.restype = ctypes.c_int
    L.hm_plan_scale.argtypes = [_vp, ctypes.c_int, _vp, _i64, _vp]
    # Float32 plans
    L.hm_plan_from_points_f32.restype = ctypes.c_int
    L.hm_plan_from_points_f32.argtypes = [ctypes.c_int, _fp, _i64, _fp, _i64, ctypes.c_float,
                                          ctypes.c_float, ctypes.c_float, ctypes.c_float,
                                          ctypes.c_int, ctypes.c_int, ctypes.c_uint, _vp,
                                          ctypes.POINTER(_vp)]
    for name in ("hm_matvec_f32", "hm_matvec_t_f32"):
        fn = getattr(L, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64, ctypes.c_int,
                       ctypes.c_uint, _vp]
    L.hm_matvec_grad_x_f32.restype = ctypes.c_int
    L.hm_matvec_grad_x_f32.argtypes = L.hm_matvec_grad_x.argtypes
    L.hm_matvec_grad_f32.restype = ctypes.c_int
    L.hm_matvec_grad_f32.argtypes = L.hm_matvec_grad.argtypes
    L.hm_matvec_t_grad_f32.restype = ctypes.c_int
    L.hm_matvec_t_grad_f32.argtypes = L.hm_matvec_t_grad.argtypes
    L.hm_graph_matvec_f32.restype = ctypes.c_int
    L.hm_graph_matvec_f32.argtypes = L.hm_graph_matvec.argtypes
    L.hm_entries_f32.restype = ctypes.c_int
    L.hm_entries_f32.argtypes = L.hm_entries.argtypes
    # EvenBarycentricMatrix leaves
    L.hm_even_nodes.restype = ctypes.c_int
    L.hm_even_nodes.argtypes = [_i64, _i64, ctypes.c_int, _dp]
    L.hm_even_weights.restype = ctypes.c_int
    L.hm_even_weights.argtypes = [_i64, _i64, ctypes.c_int, ctypes.c_int, _vp, _vp, _vp]
    L.hm_even_create.restype = ctypes.c_int
    L.hm_even_create.argtypes = [_i64, _i64, _i64, _ip, _ip, _ip, ctypes.c_int,
                                 ctypes.POINTER(_dp), _vp, ctypes.POINTER(_vp)]
    L.hm_even_matvec.restype = ctypes.c_int
    L.hm_even_matvec.argtypes = [_vp, _vp, _vp, _i64, _i64, _i64, _i64, _i64, _i64, _i64,
                                 ctypes.c_int, _vp]
    L.hm_even_leaf_data.restype = ctypes.c_int
    L.hm_even_leaf_data.argtypes = [_vp, _i64, _dp, _dp]
    L.hm_even_nodes_f32.restype = ctypes.c_int
    L.hm_even_nodes_f32.argtypes = L.hm_even_nodes.argtypes
    L.hm_even_weights_f32.restype = ctypes.c_int
    L.hm_even_weights_f32.argtypes = L.hm_even_weights.argtypes
    L.hm_even_create_f32.restype = ctypes.c_int
    L.hm_even_create_f32.argtypes = [_i64, _i64, _i64, _ip, _ip, _ip, ctypes.c_int,
                                     ctypes.POINTER(_fp), _vp, ctypes.POINTER(_vp)]
    L.hm_even_matvec_f32.restype = ctypes.c_int
    L.hm_even_matvec_f32.argtypes = L.hm_even_matvec.argtypes
    L.hm_even_leaf_data_f32.restype = ctypes.c_int
    L.hm_even_leaf_data_f32.argtypes = [_vp, _i64, _fp, _fp]
    L.hm_chebyshevpoints_f32.restype = ctypes.c_int
    L.hm_chebyshevpoints_f32.argtypes = [_fp, ctypes.c_int, ctypes.c_int]
    L.hm_chebyshevbarycentricweights_f32.restype = ctypes.c_int
    L.hm_chebyshevbarycentricweights_f32.argtypes = [_fp, ctypes.c_int, ctypes.c_int]
    L.hm_even_info.restype = ctypes.c_int
    L.hm_even_info.argtypes = [_vp, _ip, ctypes.c_int]
    L.hm_even_destroy.restype = ctypes.c_int
    L.hm_even_destroy.argtypes = [_vp]
    # LowRankMatrix algebra
    L.hm_lowrank_op.restype = ctypes.c_int
    L.hm_lowrank_op.argtypes = [ctypes.c_int, _i64, _ip, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                _ip, _vp]
    # dense leaf steps of the hierarchical Cholesky factorization
    L.hm_dense_potrf.restype = ctypes.c_int
    L.hm_dense_potrf.argtypes = [_i64, _ip, ctypes.POINTER(_vp), _ip, _ip, _vp]
    L.hm_dense_potrf_dev.restype = ctypes.c_int
    L.hm_dense_potrf_dev.argtypes = [_i64, _ip, ctypes.POINTER(_vp), _ip, _vp, _vp]
    L.hm_dense_trsm_t.restype = ctypes.c_int
    L.hm_dense_trsm_t.argtypes = [_i64, _ip, _ip, ctypes.POINTER(_vp), _ip, ctypes.POINTER(_vp),
                                  _ip, _vp]
    L.hm_dense_gemm.restype = ctypes.c_int
    L.hm_dense_gemm.argtypes = [ctypes.c_int, ctypes.c_int, _i64, _ip, ctypes.c_double,
                                ctypes.POINTER(_vp), _ip, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                _ip, ctypes.c_double, ctypes.POINTER(_vp), _ip, _vp]
    L.hm_dense_gemm_noalloc.restype = ctypes.c_int
    L.hm_dense_gemm_noalloc.argtypes = L.hm_dense_gemm.argtypes
    return L


lib = _load()


def check(code):
    if code != HM_OK:
        raise HMError(code, lib.hm_last_error().decode(errors="replace"))
    return code
