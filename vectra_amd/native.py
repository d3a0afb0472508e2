"""ctypes view of the native libraries: the layout-compatible engine types of
include/vectra_abi.h, the kernel ABI of include/vgpu.h, and helpers that build
node trees the way vectra's R bridge does (src/r_bridge.c:744-1093).

This module is plumbing for tests and benchmarks; the product is the C11
nodes and the HIP kernels.  Loading fails loudly when a library is missing.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path
from typing import Any, Sequence

import numpy as np

PKG = Path(__file__).resolve().parent
LIB = PKG / "lib"
if os.environ.get("VECTRA_AMD_LIB_DIR"):
    LIB = Path(os.environ["VECTRA_AMD_LIB_DIR"])

VEC_INT64, VEC_DOUBLE, VEC_BOOL, VEC_STRING = 0, 1, 2, 3
(EXPR_COL_REF, EXPR_LIT_INT64, EXPR_LIT_DOUBLE, EXPR_LIT_BOOL, EXPR_LIT_STRING, EXPR_LIT_NA, EXPR_ARITH,
 EXPR_CMP, EXPR_BOOL, EXPR_IS_NA, EXPR_IF_ELSE, EXPR_CAST, EXPR_NEGATE) = range(13)
AGG_COUNT, AGG_COUNT_STAR, AGG_SUM, AGG_MEAN, AGG_MIN, AGG_MAX = range(6)
JOIN_INNER, JOIN_LEFT, JOIN_FULL, JOIN_SEMI, JOIN_ANTI = range(5)
AGG_NAMES = {"count": AGG_COUNT, "n": AGG_COUNT_STAR, "count_star": AGG_COUNT_STAR, "sum": AGG_SUM,
             "mean": AGG_MEAN, "min": AGG_MIN, "max": AGG_MAX}
JOIN_NAMES = {"inner": JOIN_INNER, "left": JOIN_LEFT, "full": JOIN_FULL, "semi": JOIN_SEMI, "anti": JOIN_ANTI}


# --------------------------------------------------------------------------
# layout-compatible structs (include/vectra_abi.h)
# --------------------------------------------------------------------------

class _Str(C.Structure):
    # data as a plain address: a c_char_p field would be read back as a
    # temporary bytes object cut at the first NUL
    _fields_ = [("offsets", C.POINTER(C.c_int64)), ("data", C.c_void_p), ("data_len", C.c_int64)]


class _Buf(C.Union):
    _fields_ = [("i64", C.POINTER(C.c_int64)), ("dbl", C.POINTER(C.c_double)), ("bln", C.POINTER(C.c_uint8)),
                ("str", _Str)]


class VecArray(C.Structure):
    _fields_ = [("type", C.c_int), ("length", C.c_int64), ("validity", C.POINTER(C.c_uint8)),
                ("owns_data", C.c_uint8), ("buf", _Buf)]


class VecBatch(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("n_cols", C.c_int), ("col_names", C.POINTER(C.c_char_p)),
                ("columns", C.POINTER(VecArray)), ("sel", C.POINTER(C.c_int32)), ("sel_n", C.c_int32)]


class VecSchema(C.Structure):
    _fields_ = [("n_cols", C.c_int), ("col_names", C.POINTER(C.c_char_p)), ("col_types", C.POINTER(C.c_int)),
                ("col_annotations", C.POINTER(C.c_char_p))]


class VecNode(C.Structure):
    pass


VecNode._fields_ = [("next_batch", C.c_void_p), ("free_node", C.c_void_p), ("output_schema", VecSchema),
                    ("kind", C.c_char_p)]


class VecExpr(C.Structure):
    pass


VecExpr._fields_ = [("kind", C.c_int), ("result_type", C.c_int), ("col_name", C.c_void_p), ("lit_i64", C.c_int64),
                    ("lit_dbl", C.c_double), ("lit_bln", C.c_uint8), ("lit_str", C.c_void_p), ("op", C.c_char),
                    ("op2", C.c_char), ("left", C.POINTER(VecExpr)), ("right", C.POINTER(VecExpr)),
                    ("operand", C.POINTER(VecExpr)), ("cond", C.POINTER(VecExpr)),
                    ("then_expr", C.POINTER(VecExpr)), ("else_expr", C.POINTER(VecExpr)), ("cast_to", C.c_int)]


class AggSpec(C.Structure):
    _fields_ = [("output_name", C.c_void_p), ("kind", C.c_int), ("input_col", C.c_void_p), ("na_rm", C.c_int)]


class JoinKey(C.Structure):
    _fields_ = [("left_col", C.c_int), ("right_col", C.c_int)]


class ProjEntry(C.Structure):
    _fields_ = [("output_name", C.c_void_p), ("expr", C.c_void_p)]


# --------------------------------------------------------------------------
# kernel ABI structs (include/vgpu.h)
# --------------------------------------------------------------------------

class VgpuCol(C.Structure):
    _fields_ = [("type", C.c_int32), ("reserved", C.c_int32), ("data", C.c_void_p), ("validity", C.c_void_p)]


class _Imm(C.Union):
    _fields_ = [("i", C.c_int64), ("d", C.c_double)]


class VgpuInsn(C.Structure):
    _fields_ = [("op", C.c_uint8), ("type", C.c_uint8), ("a", C.c_uint8), ("b", C.c_uint8), ("col", C.c_int32),
                ("imm", _Imm)]


VGPU_MAX_INSN = 24


class VgpuProg(C.Structure):
    _fields_ = [("n_insn", C.c_int32), ("reserved", C.c_int32), ("insn", VgpuInsn * VGPU_MAX_INSN)]


class VgpuAgg(C.Structure):
    _fields_ = [("kind", C.c_int32), ("input", C.c_int32), ("na_rm", C.c_int32), ("reserved", C.c_int32)]


class VgpuInput(C.Structure):
    _fields_ = [("op", C.c_int32), ("side_a", C.c_int32), ("col_a", C.c_int32), ("side_b", C.c_int32),
                ("col_b", C.c_int32), ("reserved", C.c_int32)]


class VgpuQuery(C.Structure):
    _fields_ = [("key", VgpuCol), ("probe_key", VgpuCol), ("inputs", C.POINTER(VgpuCol)),
                ("pred_cols", C.POINTER(VgpuCol)), ("n_rows", C.c_int64), ("param", C.c_uint64)]


class VxRequest(C.Structure):
    _fields_ = [("cls", C.c_int32), ("pad", C.c_int32), ("threshold", C.c_double), ("arrival_s", C.c_double)]


class VxResult(C.Structure):
    _fields_ = [("start_s", C.c_double), ("end_s", C.c_double), ("pass_", C.c_int32), ("batch", C.c_int32),
                ("n_groups", C.c_int64), ("keys", C.POINTER(C.c_int64)), ("key_valid", C.POINTER(C.c_uint8)),
                ("sum", C.POINTER(C.c_double)), ("sum_valid", C.POINTER(C.c_uint8)),
                ("count", C.POINTER(C.c_double))]


class VxServeCfg(C.Structure):
    _fields_ = [("fact_key", VgpuCol), ("fact_value", VgpuCol), ("fact_group", VgpuCol), ("n_rows", C.c_int64),
                ("dim_key", VgpuCol), ("dim_attr", VgpuCol), ("n_dim", C.c_int64), ("batched", C.c_int32),
                ("max_batch", C.c_int32), ("state_budget", C.c_int64), ("keep_results", C.c_int32),
                ("in_flight", C.c_int32), ("hbm_budget", C.c_int64)]


class VxServeStats(C.Structure):
    _fields_ = [("n_done", C.c_int64), ("n_passes", C.c_int32), ("max_pass", C.c_int32), ("wall_s", C.c_double),
                ("setup_s", C.c_double), ("state_bytes", C.c_int64), ("resident_bytes", C.c_int64),
                ("max_in_flight", C.c_int32), ("slots", C.c_int32), ("hbm_budget", C.c_int64)]


class VgpuFeedSpec(C.Structure):
    _fields_ = [("n_rows", C.c_int64), ("sel", C.c_void_p), ("n_sel", C.c_int64), ("pos_base", C.c_int64),
                ("pred", C.POINTER(VgpuProg)), ("join", C.c_void_p), ("join_kind", C.c_int32),
                ("key_side", C.c_int32), ("key_col", C.c_int32), ("probe_key_col", C.c_int32),
                ("n_probe_cols", C.c_int32), ("n_build_cols", C.c_int32), ("probe_cols", C.POINTER(VgpuCol)),
                ("build_cols", C.POINTER(VgpuCol)), ("n_inputs", C.c_int32), ("reserved", C.c_int32),
                ("inputs", C.POINTER(VgpuInput))]


VGPU_OP = {name: i + 1 for i, name in enumerate(
    ["COL", "CONST", "NA", "CAST", "ADD", "SUB", "MUL", "DIV", "MOD", "EQ", "NE", "LT", "LE", "GT", "GE",
     "AND", "OR", "NOT", "ISNA", "NEG"])}
VGPU_KEY_NONE = -1
VGPU_ORDER_FIRST_SEEN, VGPU_ORDER_KEY = 0, 1


# --------------------------------------------------------------------------
# loading
# --------------------------------------------------------------------------

_libs: dict[str, C.CDLL] = {}
_libc = C.CDLL(None)
_libc.malloc.restype = C.c_void_p
_libc.malloc.argtypes = [C.c_size_t]
_libc.free.argtypes = [C.c_void_p]
_libc.calloc.restype = C.c_void_p
_libc.calloc.argtypes = [C.c_size_t, C.c_size_t]

NODE_P = C.POINTER(VecNode)
BATCH_P = C.POINTER(VecBatch)
EXPR_P = C.POINTER(VecExpr)


def _need(path: Path) -> Path:
    if not path.exists():
        raise RuntimeError(f"native library {path} is missing: build it first (python -m vectra_amd.build)")
    return path


def _proto(lib: C.CDLL, name: str, res, args) -> None:
    f = getattr(lib, name)
    f.restype = res
    f.argtypes = args


def core() -> C.CDLL:
    if "core" not in _libs:
        lib = C.CDLL(str(_need(LIB / "libvx_core.so")), mode=C.RTLD_GLOBAL)
        _proto(lib, "vx_mem_scan_create", NODE_P, [C.c_int, C.POINTER(C.c_char_p), C.POINTER(VecArray), C.c_int64,
                                                   C.c_int64])
        _proto(lib, "vx_try_collect", C.c_int, [NODE_P, C.POINTER(BATCH_P)])
        _proto(lib, "vx_try_next", C.c_int, [NODE_P, C.POINTER(BATCH_P)])
        _proto(lib, "vx_node_free", None, [NODE_P])
        _proto(lib, "vx_last_error", C.c_char_p, [])
        _proto(lib, "vx_last_warning", C.c_char_p, [])
        _proto(lib, "vx_warning_count", C.c_int64, [])
        _proto(lib, "vx_clear_warnings", None, [])
        _proto(lib, "vec_batch_free", None, [BATCH_P])
        _proto(lib, "vec_expr_alloc", EXPR_P, [C.c_int])
        _proto(lib, "vec_expr_free", None, [EXPR_P])
        _proto(lib, "vx_strdup", C.c_void_p, [C.c_char_p])
        _proto(lib, "vx_try_filter", C.c_int, [C.c_void_p, NODE_P, EXPR_P, C.POINTER(NODE_P)])
        _proto(lib, "vx_try_group", C.c_int, [C.c_void_p, NODE_P, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                              C.c_char_p, C.POINTER(NODE_P)])
        _proto(lib, "vx_try_join", C.c_int, [C.c_void_p, NODE_P, NODE_P, C.c_int, C.c_int, C.c_void_p, C.c_char_p,
                                             C.c_char_p, C.POINTER(NODE_P)])
        _proto(lib, "vx_try_scan", C.c_int, [C.c_char_p, C.POINTER(C.c_int), C.c_int, C.POINTER(NODE_P)])
        _proto(lib, "vx_try_vtr_write", C.c_int, [C.c_char_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(VecArray),
                                                  C.c_int64, C.c_int64])
        _proto(lib, "vx_pushdown_predicates", None, [NODE_P])
        _proto(lib, "vx_try_project", C.c_int, [NODE_P, C.c_int, C.c_void_p, C.POINTER(NODE_P)])
        _libs["core"] = lib
    return _libs["core"]


def vgpu() -> C.CDLL:
    if "vgpu" not in _libs:
        core()
        lib = C.CDLL(str(_need(LIB / "libvgpu.so")), mode=C.RTLD_GLOBAL)
        _proto(lib, "vgpu_init", C.c_int, [C.c_int])
        _proto(lib, "vgpu_last_error", C.c_char_p, [])
        _proto(lib, "vgpu_malloc", C.c_void_p, [C.c_size_t])
        _proto(lib, "vgpu_free", C.c_int, [C.c_void_p])
        _proto(lib, "vgpu_memcpy_h2d", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t])
        _proto(lib, "vgpu_memcpy_d2h", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t])
        _proto(lib, "vgpu_memset", C.c_int, [C.c_void_p, C.c_int, C.c_size_t])
        _proto(lib, "vgpu_synchronize", C.c_int, [])
        _proto(lib, "vgpu_filter", C.c_int, [C.POINTER(VgpuProg), C.POINTER(VgpuCol), C.c_int, C.c_int64, C.c_void_p,
                                             C.c_int64, C.c_void_p, C.POINTER(C.c_int64)])
        _proto(lib, "vgpu_filter_noalloc", C.c_int, [C.POINTER(VgpuProg), C.POINTER(VgpuCol), C.c_int, C.c_int64,
                                                     C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(C.c_int64)])
        _proto(lib, "vgpu_eval_host", C.c_int, [C.POINTER(VgpuProg), C.POINTER(VgpuCol), C.c_int, C.c_int64,
                                                C.c_void_p, C.c_int64, C.c_void_p, C.POINTER(C.c_int64)])
        _proto(lib, "vgpu_groupby_create", C.c_void_p, [C.c_int, C.c_int, C.POINTER(C.c_int32), C.c_int,
                                                        C.POINTER(VgpuAgg), C.c_int, C.c_int64])
        _proto(lib, "vgpu_groupby_feed", C.c_int, [C.c_void_p, C.POINTER(VgpuCol), C.POINTER(VgpuCol), C.c_int64,
                                                   C.c_void_p, C.c_int64, C.c_int64, C.POINTER(VgpuProg),
                                                   C.POINTER(VgpuCol), C.c_int])
        _proto(lib, "vgpu_groupby_feed_join", C.c_int, [C.c_void_p, C.POINTER(VgpuFeedSpec)])
        _proto(lib, "vgpu_groupby_finish", C.c_int64, [C.c_void_p])
        _proto(lib, "vgpu_groupby_fetch", C.c_int64, [C.c_void_p, C.c_void_p, C.c_void_p,
                                                      C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)])
        _proto(lib, "vgpu_groupby_destroy", None, [C.c_void_p])
        _proto(lib, "vgpu_groupby_record_words", C.c_int, [C.c_void_p])
        _proto(lib, "vgpu_groupby_last_path", C.c_int, [C.c_void_p])
        _proto(lib, "vgpu_join_build", C.c_void_p, [C.POINTER(VgpuCol), C.c_int64])
        _proto(lib, "vgpu_join_build_payload", C.c_void_p, [C.POINTER(VgpuCol), C.c_int64, C.POINTER(VgpuCol),
                                                            C.c_int])
        _proto(lib, "vgpu_join_probe", C.c_int, [C.c_void_p, C.c_int, C.POINTER(VgpuCol), C.c_int64, C.c_void_p,
                                                 C.c_int64, C.POINTER(VgpuCol), C.POINTER(C.c_int64)])
        _proto(lib, "vgpu_join_pairs", C.c_int, [C.c_void_p, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)])
        _proto(lib, "vgpu_join_unmatched_build", C.c_int, [C.c_void_p, C.POINTER(C.c_int64)])
        _proto(lib, "vgpu_join_build_rows", C.c_int64, [C.c_void_p])
        _proto(lib, "vgpu_join_max_dup", C.c_int64, [C.c_void_p])
        _proto(lib, "vgpu_join_destroy", None, [C.c_void_p])
        _proto(lib, "vgpu_gather", C.c_int, [C.POINTER(VgpuCol), C.c_void_p, C.c_int, C.c_int64, C.c_void_p,
                                             C.c_void_p])
        _proto(lib, "vgpu_cast", C.c_int, [C.POINTER(VgpuCol), C.c_int64, C.c_int, C.c_void_p])
        _proto(lib, "vgpu_sort_pairs_u64", C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_int, C.c_int])
        _proto(lib, "vgpu_pairs_unequal", C.c_int, [C.POINTER(VgpuCol), C.POINTER(VgpuCol), C.c_int, C.c_void_p,
                                                    C.c_void_p, C.c_int64, C.POINTER(C.c_int64)])
        _proto(lib, "vgpu_timer_start", C.c_int, [])
        _proto(lib, "vgpu_timer_stop", C.c_double, [])
        _proto(lib, "vgpu_tune_set", C.c_int, [C.c_char_p, C.c_int64])
        _proto(lib, "vgpu_tune_get", C.c_int64, [C.c_char_p])
        _proto(lib, "vgpu_tune_set_for", C.c_int, [C.c_int64, C.c_char_p, C.c_int64])
        _proto(lib, "vgpu_tune_load", C.c_int, [C.c_char_p])
        _proto(lib, "vgpu_tune_reset", None, [])
        _proto(lib, "vgpu_tune_get_for", C.c_int64, [C.c_int64, C.c_char_p])
        _proto(lib, "vgpu_tune_set_at", C.c_int, [C.c_int64, C.c_int64, C.c_int64, C.c_char_p, C.c_int64])
        _proto(lib, "vgpu_tune_get_at", C.c_int64, [C.c_int64, C.c_int64, C.c_int64, C.c_char_p])
        _proto(lib, "vgpu_kernel_timing", C.c_int, [C.c_int])
        _proto(lib, "vgpu_kernel_time_ms", C.c_double, [C.POINTER(C.c_int64), C.c_int])
        _proto(lib, "vgpu_batch_create", C.c_void_p, [C.c_int, C.c_int, C.POINTER(C.c_int32), C.c_int,
                                                      C.POINTER(VgpuAgg), C.c_int, C.POINTER(VgpuProg), C.c_int,
                                                      C.c_int])
        _proto(lib, "vgpu_batch_plan", C.c_int, [C.c_void_p, C.POINTER(VgpuQuery), C.c_int])
        _proto(lib, "vgpu_batch_run", C.c_int, [C.c_void_p])
        _proto(lib, "vgpu_batch_join", C.c_int, [C.c_void_p, C.c_void_p, C.c_int])
        _proto(lib, "vgpu_batch_rerun", C.c_int, [C.c_void_p, C.c_void_p, C.c_int])
        _proto(lib, "vgpu_batch_capture", C.c_int, [C.c_void_p])
        _proto(lib, "vgpu_batch_counts", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p])
        _proto(lib, "vgpu_batch_fetch", C.c_int64, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                                    C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)])
        _proto(lib, "vgpu_batch_destroy", None, [C.c_void_p])
        _proto(lib, "vgpu_batch_device_bytes", C.c_int64, [C.c_void_p])
        _proto(lib, "vgpu_batch_bucketed", C.c_int, [C.c_void_p])
        _proto(lib, "vgpu_strdict_create", C.c_void_p, [])
        _proto(lib, "vgpu_strdict_free", None, [C.c_void_p])
        _proto(lib, "vgpu_strdict_append", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                                     C.c_int64, C.c_void_p, C.c_void_p])
        _proto(lib, "vgpu_strdict_lookup", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                                     C.c_void_p])
        _proto(lib, "vgpu_strdict_size", C.c_int64, [C.c_void_p, C.POINTER(C.c_int64)])
        _libs["vgpu"] = lib
        # the autotuner's table for this build, unless the environment names one
        table = PKG / "tuning.txt"
        if table.exists() and not os.environ.get("VECTRA_GPU_TUNING"):
            if lib.vgpu_tune_load(str(table).encode()) < 0:
                raise VectraError(lib.vgpu_last_error().decode())
    return _libs["vgpu"]


def _node_protos(lib: C.CDLL, prefix: str) -> None:
    _proto(lib, f"{prefix}filter_node_create", NODE_P, [NODE_P, EXPR_P])
    _proto(lib, f"{prefix}group_agg_node_create", NODE_P, [NODE_P, C.c_int, C.c_void_p, C.c_int, C.c_void_p,
                                                           C.c_char_p])
    _proto(lib, f"{prefix}join_node_create", NODE_P, [NODE_P, NODE_P, C.c_int, C.c_int, C.c_void_p, C.c_char_p,
                                                      C.c_char_p])


class VxStrCmp(C.Structure):
    _fields_ = [("col", C.c_int * 2), ("lit", C.c_char_p * 2), ("na", C.c_int * 2)]


class CompiledExpr(C.Structure):
    _fields_ = [("prog", VgpuProg), ("n_cols", C.c_int), ("col_idx", C.c_int * 8), ("result_type", C.c_int),
                ("n_scmp", C.c_int), ("scmp", VxStrCmp * 4)]


def nodes() -> C.CDLL:
    if "nodes" not in _libs:
        vgpu()
        lib = C.CDLL(str(_need(LIB / "libvectra_gpu.so")))
        _node_protos(lib, "")
        _proto(lib, "vx_try_compile_expr", C.c_int, [EXPR_P, BATCH_P, C.POINTER(CompiledExpr)])
        _proto(lib, "vx_string_cmp_host", None, [C.POINTER(CompiledExpr), C.c_int, BATCH_P, C.c_void_p, C.c_void_p])
        _proto(lib, "vx_serve_run", C.c_int, [C.POINTER(VxServeCfg), C.POINTER(VxRequest), C.c_int,
                                              C.POINTER(VxResult), C.POINTER(VxServeStats)])
        _proto(lib, "vx_result_free", None, [C.POINTER(VxResult)])
        _proto(lib, "vx_explain", C.c_int64, [NODE_P, C.c_char_p, C.c_int64])
        _libs["nodes"] = lib
    return _libs["nodes"]


def explain(root: Any) -> list[str]:
    """The plan lines of a node tree (vx_explain; src/r_bridge.c:575-698)."""
    lib = nodes()
    n = lib.vx_explain(root, None, 0)
    if n < 0:
        raise VectraError("explain: no node")
    buf = C.create_string_buffer(int(n) + 1)
    lib.vx_explain(root, buf, int(n) + 1)
    return buf.value.decode().splitlines()


def host_eval_batches(table: "Table", pred: tuple) -> list[np.ndarray]:
    """Compiles pred per batch exactly as the GPU FilterNode does and runs the
    program on the host interpreter (vgpu_eval_host): a CPU check of the
    compiler and the evaluator.  Returns the non-empty selection vectors."""
    lib, vg = nodes(), vgpu()
    scan = table.scan()
    out = []
    while True:
        b = BATCH_P()
        if core().vx_try_next(scan, C.byref(b)) != 0:
            raise VectraError(core().vx_last_error().decode())
        if not b:
            break
        e = expr(pred)   # freed after the evaluation: string comparisons point at its literals
        cx = CompiledExpr()
        rc = lib.vx_try_compile_expr(e, b, C.byref(cx))
        if rc != 0 or cx.result_type != VEC_BOOL:
            core().vec_expr_free(e)
            core().vec_batch_free(b)
            raise VectraError(core().vx_last_error().decode() if rc != 0 else
                              "filter predicate must evaluate to boolean")
        bt = b.contents
        cols = (VgpuCol * 8)()
        keep = []
        for s in range(cx.n_cols):
            if cx.col_idx[s] < 0:   # a string comparison's -1 / 0 / 1 column, computed on the host
                vals = np.zeros(max(bt.n_rows, 1), np.int64)
                bm = np.zeros(max((bt.n_rows + 7) // 8, 1), np.uint8)
                lib.vx_string_cmp_host(C.byref(cx), -1 - cx.col_idx[s], b, vals.ctypes.data, bm.ctypes.data)
                keep += [vals, bm]
                cols[s].type = VEC_INT64
                cols[s].data = vals.ctypes.data
                cols[s].validity = bm.ctypes.data
                continue
            a = bt.columns[cx.col_idx[s]]
            cols[s].type = a.type
            cols[s].data = C.cast(a.buf.i64, C.c_void_p)
            cols[s].validity = C.cast(a.validity, C.c_void_p) if a.validity else None
        n = bt.sel_n if bt.sel else bt.n_rows
        sel = np.zeros(max(n, 1), np.int32)
        k = C.c_int64()
        if n > 0:
            _vcheck(vg.vgpu_eval_host(C.byref(cx.prog), cols, cx.n_cols, bt.n_rows,
                                      C.cast(bt.sel, C.c_void_p) if bt.sel else None, bt.sel_n, sel.ctypes.data,
                                      C.byref(k)))
        if k.value:
            out.append(sel[:k.value].copy())
        core().vec_expr_free(e)
        core().vec_batch_free(b)
    core().vx_node_free(scan)
    return out


def load_oracle(path: Path) -> C.CDLL:
    """Loads the oracle library (test infrastructure) sharing libvx_core."""
    if "oracle" not in _libs:
        core()
        lib = C.CDLL(str(_need(path)))
        _node_protos(lib, "oracle_")
        _proto(lib, "oracle_exact_sum", C.c_double, [C.POINTER(C.c_double), C.c_int64])
        _proto(lib, "oracle_hash_value", C.c_uint64, [C.POINTER(VecArray), C.c_int64])
        _proto(lib, "oracle_project_mul_node_create", NODE_P, [NODE_P, C.c_char_p, C.c_char_p, C.c_char_p])
        _proto(lib, "oracle_try_expr_eval", C.c_int, [EXPR_P, BATCH_P, C.POINTER(C.POINTER(VecArray))])
        _proto(lib, "oracle_array_free", None, [C.POINTER(VecArray)])
        _libs["oracle"] = lib
    return _libs["oracle"]


class VectraError(RuntimeError):
    pass


# --------------------------------------------------------------------------
# host columns and tables
# --------------------------------------------------------------------------

class Column:
    """A host column: numpy values (+ optional bool mask of valid rows) kept
    alive for as long as a VecArray borrows them."""

    def __init__(self, values: np.ndarray, valid: np.ndarray | None = None):
        values = np.asarray(values)
        if values.dtype.kind in "OU":   # strings: one offsets + data array (utf-8)
            enc = [str(v).encode() for v in values]
            self.type = VEC_STRING
            self.values = np.zeros(len(enc) + 1, np.int64)
            self.values[1:] = np.cumsum([len(e) for e in enc]) if enc else []
            self.data = C.create_string_buffer(b"".join(enc), max(1, int(self.values[-1])))
        elif values.dtype == np.bool_:
            self.type = VEC_BOOL
            self.values = np.ascontiguousarray(values.astype(np.uint8))
        elif np.issubdtype(values.dtype, np.integer):
            self.type = VEC_INT64
            self.values = np.ascontiguousarray(values.astype(np.int64))
        elif np.issubdtype(values.dtype, np.floating):
            self.type = VEC_DOUBLE
            self.values = np.ascontiguousarray(values.astype(np.float64))
        else:
            raise TypeError(f"unsupported column dtype {values.dtype}")
        n = len(self.values) - (1 if self.type == VEC_STRING else 0)
        if valid is None:
            self.bitmap = None
        else:
            valid = np.asarray(valid, dtype=bool)
            self.bitmap = np.ascontiguousarray(np.packbits(valid, bitorder="little"))
            if len(self.bitmap) == 0:
                self.bitmap = np.zeros(1, np.uint8)
        self.n = n

    def as_vecarray(self) -> VecArray:
        a = VecArray()
        a.type = self.type
        a.length = self.n
        a.owns_data = 1
        a.validity = self.bitmap.ctypes.data_as(C.POINTER(C.c_uint8)) if self.bitmap is not None else None
        if self.type == VEC_STRING:
            a.buf.str.offsets = self.values.ctypes.data_as(C.POINTER(C.c_int64))
            a.buf.str.data = C.addressof(self.data)
            a.buf.str.data_len = int(self.values[-1])
            a.owns_data = 0
        elif self.type == VEC_BOOL:
            a.buf.bln = self.values.ctypes.data_as(C.POINTER(C.c_uint8))
        elif self.type == VEC_INT64:
            a.buf.i64 = self.values.ctypes.data_as(C.POINTER(C.c_int64))
        else:
            a.buf.dbl = self.values.ctypes.data_as(C.POINTER(C.c_double))
        return a


class Table:
    """Named host columns, emitted in row groups of rg_rows like a .vtr scan."""

    def __init__(self, cols: dict[str, Any], rg_rows: int = 0):
        self.cols = {k: (v if isinstance(v, Column) else Column(*v) if isinstance(v, tuple) else Column(v))
                     for k, v in cols.items()}
        lens = {c.n for c in self.cols.values()}
        if len(lens) > 1:
            raise ValueError(f"ragged table: {lens}")
        self.n = lens.pop() if lens else 0
        self.rg_rows = rg_rows
        self._keep: list[Any] = []

    def scan(self) -> Any:
        names = list(self.cols)
        arr = (VecArray * max(1, len(names)))()
        for i, k in enumerate(names):
            arr[i] = self.cols[k].as_vecarray()
        cnames = (C.c_char_p * max(1, len(names)))(*[k.encode() for k in names])
        self._keep.append((arr, cnames))
        node = core().vx_mem_scan_create(len(names), cnames, arr, self.n, self.rg_rows)
        node._keep = self   # the scan borrows these buffers for its lifetime
        return node


def vtr_write(path: str | os.PathLike, cols: dict[str, Any], rg_rows: int = 0) -> None:
    """Writes named host columns as a version-3 .vtr file of row groups of
    rg_rows rows (src/vtr1.c:39-135 layout, statistics per row group)."""
    t = Table(cols)
    names = list(t.cols)
    arr = (VecArray * max(1, len(names)))()
    for i, k in enumerate(names):
        arr[i] = t.cols[k].as_vecarray()
    cnames = (C.c_char_p * max(1, len(names)))(*[k.encode() for k in names])
    if core().vx_try_vtr_write(str(path).encode(), len(names), cnames, arr, t.n, rg_rows) != 0:
        raise VectraError(core().vx_last_error().decode())


def scan_vtr(path: str | os.PathLike, col_indices: Sequence[int] | None = None) -> Any:
    """scan_node_create over a .vtr file: all columns, or the given file
    column indices (output in file order)."""
    out = NODE_P()
    if col_indices is None:
        rc = core().vx_try_scan(str(path).encode(), None, 0, C.byref(out))
    else:
        idx = (C.c_int * max(1, len(col_indices)))(*col_indices)
        rc = core().vx_try_scan(str(path).encode(), idx, len(col_indices), C.byref(out))
    if rc != 0:
        raise VectraError(core().vx_last_error().decode())
    return out


def pushdown(node: Any) -> Any:
    """The optimizer's predicate pushdown onto a .vtr scan below a filter."""
    core().vx_pushdown_predicates(node)
    return node


# --------------------------------------------------------------------------
# expressions (the R bridge's parse_expr, src/r_bridge.c:744-876)
# --------------------------------------------------------------------------

def _malloc_str(s: str) -> int:
    return core().vx_strdup(s.encode())


def expr(tree: Any) -> Any:
    """Builds a malloc'd VecExpr from a nested tuple:
    ("col", name) ("dbl", v) ("int", v) ("lgl", v) ("str", s) ("na",) ("na_str",) ("cmp", op, l, r)
    ("arith", op, l, r) ("and", l, r) ("or", l, r) ("not", x) ("is_na", x) ("neg", x)."""
    lib = core()
    kind = tree[0]
    if kind == "col":
        e = lib.vec_expr_alloc(EXPR_COL_REF)
        e.contents.col_name = _malloc_str(tree[1])
        return e
    if kind == "dbl":
        e = lib.vec_expr_alloc(EXPR_LIT_DOUBLE)
        e.contents.lit_dbl = float(tree[1])
        e.contents.result_type = VEC_DOUBLE
        return e
    if kind == "int":
        e = lib.vec_expr_alloc(EXPR_LIT_INT64)
        e.contents.lit_i64 = int(tree[1])
        e.contents.result_type = VEC_INT64
        return e
    if kind == "lgl":
        e = lib.vec_expr_alloc(EXPR_LIT_BOOL)
        e.contents.lit_bln = 1 if tree[1] else 0
        e.contents.result_type = VEC_BOOL
        return e
    if kind == "na":
        e = lib.vec_expr_alloc(EXPR_LIT_NA)
        e.contents.result_type = VEC_DOUBLE
        return e
    if kind == "str":
        e = lib.vec_expr_alloc(EXPR_LIT_STRING)
        e.contents.lit_str = _malloc_str(tree[1])
        e.contents.result_type = VEC_STRING
        return e
    if kind == "na_str":
        e = lib.vec_expr_alloc(EXPR_LIT_NA)
        e.contents.result_type = VEC_STRING
        return e
    if kind in ("cmp", "arith"):
        e = lib.vec_expr_alloc(EXPR_CMP if kind == "cmp" else EXPR_ARITH)
        op = tree[1]
        e.contents.op = op[0].encode()
        e.contents.op2 = (op[1] if len(op) > 1 else " ").encode()
        e.contents.left = expr(tree[2])
        e.contents.right = expr(tree[3])
        e.contents.result_type = VEC_BOOL if kind == "cmp" else VEC_DOUBLE
        return e
    if kind in ("and", "or"):
        e = lib.vec_expr_alloc(EXPR_BOOL)
        e.contents.op = b"&" if kind == "and" else b"|"
        e.contents.left = expr(tree[1])
        e.contents.right = expr(tree[2])
        e.contents.result_type = VEC_BOOL
        return e
    if kind == "not":
        e = lib.vec_expr_alloc(EXPR_BOOL)
        e.contents.op = b"!"
        e.contents.operand = expr(tree[1])
        e.contents.result_type = VEC_BOOL
        return e
    if kind == "is_na":
        e = lib.vec_expr_alloc(EXPR_IS_NA)
        e.contents.operand = expr(tree[1])
        e.contents.result_type = VEC_BOOL
        return e
    if kind == "neg":
        e = lib.vec_expr_alloc(EXPR_NEGATE)
        e.contents.operand = expr(tree[1])
        return e
    raise ValueError(f"unknown expression node {kind}")


# --------------------------------------------------------------------------
# node construction and execution
# --------------------------------------------------------------------------

def _lib_for(engine: str, oracle_path: Path | None = None) -> tuple[C.CDLL, str]:
    if engine == "gpu":
        return nodes(), ""
    if engine == "oracle":
        if "oracle" not in _libs:
            if oracle_path is None:
                raise RuntimeError("oracle library not loaded")
            load_oracle(oracle_path)
        return _libs["oracle"], "oracle_"
    raise ValueError(engine)


def _ctor(lib: C.CDLL, name: str) -> C.c_void_p:
    return C.cast(getattr(lib, name), C.c_void_p)


def _check(rc: int) -> None:
    if rc != 0:
        raise VectraError(core().vx_last_error().decode())


def filter_node(engine: str, child: Any, pred: Any) -> Any:
    lib, p = _lib_for(engine)
    out = NODE_P()
    e = expr(pred) if isinstance(pred, tuple) else pred
    _check(core().vx_try_filter(_ctor(lib, f"{p}filter_node_create"), child, e, C.byref(out)))
    out._keep = child
    return out


def group_agg_node(engine: str, child: Any, keys: Sequence[str], aggs: Sequence[tuple], temp_dir: str | None = "/tmp"
                   ) -> Any:
    """aggs: (output_name, kind, input_col or None, na_rm)."""
    lib, p = _lib_for(engine)
    nk = len(keys)
    key_names = _libc.malloc(max(1, nk) * C.sizeof(C.c_void_p))
    kn = C.cast(key_names, C.POINTER(C.c_void_p))
    for i, k in enumerate(keys):
        kn[i] = _malloc_str(k)
    na = len(aggs)
    specs = _libc.calloc(max(1, na), C.sizeof(AggSpec))
    sp = C.cast(specs, C.POINTER(AggSpec))
    for i, (name, kind, col, na_rm) in enumerate(aggs):
        sp[i].output_name = _malloc_str(name)
        sp[i].kind = AGG_NAMES[kind] if isinstance(kind, str) else kind
        sp[i].input_col = _malloc_str(col) if col is not None else None
        sp[i].na_rm = int(bool(na_rm))
    td = temp_dir.encode() if temp_dir is not None else None
    out = NODE_P()
    _check(core().vx_try_group(_ctor(lib, f"{p}group_agg_node_create"), child, nk, key_names, na, specs, td,
                               C.byref(out)))
    out._keep = child
    return out


def join_node(engine: str, left: Any, right: Any, kind: str, keys: Sequence[tuple[int, int]],
              suffix_x: str = ".x", suffix_y: str = ".y") -> Any:
    lib, p = _lib_for(engine)
    jk = _libc.malloc(max(1, len(keys)) * C.sizeof(JoinKey))
    kp = C.cast(jk, C.POINTER(JoinKey))
    for i, (lc, rc) in enumerate(keys):
        kp[i].left_col = lc
        kp[i].right_col = rc
    out = NODE_P()
    _check(core().vx_try_join(_ctor(lib, f"{p}join_node_create"), left, right, JOIN_NAMES[kind], len(keys), jk,
                              suffix_x.encode(), suffix_y.encode(), C.byref(out)))
    out._keep = (left, right)
    return out


def project_node(child: Any, entries: Sequence[tuple[str, Any]]) -> Any:
    """The reference's ProjectNode (src/project.h, the harness's stand-in):
    entries are (output_name, None) to pass a child column through or
    (output_name, expression tuple) for a computed column -- the hidden
    mutate summarise() inserts is every child column passed through plus
    (".vectra_tmp_<name>", ("arith", "*", ("col", a), ("col", b)))."""
    n = len(entries)
    arr = _libc.calloc(max(1, n), C.sizeof(ProjEntry))
    ep = C.cast(arr, C.POINTER(ProjEntry))
    for i, (name, tree) in enumerate(entries):
        ep[i].output_name = _malloc_str(name)
        ep[i].expr = C.cast(expr(tree), C.c_void_p).value if tree is not None else None
    out = NODE_P()
    _check(core().vx_try_project(child, n, arr, C.byref(out)))
    out._keep = child
    return out


def hidden_mutate(child: Any, tmp_name: str, a: str, b: str) -> Any:
    """Project(child columns..., tmp_name = a * b), as summarise(sum(a * b))
    builds it (R/verbs.R:266-284)."""
    names = [child.contents.output_schema.col_names[i].decode()
             for i in range(child.contents.output_schema.n_cols)]
    return project_node(child, [(nm, None) for nm in names] + [(tmp_name, ("arith", "*", ("col", a), ("col", b)))])


def batch_to_numpy(b: Any) -> dict[str, tuple[np.ndarray, np.ndarray]]:
    """Dense batch -> {name: (values, valid)}; strings come back as object arrays."""
    out = {}
    bt = b.contents
    n = bt.n_rows
    for c in range(bt.n_cols):
        a = bt.columns[c]
        name = bt.col_names[c].decode()
        if a.validity and n > 0:
            bits = np.ctypeslib.as_array(a.validity, shape=((n + 7) // 8,))
            valid = np.unpackbits(bits, bitorder="little")[:n].astype(bool)
        else:
            valid = np.ones(n, dtype=bool)
        if n == 0:
            vals = np.zeros(0, dtype={VEC_INT64: np.int64, VEC_DOUBLE: np.float64, VEC_BOOL: np.uint8}.get(a.type, object))
        elif a.type == VEC_INT64:
            vals = np.ctypeslib.as_array(a.buf.i64, shape=(n,)).copy()
        elif a.type == VEC_DOUBLE:
            vals = np.ctypeslib.as_array(a.buf.dbl, shape=(n,)).copy()
        elif a.type == VEC_BOOL:
            vals = np.ctypeslib.as_array(a.buf.bln, shape=(n,)).copy()
        else:
            offs = np.ctypeslib.as_array(a.buf.str.offsets, shape=(n + 1,))
            data = C.string_at(a.buf.str.data, int(offs[-1])) if offs[-1] > 0 else b""
            vals = np.array([data[offs[i]:offs[i + 1]].decode() for i in range(n)], dtype=object)
        vals = np.where(valid, vals, 0) if a.type != VEC_STRING else vals
        out[name] = (vals, valid)
    return out


def collect(root: Any, free: bool = True) -> dict[str, tuple[np.ndarray, np.ndarray]]:
    """Pull loop of vec_collect (src/collect.c:188-206); raises VectraError."""
    lib = core()
    out = BATCH_P()
    rc = lib.vx_try_collect(root, C.byref(out))
    if rc != 0:
        msg = lib.vx_last_error().decode()
        if free:
            lib.vx_node_free(root)
        raise VectraError(msg)
    try:
        return batch_to_numpy(out)
    finally:
        lib.vec_batch_free(out)
        if free:
            lib.vx_node_free(root)


def next_batches(root: Any, free: bool = True) -> list[dict[str, Any]]:
    """Pulls every batch separately, keeping sel: [{name: (values, valid)}, 'sel': array|None, 'n_rows': n]."""
    lib = core()
    res = []
    while True:
        out = BATCH_P()
        rc = lib.vx_try_next(root, C.byref(out))
        if rc != 0:
            msg = lib.vx_last_error().decode()
            if free:
                lib.vx_node_free(root)
            raise VectraError(msg)
        if not out:
            break
        bt = out.contents
        sel = None
        if bt.sel:
            sel = np.ctypeslib.as_array(bt.sel, shape=(max(bt.sel_n, 1),))[:bt.sel_n].copy()
        sel_ptr = bt.sel
        bt.sel = None
        d = batch_to_numpy(out)
        bt.sel = sel_ptr
        res.append({"cols": d, "sel": sel, "n_rows": bt.n_rows})
        lib.vec_batch_free(out)
    if free:
        lib.vx_node_free(root)
    return res


# --------------------------------------------------------------------------
# device-resident columns and the kernel-level API (include/vgpu.h)
# --------------------------------------------------------------------------

def _vcheck(rc: int) -> None:
    if rc < 0:
        raise VectraError(vgpu().vgpu_last_error().decode())


class DeviceColumn:
    """A column copied once into HBM (values + optional validity bitmap)."""

    def __init__(self, values: np.ndarray, valid: np.ndarray | None = None):
        host = Column(values, valid)
        lib = vgpu()
        nbytes = host.values.nbytes
        self.n = host.n
        self.type = host.type
        self.data = lib.vgpu_malloc(max(nbytes, 8))
        if not self.data:
            raise VectraError(lib.vgpu_last_error().decode())
        _vcheck(lib.vgpu_memcpy_h2d(self.data, host.values.ctypes.data, nbytes))
        self.valid = None
        if host.bitmap is not None:
            self.valid = lib.vgpu_malloc(max(host.bitmap.nbytes, 8))
            _vcheck(lib.vgpu_memcpy_h2d(self.valid, host.bitmap.ctypes.data, host.bitmap.nbytes))

    @classmethod
    def wrap(cls, data_ptr: int, n: int, type_: int, valid_ptr: int | None = None, keep: Any = None) -> "DeviceColumn":
        """A view of device memory the caller owns (e.g. a torch tensor's
        storage, kept alive through `keep`); free() leaves it alone."""
        self = cls.__new__(cls)
        self.n, self.type, self.data, self.valid = n, type_, data_ptr, valid_ptr
        self._keep, self._view = keep, True
        return self

    @property
    def col(self) -> VgpuCol:
        c = VgpuCol()
        c.type = self.type
        c.data = self.data
        c.validity = self.valid
        return c

    def free(self) -> None:
        lib = vgpu()
        if getattr(self, "_view", False):
            self.data = self.valid = self._keep = None
            return
        if self.data:
            lib.vgpu_free(self.data)
        if self.valid:
            lib.vgpu_free(self.valid)
        self.data = self.valid = None


class StrDict:
    """The exact device string dictionary (include/vgpu.h, vgpu_strdict_*):
    host string arrays (None = NA) in, int64 codes out."""

    def __init__(self):
        lib = vgpu()
        self.h = lib.vgpu_strdict_create()
        if not self.h:
            raise VectraError(lib.vgpu_last_error().decode())

    @staticmethod
    def _pack(strings: Sequence[str | None]):
        raw = [b"" if x is None else x.encode() for x in strings]
        off = np.zeros(len(raw) + 1, np.int64)
        off[1:] = np.cumsum([len(b) for b in raw])
        chars = np.frombuffer(b"".join(raw) or b"\0", np.uint8).copy()
        valid = np.array([x is not None for x in strings], bool)
        return off, chars, valid

    def _call(self, run, strings: Sequence[str | None]) -> np.ndarray:
        lib = vgpu()
        off, chars, valid = self._pack(strings)
        n = len(strings)
        dcodes = lib.vgpu_malloc(max(n, 1) * 8)
        dval = None
        if not valid.all():
            bm = np.packbits(valid, bitorder="little")
            dval = lib.vgpu_malloc(max(bm.nbytes, 8))
            _vcheck(lib.vgpu_memcpy_h2d(dval, bm.ctypes.data, bm.nbytes))
        try:
            _vcheck(run(off.ctypes.data, chars.ctypes.data, dval, n, dcodes))
            codes = np.empty(n, np.int64)
            _vcheck(lib.vgpu_memcpy_d2h(codes.ctypes.data, dcodes, n * 8))
        finally:
            lib.vgpu_free(dcodes)
            if dval:
                lib.vgpu_free(dval)
        return codes

    def append(self, strings: Sequence[str | None]) -> np.ndarray:
        lib = vgpu()
        return self._call(lambda o, c, v, n, d: lib.vgpu_strdict_append(self.h, o, c, v, n, 0, d, None), strings)

    def lookup(self, strings: Sequence[str | None]) -> np.ndarray:
        lib = vgpu()
        return self._call(lambda o, c, v, n, d: lib.vgpu_strdict_lookup(self.h, o, c, v, n, d), strings)

    def size(self) -> tuple[int, int]:
        rows = C.c_int64()
        d = vgpu().vgpu_strdict_size(self.h, C.byref(rows))
        return int(d), int(rows.value)

    def free(self) -> None:
        if self.h:
            vgpu().vgpu_strdict_free(self.h)
            self.h = None


def col_array(cols: Sequence[DeviceColumn | None]) -> Any:
    arr = (VgpuCol * max(1, len(cols)))()
    for i, c in enumerate(cols):
        if c is not None:
            arr[i] = c.col
    return arr


def compile_prog(insns: Sequence[tuple]) -> VgpuProg:
    """(op, type, a, b, col, imm) tuples -> VgpuProg (op by name)."""
    p = VgpuProg()
    p.n_insn = len(insns)
    for i, (op, typ, a, b, col, imm) in enumerate(insns):
        ins = p.insn[i]
        ins.op = VGPU_OP[op]
        ins.type = typ
        ins.a, ins.b, ins.col = a, b, col
        if isinstance(imm, float):
            ins.imm.d = imm
        else:
            ins.imm.i = int(imm)
    return p


def cmp_const_prog(col_slot: int, col_type: int, op: str, value: float | int) -> VgpuProg:
    """col <op> constant, compared in the column's type (a double constant
    against an int64 column widens the column, as vec_cmp does)."""
    ctype = VEC_DOUBLE if (col_type == VEC_DOUBLE or isinstance(value, float)) else VEC_INT64
    ins = [("COL", col_type, 0, 0, col_slot, 0)]
    if ctype != col_type:
        ins.append(("CAST", ctype, 0, 0, 0, 0))
    a = len(ins) - 1
    ins.append(("CONST", ctype, 0, 0, 0, float(value) if ctype == VEC_DOUBLE else int(value)))
    ins.append((op, ctype, a, len(ins) - 1, 0, 0))
    return compile_prog(ins)


class JoinTable:
    def __init__(self, key: DeviceColumn, payload: Sequence[DeviceColumn] = ()):
        lib = vgpu()
        if payload:
            self._pay = col_array(payload)
            self.h = lib.vgpu_join_build_payload(C.byref(key.col), key.n, self._pay, len(payload))
        else:
            self.h = lib.vgpu_join_build(C.byref(key.col), key.n)
        if not self.h:
            raise VectraError(lib.vgpu_last_error().decode())
        self.key_type = key.type

    def free(self) -> None:
        if self.h:
            vgpu().vgpu_join_destroy(self.h)
            self.h = None


class GroupBy:
    """Kernel-level group-by: create / feed / feed_join / finish / fetch."""

    def __init__(self, key_type: int, input_types: Sequence[int], aggs: Sequence[tuple], order: int,
                 expected_groups: int = 0):
        lib = vgpu()
        self.n_aggs = len(aggs)
        self.key_type = key_type
        it = (C.c_int32 * max(1, len(input_types)))(*input_types)
        ag = (VgpuAgg * max(1, len(aggs)))()
        for i, (kind, inp, na_rm) in enumerate(aggs):
            ag[i].kind = AGG_NAMES[kind] if isinstance(kind, str) else kind
            ag[i].input = inp
            ag[i].na_rm = na_rm
        self.h = lib.vgpu_groupby_create(key_type, len(input_types), it, len(aggs), ag, order, expected_groups)
        if not self.h:
            raise VectraError(lib.vgpu_last_error().decode())

    def feed(self, key: DeviceColumn | None, inputs: Sequence[DeviceColumn], n_rows: int, pos_base: int = 0,
             pred: VgpuProg | None = None, pred_cols: Sequence[DeviceColumn] = ()) -> None:
        ins = col_array(inputs)
        pc = col_array(pred_cols)
        kc = C.byref(key.col) if key is not None else None
        _vcheck(vgpu().vgpu_groupby_feed(self.h, kc, ins, n_rows, None, 0, pos_base,
                                         C.byref(pred) if pred is not None else None, pc, len(pred_cols)))

    def feed_join(self, n_rows: int, probe_cols: Sequence[DeviceColumn], build_cols: Sequence[DeviceColumn],
                  inputs: Sequence[tuple], key: tuple[int, int], join: JoinTable | None = None,
                  join_kind: int = JOIN_INNER, probe_key_col: int = 0, pred: VgpuProg | None = None,
                  pos_base: int = 0) -> None:
        sp = VgpuFeedSpec()
        sp.n_rows = n_rows
        sp.pos_base = pos_base
        sp.pred = C.pointer(pred) if pred is not None else None
        sp.join = join.h if join is not None else None
        sp.join_kind = join_kind
        sp.key_side, sp.key_col = key
        sp.probe_key_col = probe_key_col
        pcs, bcs = col_array(probe_cols), col_array(build_cols)
        sp.n_probe_cols, sp.n_build_cols = len(probe_cols), len(build_cols)
        sp.probe_cols = C.cast(pcs, C.POINTER(VgpuCol))
        sp.build_cols = C.cast(bcs, C.POINTER(VgpuCol))
        inp = (VgpuInput * max(1, len(inputs)))()
        for i, t in enumerate(inputs):
            if len(t) == 2:
                inp[i].op, inp[i].side_a, inp[i].col_a = 0, t[0], t[1]
            else:
                inp[i].op, inp[i].side_a, inp[i].col_a, inp[i].side_b, inp[i].col_b = 1, t[0], t[1], t[2], t[3]
        sp.n_inputs = len(inputs)
        sp.inputs = C.cast(inp, C.POINTER(VgpuInput))
        self._keep = (pcs, bcs, inp, pred)
        _vcheck(vgpu().vgpu_groupby_feed_join(self.h, C.byref(sp)))

    def finish(self) -> int:
        n = vgpu().vgpu_groupby_finish(self.h)
        _vcheck(n)
        self.n_groups = n
        return n

    def last_path(self) -> int:
        """0 hash table, 1 hash partitions, 2 key-range partitions"""
        return vgpu().vgpu_groupby_last_path(self.h)

    def fetch(self) -> dict[str, tuple[np.ndarray, np.ndarray]]:
        n = self.n_groups
        keys = np.zeros(max(n, 1), np.int64)
        kv = np.zeros(max((n + 7) // 8, 1), np.uint8)
        vals = [np.zeros(max(n, 1), np.float64) for _ in range(self.n_aggs)]
        valid = [np.zeros(max((n + 7) // 8, 1), np.uint8) for _ in range(self.n_aggs)]
        vp = (C.c_void_p * max(1, self.n_aggs))(*[v.ctypes.data for v in vals])
        mp = (C.c_void_p * max(1, self.n_aggs))(*[m.ctypes.data for m in valid])
        w = vgpu().vgpu_groupby_fetch(self.h, keys.ctypes.data, kv.ctypes.data, vp, mp)
        _vcheck(w)
        self.warnings = w
        unpack = lambda b: np.unpackbits(b, bitorder="little")[:n].astype(bool)
        out = {}
        if self.key_type == VEC_DOUBLE:
            out["key"] = (keys[:n].view(np.float64), unpack(kv))
        elif self.key_type != VGPU_KEY_NONE:
            out["key"] = (keys[:n], unpack(kv))
        for i in range(self.n_aggs):
            out[f"a{i}"] = (vals[i][:n], unpack(valid[i]))
        return out

    def free(self) -> None:
        if self.h:
            vgpu().vgpu_groupby_destroy(self.h)
            self.h = None


def _param_bits(p) -> int:
    """the 8 bytes of a predicate constant: a float's double bits, an int's two's complement"""
    if isinstance(p, float):
        return int(np.array([p], np.float64).view(np.uint64)[0])
    return int(p) & (2**64 - 1)


class Batch:
    """Kernel-level batch of filter -> group-by queries of one shape
    (include/vgpu.h, vgpu_batch_*)."""

    def __init__(self, key_type: int, input_types: Sequence[int], aggs: Sequence[tuple], order: int,
                 pred: VgpuProg | None = None, n_pred_cols: int = 0, param_insn: int = -1):
        lib = vgpu()
        self.n_aggs = len(aggs)
        self.key_type = key_type
        it = (C.c_int32 * max(1, len(input_types)))(*input_types)
        ag = (VgpuAgg * max(1, len(aggs)))()
        for i, (kind, inp, na_rm) in enumerate(aggs):
            ag[i].kind = AGG_NAMES[kind] if isinstance(kind, str) else kind
            ag[i].input = inp
            ag[i].na_rm = na_rm
        self._pred = pred
        self.h = lib.vgpu_batch_create(key_type, len(input_types), it, len(aggs), ag, order,
                                       C.byref(pred) if pred is not None else None, n_pred_cols, param_insn)
        if not self.h:
            raise VectraError(lib.vgpu_last_error().decode())

    def plan(self, queries: Sequence[dict]) -> None:
        """queries: dicts with key (DeviceColumn or None), inputs, pred_cols,
        n_rows and param (float or int: its 8 bytes)."""
        qa = (VgpuQuery * len(queries))()
        keep = []
        for i, q in enumerate(queries):
            if q.get("key") is not None:
                qa[i].key = q["key"].col
            if q.get("probe_key") is not None:
                qa[i].probe_key = q["probe_key"].col
            ins = col_array(q.get("inputs", ()))
            pcs = col_array(q.get("pred_cols", ()))
            keep += [ins, pcs]
            qa[i].inputs = C.cast(ins, C.POINTER(VgpuCol))
            qa[i].pred_cols = C.cast(pcs, C.POINTER(VgpuCol))
            qa[i].n_rows = q["n_rows"]
            qa[i].param = _param_bits(q.get("param", 0))
        self.n_queries = len(queries)
        self._keep = (qa, keep)
        _vcheck(vgpu().vgpu_batch_plan(self.h, qa, len(queries)))

    def join(self, table: "JoinTable", key_payload: int) -> None:
        self._join = table
        _vcheck(vgpu().vgpu_batch_join(self.h, table.h, key_payload))

    def run(self) -> None:
        _vcheck(vgpu().vgpu_batch_run(self.h))

    def rerun(self, params: Sequence[float | int], n_active: int | None = None) -> None:
        """params: floats (double constants) or ints (int64 constants)"""
        p = np.array([_param_bits(x) for x in params], np.uint64)
        _vcheck(vgpu().vgpu_batch_rerun(self.h, p.ctypes.data, len(p) if n_active is None else n_active))

    def bucketed(self) -> bool:
        return bool(vgpu().vgpu_batch_bucketed(self.h))

    def capture(self) -> None:
        _vcheck(vgpu().vgpu_batch_capture(self.h))

    def counts(self) -> tuple[np.ndarray, np.ndarray]:
        n = np.zeros(self.n_queries, np.int64)
        off = np.zeros(self.n_queries, np.int64)
        _vcheck(vgpu().vgpu_batch_counts(self.h, n.ctypes.data, off.ctypes.data))
        return n, off

    def fetch(self, q: int) -> dict[str, tuple[np.ndarray, np.ndarray]]:
        n_groups, _ = self.counts()
        n = int(n_groups[q])
        keys = np.zeros(max(n, 1), np.int64)
        kv = np.zeros(max((n + 7) // 8, 1), np.uint8)
        vals = [np.zeros(max(n, 1), np.float64) for _ in range(self.n_aggs)]
        valid = [np.zeros(max((n + 7) // 8, 1), np.uint8) for _ in range(self.n_aggs)]
        vp = (C.c_void_p * max(1, self.n_aggs))(*[v.ctypes.data for v in vals])
        mp = (C.c_void_p * max(1, self.n_aggs))(*[m.ctypes.data for m in valid])
        w = vgpu().vgpu_batch_fetch(self.h, q, keys.ctypes.data, kv.ctypes.data, vp, mp)
        _vcheck(w)
        unpack = lambda b: np.unpackbits(b, bitorder="little")[:n].astype(bool)
        out = {}
        if self.key_type == VEC_DOUBLE:
            out["key"] = (keys[:n].view(np.float64), unpack(kv))
        elif self.key_type == VEC_BOOL:
            out["key"] = (keys[:n].astype(bool), unpack(kv))
        elif self.key_type != VGPU_KEY_NONE:
            out["key"] = (keys[:n], unpack(kv))
        for i in range(self.n_aggs):
            out[f"a{i}"] = (vals[i][:n], unpack(valid[i]))
        return out

    def free(self) -> None:
        if self.h:
            vgpu().vgpu_batch_destroy(self.h)
            self.h = None


def serve(fact: dict, dim: dict, requests: Sequence[tuple], batched: bool, max_batch: int = 32,
          state_budget: int = 0, keep_results: bool = True, in_flight: int = 2,
          hbm_budget: int = 0) -> tuple[list[dict], dict]:
    """Runs the C request loop (include/vx_serve.h) over device columns.
    fact: {"key", "value", "group"} DeviceColumns; dim: {"key", "attr"};
    requests: (cls, threshold, arrival_s) in arrival order."""
    n = len(requests)
    cfg = VxServeCfg()
    cfg.fact_key, cfg.fact_value, cfg.fact_group = fact["key"].col, fact["value"].col, fact["group"].col
    cfg.n_rows = fact["value"].n
    cfg.dim_key, cfg.dim_attr = dim["key"].col, dim["attr"].col
    cfg.n_dim = dim["key"].n
    cfg.batched = 1 if batched else 0
    cfg.max_batch = max_batch
    cfg.state_budget = state_budget
    cfg.keep_results = 1 if keep_results else 0
    cfg.in_flight = in_flight
    cfg.hbm_budget = hbm_budget
    reqs = (VxRequest * max(n, 1))()
    for i, (cls, t, at) in enumerate(requests):
        reqs[i].cls, reqs[i].threshold, reqs[i].arrival_s = cls, t, at
    out = (VxResult * max(n, 1))()
    st = VxServeStats()
    lib = nodes()
    if lib.vx_serve_run(C.byref(cfg), reqs, n, out, C.byref(st)) != 0:
        raise VectraError(vgpu().vgpu_last_error().decode())
    res = []
    for i in range(n):
        r = out[i]
        d = {"start_s": r.start_s, "end_s": r.end_s, "pass": r.pass_, "batch": r.batch, "n_groups": r.n_groups}
        if keep_results:
            g = r.n_groups
            bits = lambda p: np.unpackbits(np.ctypeslib.as_array(p, ((g + 7) // 8 + 1,)).copy(),
                                           bitorder="little")[:g].astype(bool)
            d["key"] = (np.ctypeslib.as_array(r.keys, (max(g, 1),))[:g].copy(), bits(r.key_valid))
            d["s"] = (np.ctypeslib.as_array(r.sum, (max(g, 1),))[:g].copy(), bits(r.sum_valid))
            d["n"] = (np.ctypeslib.as_array(r.count, (max(g, 1),))[:g].copy(), np.ones(g, bool))
            lib.vx_result_free(C.byref(out[i]))
        res.append(d)
    stats = {k: getattr(st, k) for k, _ in VxServeStats._fields_}
    return res, stats
