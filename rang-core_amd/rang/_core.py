"""The gridder slice of rang-core's extension module `rang._core`, on the
MI355X library.

Same function names, argument order, checks and return values as the
reference's PyO3 bindings (core/py/src/lib.rs:1055-1446): option lists of 18
floats, NumPy buffers borrowed in place, status codes returned. The calls go
through the C ABI of include/rang_core_wgridder.h (ctypes), whose work runs on
the GPU. There is no CPU fallback: without the built library every call
raises.
"""

import ctypes
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parents[1] / "lib" / "librang_core_amd.so"
N_INFO = 14

VIS_UNIT, VIS_C64, VIS_C128 = 0, 1, 2
WGT_NONE, WGT_F32, WGT_F64, WGT_F32_PRODUCT = 0, 1, 2, 3


class WgridderOptions(ctypes.Structure):
    """core/src/ffi_wgridder.rs:22-43."""

    _fields_ = [
        ("nx", ctypes.c_size_t),
        ("ny", ctypes.c_size_t),
        ("pixsize_x", ctypes.c_double),
        ("pixsize_y", ctypes.c_double),
        ("epsilon", ctypes.c_double),
        ("center_x", ctypes.c_double),
        ("center_y", ctypes.c_double),
        ("sigma_min", ctypes.c_double),
        ("sigma_max", ctypes.c_double),
        ("do_wgridding", ctypes.c_int32),
        ("divide_by_n", ctypes.c_int32),
        ("flip_u", ctypes.c_int32),
        ("flip_v", ctypes.c_int32),
        ("flip_w", ctypes.c_int32),
        ("allow_nshift", ctypes.c_int32),
        ("kernel_index", ctypes.c_int64),
        ("nu", ctypes.c_size_t),
        ("nv", ctypes.c_size_t),
    ]


class BatchImage(ctypes.Structure):
    """include/rang_core_batch.h RangBatchImage."""

    _fields_ = [
        ("opts", WgridderOptions),
        ("nrow", ctypes.c_size_t),
        ("nchan", ctypes.c_size_t),
        ("uvw", ctypes.c_void_p),
        ("freq", ctypes.c_void_p),
        ("vis", ctypes.c_void_p),
        ("wgt", ctypes.c_void_p),
        ("mask", ctypes.c_void_p),
        ("dirty", ctypes.c_void_p),
        ("vis_kind", ctypes.c_int32),
        ("wgt_kind", ctypes.c_int32),
        ("tile", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("info", ctypes.c_double * 14),
    ]


class TuneReport(ctypes.Structure):
    """include/rang_core_batch.h RangTuneReport."""

    _fields_ = [
        ("ms", ctypes.c_double),
        ("default_ms", ctypes.c_double),
        ("cached", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


# Every symbol of include/rang_core_wgridder.h with its ctypes signature.
_vp, _sz, _i32, _dp = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int32, ctypes.c_double
_OPT = ctypes.POINTER(WgridderOptions)


class FistaSettings(ctypes.Structure):
    """RangFistaSettings (include/rang_core_wgridder.h): prox_solver::Settings
    (core/src/prox_solver.rs:175-208), NaN for a tolerance that is None."""

    _fields_ = [
        ("step", ctypes.c_double),
        ("max_iterations", ctypes.c_uint64),
        ("momentum", ctypes.c_int32),
        ("restart", ctypes.c_int32),
        ("gap_tolerance", ctypes.c_double),
        ("kkt_tolerance", ctypes.c_double),
        ("step_tolerance", ctypes.c_double),
    ]


class FistaRecord(ctypes.Structure):
    """RangFistaRecord: prox_solver::Record (core/src/prox_solver.rs:211-231)."""

    _fields_ = [
        ("iteration", ctypes.c_uint64),
        ("objective", ctypes.c_double),
        ("misfit", ctypes.c_double),
        ("penalty", ctypes.c_double),
        ("gap", ctypes.c_double),
        ("kkt", ctypes.c_double),
        ("nonzero", ctypes.c_uint64),
        ("flux", ctypes.c_double),
        ("curvature", ctypes.c_double),
        ("relative_step", ctypes.c_double),
        ("restarted", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
    ]


class FistaOutcome(ctypes.Structure):
    """RangFistaOutcome: prox_solver::Outcome without x and G x
    (core/src/prox_solver.rs:234-249)."""

    _fields_ = [
        ("iterations", ctypes.c_uint64),
        ("converged", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("restarts", ctypes.c_uint64),
        ("objective_rises", ctypes.c_uint64),
        ("largest_rise", ctypes.c_double),
        ("negative_curvatures", ctypes.c_uint64),
        ("min_curvature", ctypes.c_double),
        ("max_curvature", ctypes.c_double),
        ("last", FistaRecord),
    ]


class OuterSchedule(ctypes.Structure):
    """RangOuterSchedule: prox_solver::OuterSchedule (core/src/prox_solver.rs:760-772)."""

    _fields_ = [
        ("cycles", ctypes.c_uint64),
        ("kkt", ctypes.c_double),
        ("lambda_", ctypes.c_double),
        ("gamma", ctypes.c_double),
    ]


class OuterRecord(ctypes.Structure):
    """RangOuterRecord: prox_solver::OuterRecord (core/src/prox_solver.rs:447-462)."""

    _fields_ = [
        ("outer", ctypes.c_uint64),
        ("objective", ctypes.c_double),
        ("misfit", ctypes.c_double),
        ("gap", ctypes.c_double),
        ("kkt", ctypes.c_double),
        ("inner_iterations", ctypes.c_uint64),
        ("defect_sup", ctypes.c_double),
    ]


class DefectOutcome(ctypes.Structure):
    """RangDefectOutcome: how defect_correction ended (core/src/prox_solver.rs:778-875)."""

    _fields_ = [
        ("iterations", ctypes.c_uint64),
        ("converged", ctypes.c_int32),
        ("reserved", ctypes.c_int32),
        ("cycles", ctypes.c_uint64),
        ("restarts", ctypes.c_uint64),
        ("negative_curvatures", ctypes.c_uint64),
        ("exact_applies", ctypes.c_uint64),
        ("cheap_applies", ctypes.c_uint64),
        ("last", OuterRecord),
    ]


class GridInputs(ctypes.Structure):
    """`GridInputs` (core/src/ffi_grid.rs:7-29)."""

    _fields_ = [(n, ctypes.c_size_t) for n in ("nrow", "nchan", "width", "height", "kernel_size",
                                                "oversampling", "n_w_layers")] + [
        ("pixel_size_x", ctypes.c_double), ("pixel_size_y", ctypes.c_double),
        ("padding", ctypes.c_double), ("uvw", ctypes.c_void_p), ("frequencies", ctypes.c_void_p),
        ("data", ctypes.c_void_p), ("weights", ctypes.c_void_p), ("imaging_weights", ctypes.c_void_p)]


class GridOutputs(ctypes.Structure):
    """`GridOutputs` (core/src/ffi_grid.rs:31-45)."""

    _fields_ = [("dirty", ctypes.c_void_p), ("psf", ctypes.c_void_p), ("stats", ctypes.c_void_p)]


class BloscChunk(ctypes.Structure):
    """`RangBloscChunk` (include/rang_core_blosc.h)."""

    _fields_ = [("src", ctypes.c_void_p), ("len", ctypes.c_size_t), ("dst", ctypes.c_void_p),
                ("dst_len", ctypes.c_size_t)]


class MsColumn(ctypes.Structure):
    """`MsColumn` (core/src/ffi_ms.rs:24-33)."""

    _fields_ = [("data", ctypes.c_void_p), ("nbytes", ctypes.c_size_t), ("nrows", ctypes.c_uint64),
                ("ndim", ctypes.c_size_t), ("shape", ctypes.c_int64 * 8), ("dtype", ctypes.c_int32),
                ("nstrings", ctypes.c_size_t)]


GRID_N_STATS = 8
GRID_N_TIMINGS = 14
_i64p = ctypes.POINTER(ctypes.c_int64)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_colp = ctypes.POINTER(MsColumn)
SIGNATURES = {
    "rang_core_ms_last_error": ([_vp, _sz], _sz),
    "rang_core_ms_open": ([ctypes.c_char_p, ctypes.POINTER(_vp)], _i32),
    "rang_core_ms_close": ([_vp], None),
    "rang_core_ms_describe": ([_vp, _vp, _sz], _sz),
    "rang_core_ms_read": ([_vp, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64, _colp], _i32),
    "rang_core_ms_free": ([_colp], None),
    "rang_amd_ms_read_slice": ([_vp, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64, _i64p, _sz, _colp],
                               _i32),
    "rang_amd_ms_read_rows": ([_vp, ctypes.c_char_p, _u64p, _sz, _i64p, _sz, _colp], _i32),
    "rang_amd_ms_read_dev": ([_vp, ctypes.c_char_p, ctypes.c_uint64, ctypes.c_uint64, _i64p, _sz, _vp, _sz,
                              _vp, _colp], _i32),
    "rang_core_blosc_size": ([_vp, _sz, ctypes.POINTER(_sz)], _i32),
    "rang_core_blosc_header": ([_vp, _sz, ctypes.POINTER(_sz)], _i32),
    "rang_core_blosc_decode": ([_vp, _sz, _vp, _sz], _i32),
    "rang_amd_blosc_decode_batch": ([_vp, _sz, _vp], _i32),
    "rang_amd_blosc_decode_resident": ([_vp, _vp, _sz, _vp], _i32),
    "rang_amd_blosc_last_error": ([_vp, _sz], _sz),
    "rang_amd_release_workspaces": ([], _i32),
    "rang_amd_wgridder_batch_vis2dirty_f32": ([ctypes.POINTER(BatchImage), _sz, _sz, _vp], _i32),
    "rang_amd_wgridder_batch_vis2dirty_f64": ([ctypes.POINTER(BatchImage), _sz, _sz, _vp], _i32),
    "rang_amd_wgridder_tune_f32": (
        [ctypes.POINTER(BatchImage), ctypes.c_char_p, _i32, ctypes.POINTER(TuneReport), _vp],
        _i32,
    ),
    "rang_amd_wgridder_tune_f64": (
        [ctypes.POINTER(BatchImage), ctypes.c_char_p, _i32, ctypes.POINTER(TuneReport), _vp],
        _i32,
    ),
    "rang_core_wgridder_vis2dirty_f64": ([_OPT, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i32),
    "rang_core_wgridder_vis2dirty_f32": ([_OPT, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i32),
    "rang_core_wgridder_dirty2vis_f64": ([_OPT, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i32),
    "rang_core_wgridder_dirty2vis_f32": ([_OPT, _sz, _sz, _vp, _vp, _vp, _vp, _vp, _vp, _vp], _i32),
    "rang_core_wgridder_vis2dirty_ex_f64": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp, _i32, _vp, _vp, _vp, _vp], _i32),
    "rang_core_wgridder_vis2dirty_ex_f32": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp, _i32, _vp, _vp, _vp, _vp], _i32),
    "rang_core_wgridder_vis2dirty_ex2_f64": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp], _i32),
    "rang_core_wgridder_vis2dirty_ex2_f32": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp], _i32),
    "rang_core_wgridder_dirty2vis_ex_f64": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp], _i32),
    "rang_core_wgridder_dirty2vis_ex_f32": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp], _i32),
    "rang_core_wgridder_kernel_row": ([_sz, _vp], _i32),
    "rang_core_wgridder_last_error": ([_vp, _sz], _sz),
    "rang_amd_wgridder_vis2dirty_dev_f64": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_vis2dirty_dev_f32": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_vis2dirty_noalloc_f32": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp, _vp, _i32, _vp, _vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_dirty2vis_dev_f64": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_dirty2vis_dev_f32": (
        [_OPT, _sz, _sz, _vp, _vp, _vp, _vp, _i32, _vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_plan_f64": ([_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp, _i32, _vp, _vp], _i32),
    "rang_amd_wgridder_plan_f32": ([_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp, _i32, _vp, _vp], _i32),
    "rang_amd_wgridder_footprint_f64": ([_OPT, _sz, _sz, _vp, _vp, _vp, _sz, _vp, _vp], _i32),
    "rang_amd_wgridder_footprint_f32": ([_OPT, _sz, _sz, _vp, _vp, _vp, _sz, _vp, _vp], _i32),
    "rang_amd_wgridder_prepare_dev_f64": ([_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp], _vp),
    "rang_amd_wgridder_prepare_dev_f32": ([_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _vp], _vp),
    "rang_amd_wgridder_prepared_vis2dirty_f64": ([_vp, _vp, _i32, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_vis2dirty_f32": ([_vp, _vp, _i32, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_dirty2vis_f64": ([_vp, _vp, _i32, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_dirty2vis_f32": ([_vp, _vp, _i32, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_residual_f64": ([_vp, _vp, _i32, _vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_residual_f32": ([_vp, _vp, _i32, _vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_residual_sorted_f64": ([_vp, _vp, _i32, _vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_residual_sorted_f32": ([_vp, _vp, _i32, _vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_sorted_f64": ([_vp, _vp, _sz, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_sorted_f32": ([_vp, _vp, _sz, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_gram_f64": ([_vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_gram_f32": ([_vp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_cg_f64": ([_vp, _vp, _vp, _i32, _dp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_cg_f32": ([_vp, _vp, _vp, _i32, _dp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_fista_f64": (
        [_vp, _vp, _dp, _dp, _vp, _i32, _vp, _vp, _vp, _vp, _sz, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_fista_f32": (
        [_vp, _vp, _dp, _dp, _vp, _i32, _vp, _vp, _vp, _vp, _sz, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_power_f64": ([_vp, ctypes.c_uint64, _dp, ctypes.c_uint64, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_power_f32": ([_vp, ctypes.c_uint64, _dp, ctypes.c_uint64, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_active": ([_vp], _sz),
    "rang_amd_wgridder_prepared_memory": ([_vp], _sz),
    "rang_amd_wgridder_prepared_release_grids": ([_vp], None),
    "rang_amd_wgridder_prepared_free": ([_vp], None),
    "rang_core_grid_invert": ([ctypes.POINTER(GridInputs), ctypes.POINTER(GridOutputs)], _i32),
    "rang_core_grid_invert_gpu": (
        [ctypes.POINTER(GridInputs), ctypes.POINTER(GridOutputs), _i32, _vp], _i32),
    "rang_core_grid_kernel": ([_sz, _sz, _vp], _i32),
    "rang_amd_grid_last_error": ([_vp, _sz], _sz),
    "rang_amd_device_name": ([_vp, _sz], _sz),
    # the learned preconditioner's kernels (include/rang_core_precond.h)
    "rang_amd_nn_last_error": ([_vp, _sz], _sz),
    "rang_amd_nn_conv_fwd": ([_vp, _i32, _i32, _vp, _vp, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _i32, _i32,
                              _i32, _vp], _i32),
    "rang_amd_nn_conv_wgrad": ([_vp, _i32, _i32, _vp, _i32, _i32, _vp, _i32, _i32, _i32, _i32, _i32, _i32,
                                _vp], _i32),
    "rang_amd_nn_flip": ([_vp, _vp, _i32, _i32, _i32, _vp], _i32),
    "rang_amd_nn_to_bf16": ([_vp, _vp, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_relu_bwd": ([_vp, _i32, _i32, _vp, _i32, _i32, _vp, _i32, _i32, ctypes.c_int64, _i32, _vp],
                             _i32),
    "rang_amd_nn_pool": ([_vp, _i32, _i32, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp], _i32),
    "rang_amd_nn_pool_bwd": ([_vp, _i32, _i32, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp], _i32),
    "rang_amd_nn_up": ([_vp, _i32, _i32, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp], _i32),
    "rang_amd_nn_up_bwd": ([_vp, _i32, _i32, _vp, _i32, _i32, _i32, _i32, _i32, _i32, _vp], _i32),
    "rang_amd_nn_adam": ([_vp, _vp, _vp, _vp, _vp, ctypes.c_int64, _dp, _dp, _dp, _dp, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_pack_f64": ([_vp, _vp, _vp, _i32, _i32, _i32, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_pack_f32": ([_vp, _vp, _vp, _i32, _i32, _i32, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_unpack_f64": ([_vp, _i32, _i32, _vp, _vp, _i32, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_unpack_f32": ([_vp, _i32, _i32, _vp, _vp, _i32, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_image_dot_f64": ([_vp, _vp, _vp, _i32, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_image_dot_f32": ([_vp, _vp, _vp, _i32, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_loss": ([_vp, _i32, _i32, _vp, _vp, _vp, _i32, _i32, ctypes.c_int64, _vp, _vp], _i32),
    "rang_amd_nn_sky_f64": ([_vp, _i32, _i32, _i32, _i32, _dp, _dp, _i32, ctypes.c_uint64, _vp], _i32),
    "rang_amd_nn_sky_f32": ([_vp, _i32, _i32, _i32, _i32, _dp, _dp, _i32, ctypes.c_uint64, _vp], _i32),
    "rang_amd_nn_scale_images": ([_vp, _vp, _vp, _i32, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_grad_images": ([_vp, _vp, _vp, _i32, _i32, _i32, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_axpby_f64": ([_vp, _vp, _dp, _dp, ctypes.c_int64, _vp], _i32),
    "rang_amd_nn_axpby_f32": ([_vp, _vp, _dp, _dp, ctypes.c_int64, _vp], _i32),
    # the normal operator without visibilities (include/rang_core_gram.h)
    "rang_amd_wexpansion": ([_i32, _dp, _dp, _vp, _vp, _vp], _i32),
    "rang_amd_gram_new_f64": ([_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _i32, _dp, _dp, _dp, _vp], _vp),
    "rang_amd_gram_new_f32": ([_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _i32, _dp, _dp, _dp, _vp], _vp),
    "rang_amd_gram_no_w_f64": ([_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _dp, _vp], _vp),
    "rang_amd_gram_no_w_f32": ([_OPT, _sz, _sz, _vp, _vp, _vp, _i32, _dp, _vp], _vp),
    "rang_amd_gram_inverse_filter": ([_vp, _dp, _vp], _vp),
    "rang_amd_gram_apply_f64": ([_vp, _vp, _vp, _vp], _i32),
    "rang_amd_gram_apply_f32": ([_vp, _vp, _vp, _vp], _i32),
    "rang_amd_gram_stats": ([_vp, _vp], _i32),
    "rang_amd_gram_side": ([_vp], _sz),
    "rang_amd_gram_free": ([_vp], None),
    "rang_amd_wgridder_prepared_pcg_f64": ([_vp, _vp, _vp, _vp, _i32, _dp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_pcg_f32": ([_vp, _vp, _vp, _vp, _i32, _dp, _vp, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_defect_f64": (
        [_vp, _vp, _vp, _dp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp], _i32),
    "rang_amd_wgridder_prepared_defect_f32": (
        [_vp, _vp, _vp, _dp, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _sz, _vp, _vp], _i32),
    # imaging weights and Stokes I (include/rang_core_weights.h)
    "rang_core_weight_grid_shape": ([_sz, _sz, _dp, _vp], _i32),
    "rang_core_padded_size": ([_sz, _dp], _sz),
    "rang_core_stokes_i": ([_vp, _vp], _i32),
    "rang_core_imaging_weights": ([_vp, _vp], _i32),
    "rang_amd_stokes_i_dev": ([_vp, _vp, _vp], _i32),
    "rang_amd_imaging_weights_dev": ([_vp, _vp, _vp], _i32),
    "rang_amd_stage_timing": ([_i32], None),
    "rang_amd_stage_count": ([], _i32),
    "rang_amd_stage_name": ([_i32], ctypes.c_char_p),
    "rang_amd_stage_ms": ([_i32], _dp),
}

_lib = None


def library():
    """The loaded C-ABI library; raises when it was not built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"the MI355X gridder library is not built ({LIB_PATH}); "
                "run __graft_entry__.build() or make -C rang-core_amd"
            )
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, (args, res) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = res
        _lib = lib
    return _lib


def library_path():
    return str(LIB_PATH)


def options(o):
    """`WgridderOptions` from the 18 option floats in declaration order
    (core/py/src/lib.rs:1055-1080)."""
    if len(o) != 18:
        raise ValueError("18 gridder options required")
    return WgridderOptions(
        nx=int(o[0]), ny=int(o[1]), pixsize_x=o[2], pixsize_y=o[3], epsilon=o[4],
        center_x=o[5], center_y=o[6], sigma_min=o[7], sigma_max=o[8],
        do_wgridding=int(o[9]), divide_by_n=int(o[10]), flip_u=int(o[11]), flip_v=int(o[12]),
        flip_w=int(o[13]), allow_nshift=int(o[14]), kernel_index=int(o[15]), nu=int(o[16]),
        nv=int(o[17]),
    )


def _contig(a, name):
    """The reference refuses non-C-contiguous arrays (core/py/src/lib.rs:113-137)."""
    if not a.flags.c_contiguous:
        raise ValueError(f"{name} must be C-contiguous")
    return a


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _vis_arg(vis):
    if vis is None:
        return None, VIS_UNIT
    vis = np.asarray(vis)
    if vis.dtype == np.complex64:
        return _contig(vis, "vis"), VIS_C64
    if vis.dtype == np.complex128:
        return _contig(vis, "vis"), VIS_C128
    raise ValueError("vis must be complex64 or complex128")


def _wgt_arg(wgt):
    if wgt is None:
        return None, WGT_NONE
    wgt = np.asarray(wgt)
    if wgt.dtype == np.float32:
        return _contig(wgt, "wgt"), WGT_F32
    if wgt.dtype == np.float64:
        return _contig(wgt, "wgt"), WGT_F64
    raise ValueError("weights must be float32 or float64")


def _checked(a, dtype, name):
    if a is None:
        return None
    if not isinstance(a, np.ndarray) or a.dtype != dtype:
        raise TypeError(f"{name} must be a {np.dtype(dtype).name} array")
    return _contig(a, name)


def _v2d(opts, uvw, freq, vis, wgt, wgt2, mask, dirty, psf, info, real, fn):
    """core/py/src/lib.rs:1224-1280 (wgridder_v2d)."""
    o = options(opts)
    uvw = _checked(uvw, np.float64, "uvw")
    freq = _checked(freq, np.float64, "freq")
    nrow, nchan = uvw.size // 3, freq.size
    n, npix = nrow * nchan, o.nx * o.ny
    vis, vis_kind = _vis_arg(vis)
    wgt, wgt_kind = _wgt_arg(wgt)
    wgt2 = _checked(wgt2, np.float32, "wgt2")
    if wgt2 is not None:
        if wgt_kind != WGT_F32:
            raise ValueError("two weights must both be float32")
        wgt_kind = WGT_F32_PRODUCT
    mask = _checked(mask, np.uint8, "mask")
    dirty = _checked(dirty, real, "dirty")
    psf = _checked(psf, real, "psf")
    info = _checked(info, np.float64, "info")
    ok = (
        uvw.size == nrow * 3
        and (vis is None or vis.size == n)
        and (wgt is None or wgt.size == n)
        and (wgt2 is None or wgt2.size == n)
        and (mask is None or mask.size == n)
        and (dirty is None or dirty.size == npix)
        and (psf is None or psf.size == npix)
        and info.size == N_INFO
    )
    if not ok:
        raise ValueError("gridder array sizes disagree")
    return int(fn(ctypes.byref(o), nrow, nchan, _ptr(uvw), _ptr(freq), _ptr(vis), vis_kind,
                  _ptr(wgt), _ptr(wgt2), wgt_kind, _ptr(mask), _ptr(dirty), _ptr(psf), _ptr(info)))


def wgridder_vis2dirty_f64(opts, uvw, freq, vis, wgt, wgt2, mask, dirty, psf, info):
    """`rang_core_wgridder_vis2dirty_ex2_f64`: data of either complex precision
    and weights of either real precision read in place; float64 outputs;
    returns the status code (core/py/src/lib.rs:1283-1318)."""
    return _v2d(opts, uvw, freq, vis, wgt, wgt2, mask, dirty, psf, info, np.float64,
                library().rang_core_wgridder_vis2dirty_ex2_f64)


def wgridder_vis2dirty_f32(opts, uvw, freq, vis, wgt, wgt2, mask, dirty, psf, info):
    """As `wgridder_vis2dirty_f64`, working in single precision
    (core/py/src/lib.rs:1320-1355)."""
    return _v2d(opts, uvw, freq, vis, wgt, wgt2, mask, dirty, psf, info, np.float32,
                library().rang_core_wgridder_vis2dirty_ex2_f32)


def _d2v(opts, uvw, freq, dirty, wgt, mask, vis, info, real, fn):
    """core/py/src/lib.rs:1357-1395 (wgridder_d2v); `vis` is the real view."""
    o = options(opts)
    uvw = _checked(uvw, np.float64, "uvw")
    freq = _checked(freq, np.float64, "freq")
    nrow, nchan = uvw.size // 3, freq.size
    n = nrow * nchan
    dirty = _checked(dirty, real, "dirty")
    wgt, wgt_kind = _wgt_arg(wgt)
    mask = _checked(mask, np.uint8, "mask")
    vis = _checked(vis, real, "vis")
    info = _checked(info, np.float64, "info")
    ok = (
        uvw.size == nrow * 3
        and dirty.size == o.nx * o.ny
        and vis.size == 2 * n
        and (wgt is None or wgt.size == n)
        and (mask is None or mask.size == n)
        and info.size == N_INFO
    )
    if not ok:
        raise ValueError("gridder array sizes disagree")
    return int(fn(ctypes.byref(o), nrow, nchan, _ptr(uvw), _ptr(freq), _ptr(dirty), _ptr(wgt),
                  wgt_kind, _ptr(mask), _ptr(vis), _ptr(info)))


def wgridder_dirty2vis_f64(opts, uvw, freq, dirty, wgt, mask, vis, info):
    """`rang_core_wgridder_dirty2vis_ex_f64` (core/py/src/lib.rs:1397-1420)."""
    return _d2v(opts, uvw, freq, dirty, wgt, mask, vis, info, np.float64,
                library().rang_core_wgridder_dirty2vis_ex_f64)


def wgridder_dirty2vis_f32(opts, uvw, freq, dirty, wgt, mask, vis, info):
    """`rang_core_wgridder_dirty2vis_ex_f32` (core/py/src/lib.rs:1422-1445)."""
    return _d2v(opts, uvw, freq, dirty, wgt, mask, vis, info, np.float32,
                library().rang_core_wgridder_dirty2vis_ex_f32)


def wgridder_kernel_row(index, out):
    """Row `index` of the kernel table into seven doubles; 1 past the end."""
    out = _checked(out, np.float64, "out")
    if out.size != 7:
        return 1
    return int(library().rang_core_wgridder_kernel_row(int(index), _ptr(out)))


def wgridder_last_error():
    """The message of this thread's last status-1 gridder failure."""
    lib = library()
    n = lib.rang_core_wgridder_last_error(None, 0)
    buf = ctypes.create_string_buffer(max(int(n), 1))
    lib.rang_core_wgridder_last_error(ctypes.cast(buf, ctypes.c_void_p), n)
    return buf.raw[:n].decode("utf-8", "replace")


# ---------------------------------------------------------------- w-stacking
# The grid_* functions of rang._core (core/py/src/lib.rs:906-1049).


def _grid_structs(nrow, nchan, width, height, kernel_size, oversampling, n_w_layers, pixel_size_x,
                  pixel_size_y, padding, uvw, frequencies, data, weights, imaging_weights, dirty, psf,
                  stats):
    cells, pixels = nrow * nchan, width * height
    uvw = _checked(uvw, np.float64, "uvw")
    frequencies = _checked(frequencies, np.float64, "frequencies")
    data = _checked(data, np.float32, "data")
    weights = _checked(weights, np.float32, "weights")
    imaging_weights = _checked(imaging_weights, np.float32, "imaging_weights")
    dirty = _checked(dirty, np.float64, "dirty")
    psf = _checked(psf, np.float64, "psf")
    stats = _checked(stats, np.float64, "stats")
    if (uvw.size != 3 * nrow or frequencies.size != nchan or data.size != 2 * cells
            or weights.size != cells or imaging_weights.size != cells or dirty.size != pixels
            or psf.size != pixels or stats.size != GRID_N_STATS):
        raise ValueError("gridder array sizes disagree")
    inp = GridInputs(nrow, nchan, width, height, kernel_size, oversampling, n_w_layers,
                     pixel_size_x, pixel_size_y, padding, uvw.ctypes.data, frequencies.ctypes.data,
                     data.ctypes.data, weights.ctypes.data, imaging_weights.ctypes.data)
    out = GridOutputs(dirty.ctypes.data, psf.ctypes.data, stats.ctypes.data)
    return inp, out


def grid_invert(nrow, nchan, width, height, kernel_size, oversampling, n_w_layers, pixel_size_x,
                pixel_size_y, padding, uvw, frequencies, data, weights, imaging_weights, dirty, psf,
                stats):
    """`rang_core_grid_invert` into caller-owned arrays; returns its status."""
    inp, out = _grid_structs(nrow, nchan, width, height, kernel_size, oversampling, n_w_layers,
                             pixel_size_x, pixel_size_y, padding, uvw, frequencies, data, weights,
                             imaging_weights, dirty, psf, stats)
    return int(library().rang_core_grid_invert(ctypes.byref(inp), ctypes.byref(out)))


def grid_invert_gpu(nrow, nchan, width, height, kernel_size, oversampling, n_w_layers, pixel_size_x,
                    pixel_size_y, padding, uvw, frequencies, data, weights, imaging_weights,
                    precision, dirty, psf, stats, timings):
    """`rang_core_grid_invert_gpu`; `precision` "f64" or "f32"."""
    if precision not in ("f64", "f32"):
        raise ValueError(f"precision must be 'f64' or 'f32', not {precision!r}")
    timings = _checked(timings, np.float64, "timings")
    if timings.size != GRID_N_TIMINGS:
        raise ValueError("gridder array sizes disagree")
    inp, out = _grid_structs(nrow, nchan, width, height, kernel_size, oversampling, n_w_layers,
                             pixel_size_x, pixel_size_y, padding, uvw, frequencies, data, weights,
                             imaging_weights, dirty, psf, stats)
    return int(library().rang_core_grid_invert_gpu(ctypes.byref(inp), ctypes.byref(out),
                                                   0 if precision == "f64" else 1, _ptr(timings)))


def grid_kernel(kernel_size, oversampling, kernel):
    """`rang_core_grid_kernel` into a float64 array of kernel_size * oversampling."""
    kernel = _checked(kernel, np.float64, "kernel")
    if kernel.size != kernel_size * oversampling:
        return 1
    return int(library().rang_core_grid_kernel(int(kernel_size), int(oversampling), _ptr(kernel)))


def grid_last_error():
    lib = library()
    n = lib.rang_amd_grid_last_error(None, 0)
    buf = ctypes.create_string_buffer(max(int(n), 1))
    lib.rang_amd_grid_last_error(ctypes.cast(buf, ctypes.c_void_p), n)
    return buf.raw[:n].decode("utf-8", "replace")


def grid_gpu_adapter():
    """The GPU the library runs on (core/py/src/lib.rs:1028-1030)."""
    lib = library()
    buf = ctypes.create_string_buffer(256)
    n = lib.rang_amd_device_name(ctypes.cast(buf, ctypes.c_void_p), 255)
    if n == 0:
        raise RuntimeError("no HIP device")
    return buf.raw[:min(int(n), 255)].decode("utf-8", "replace")


# ---------------------------------------------------------------- weights
# The weight functions of rang._core that rang.core_weights calls
# (core/src/ffi_weights.rs through the PyO3 module).


class StokesIInputs(ctypes.Structure):
    """`StokesIInputs` (core/src/ffi_weights.rs:9-19)."""

    _fields_ = [("ncell", ctypes.c_size_t), ("data", ctypes.c_void_p), ("weights", ctypes.c_void_p),
                ("flags", ctypes.c_void_p)]


class StokesIOutputs(ctypes.Structure):
    """`StokesIOutputs` (core/src/ffi_weights.rs:21-27)."""

    _fields_ = [("vis", ctypes.c_void_p), ("weight", ctypes.c_void_p)]


class ImagingWeightInputs(ctypes.Structure):
    """`ImagingWeightInputs` (core/src/ffi_weights.rs:29-55)."""

    _fields_ = [("nrow", ctypes.c_size_t), ("nchan", ctypes.c_size_t), ("uvw", ctypes.c_void_p),
                ("wavelength", ctypes.c_void_p), ("weights", ctypes.c_void_p), ("row_selected", ctypes.c_void_p),
                ("class_", ctypes.c_uint32), ("robust", ctypes.c_double), ("width", ctypes.c_size_t),
                ("height", ctypes.c_size_t), ("pixel_scale_x", ctypes.c_double), ("pixel_scale_y", ctypes.c_double),
                ("super_weight", ctypes.c_double), ("rank_filter_level", ctypes.c_double),
                ("rank_filter_size", ctypes.c_size_t)]


class ImagingWeightOutputs(ctypes.Structure):
    """`ImagingWeightOutputs` (core/src/ffi_weights.rs:57-69)."""

    _fields_ = [("weights", ctypes.c_void_p), ("grid", ctypes.c_void_p), ("sums", ctypes.c_void_p)]


def weight_grid_shape(width, height, super_weight):
    """(status, width, height) of the weight grid after the super weight."""
    shape = np.zeros(2, np.uintp)
    st = int(library().rang_core_weight_grid_shape(int(width), int(height), float(super_weight), _ptr(shape)))
    return st, int(shape[0]), int(shape[1])


def padded_size(trimmed, padding):
    return int(library().rang_core_padded_size(int(trimmed), float(padding)))


def stokes_i(ncell, data, weights, flags, vis, weight):
    """`rang_core_stokes_i` into caller-owned float32 arrays; returns its status."""
    data = _checked(data, np.float32, "data")
    weights = _checked(weights, np.float32, "weights")
    flags = _checked(flags, np.uint8, "flags")
    vis = _checked(vis, np.float32, "vis")
    weight = _checked(weight, np.float32, "weight")
    if (data.size != 8 * ncell or weights.size != 4 * ncell or flags.size != 4 * ncell
            or vis.size != 2 * ncell or weight.size != ncell):
        raise ValueError("stokes I array sizes disagree")
    inp = StokesIInputs(ncell, data.ctypes.data, weights.ctypes.data, flags.ctypes.data)
    out = StokesIOutputs(vis.ctypes.data, weight.ctypes.data)
    return int(library().rang_core_stokes_i(ctypes.byref(inp), ctypes.byref(out)))


def imaging_weights(nrow, nchan, uvw, wavelength, weights, selected, class_, robust, width, height, scale_x,
                    scale_y, super_weight, level, size, out, grid, sums):
    """`rang_core_imaging_weights` into caller-owned arrays; returns its status."""
    uvw = _checked(uvw, np.float64, "uvw")
    wavelength = _checked(wavelength, np.float64, "wavelength")
    weights = _checked(weights, np.float32, "weights")
    selected = _checked(selected, np.uint8, "row_selected")
    out = _checked(out, np.float32, "out")
    sums = _checked(sums, np.float64, "sums")
    if (uvw.size != 3 * nrow or wavelength.size != nchan or weights.size != nrow * nchan
            or selected.size != nrow or out.size != nrow * nchan or sums.size != 4):
        raise ValueError("imaging weight array sizes disagree")
    g = None
    if grid is not None:
        g = _checked(grid, np.float64, "grid")
        st, gw, gh = weight_grid_shape(width, height, super_weight)
        if st or g.size != gw * (gh // 2):
            raise ValueError("imaging weight array sizes disagree")
    inp = ImagingWeightInputs(nrow, nchan, uvw.ctypes.data, wavelength.ctypes.data, weights.ctypes.data,
                              selected.ctypes.data, int(class_), float(robust), int(width), int(height),
                              float(scale_x), float(scale_y), float(super_weight), float(level), int(size))
    o = ImagingWeightOutputs(out.ctypes.data, None if g is None else g.ctypes.data, sums.ctypes.data)
    return int(library().rang_core_imaging_weights(ctypes.byref(inp), ctypes.byref(o)))
