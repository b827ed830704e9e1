"""ctypes binding of libepr_hip.so, the HIP/gfx950 library behind include/epr.h.

The library is loaded from this package directory; a missing or unloadable
library is an error (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libepr_hip.so")

EXPORTS = [
    "epr_default_params", "epr_create", "epr_destroy", "epr_last_error", "epr_set_samples",
    "epr_build", "epr_init", "epr_gs_sweeps", "epr_prolong_add", "epr_energies", "epr_solve",
    "epr_num_functions", "epr_get_coeffs", "epr_set_coeffs", "epr_quadratic_fits", "epr_apply",
    "epr_get_array", "epr_write_tree", "epr_write_grid", "epr_last_seconds", "epr_time_gs_sweeps",
    "epr_time_copy", "epr_subgrid_triangulation", "epr_level_set", "epr_mesh_info", "epr_mesh_copy",
    "epr_mesh_destroy", "epr_mg_create", "epr_mg_num_levels", "epr_mg_level_size", "epr_mg_set_colours",
    "epr_mg_vcycle", "epr_mg_relax", "epr_mg_destroy", "epr_adapted_indexer_write",
    "epr_iso_open", "epr_iso_last_error", "epr_iso_info", "epr_iso_values", "epr_iso_level_set_nodes",
    "epr_iso_extract", "epr_iso_level_set", "epr_iso_write_grid", "epr_iso_seconds", "epr_iso_destroy",
    "epr_indexer_hash", "epr_orient", "epr_write_density", "epr_iso_read_density", "epr_iso_extract_trimmed", "epr_iso_component_densities",
    "epr_build_batch", "epr_solve_batch", "epr_density_grid", "epr_gs_colour", "epr_energies_noalloc",
    "epr_unet_create", "epr_unet_destroy", "epr_unet_last_error", "epr_unet_num_params", "epr_unet_num_nodes",
    "epr_unet_set_params", "epr_unet_get_params", "epr_unet_forward", "epr_unet_loss_gradients",
    "epr_unet_set_dataset", "epr_unet_train", "epr_learned_features", "epr_init_learned", "epr_solve_learned", "epr_time_gs_sweeps_fp32",
    "epr_autotune", "epr_autotune_batch", "epr_tune_get", "epr_tune_set", "epr_time_op", "epr_time_batch_sweeps",
    "epr_solve_schedule", "epr_sample_generate",
]


class Params(ctypes.Structure):
    """epr_params (include/epr.h), the reference's command-line parameters."""
    _fields_ = [("depth", ctypes.c_int), ("seed", ctypes.c_int), ("tRadius", ctypes.c_int),
                ("dRadius", ctypes.c_int), ("minSolveDepth", ctypes.c_int),
                ("maxSolveDepth", ctypes.c_int), ("gsIters", ctypes.c_int),
                ("singleLevel", ctypes.c_int), ("verbose", ctypes.c_int),
                ("nSigma", ctypes.c_double), ("scale", ctypes.c_double),
                ("sWeight", ctypes.c_double), ("dWeight", ctypes.c_double),
                ("iMult", ctypes.c_double), ("aRadius", ctypes.c_int), ("separate", ctypes.c_int),
                ("jitter", ctypes.c_double), ("fp32Sweeps", ctypes.c_int)]


class EprError(RuntimeError):
    pass


_lib = None


def load():
    """Loads libepr_hip.so (raises if it is missing: build it with __graft_entry__.build())."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise EprError(f"{LIB_PATH} is missing; build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, sz, ci = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_size_t, ctypes.c_int
    L.epr_default_params.argtypes = [ctypes.POINTER(Params)]
    L.epr_create.argtypes = [ci, ctypes.POINTER(Params), ctypes.POINTER(vp)]
    L.epr_destroy.argtypes = [vp]
    L.epr_last_error.argtypes = [vp]
    L.epr_last_error.restype = ctypes.c_char_p
    L.epr_set_samples.argtypes = [vp, sz, dp, dp]
    L.epr_build.argtypes = [vp]
    L.epr_init.argtypes = [vp, ci, dp, dp]
    L.epr_gs_sweeps.argtypes = [vp, ci, ci]
    L.epr_gs_colour.argtypes = [vp, ci, ci]
    L.epr_energies_noalloc.argtypes = [vp, ci, dp]
    L.epr_prolong_add.argtypes = [vp, ci]
    fp = ctypes.POINTER(ctypes.c_float)
    L.epr_unet_create.argtypes = [ci, ci, ci, ctypes.POINTER(vp)]
    L.epr_unet_destroy.argtypes = [vp]
    L.epr_unet_last_error.argtypes = [vp]
    L.epr_unet_last_error.restype = ctypes.c_char_p
    L.epr_unet_num_params.argtypes = [vp]
    L.epr_unet_num_params.restype = sz
    L.epr_unet_num_nodes.argtypes = [vp]
    L.epr_unet_num_nodes.restype = sz
    L.epr_unet_set_params.argtypes = [vp, fp]
    L.epr_unet_get_params.argtypes = [vp, fp]
    L.epr_unet_forward.argtypes = [vp, ci, fp, fp]
    L.epr_unet_loss_gradients.argtypes = [vp, ci, fp, fp, dp, fp]
    L.epr_unet_set_dataset.argtypes = [vp, sz, fp, fp]
    L.epr_unet_train.argtypes = [vp, ci, ci, ctypes.c_double, ctypes.c_uint, dp]
    L.epr_learned_features.argtypes = [vp, ci, fp]
    L.epr_init_learned.argtypes = [vp, vp, ci, dp]
    L.epr_solve_learned.argtypes = [vp, vp, dp]
    L.epr_solve_schedule.argtypes = [vp, vp, ctypes.POINTER(ci), dp]
    L.epr_sample_generate.argtypes = [ctypes.c_char_p, ctypes.c_uint, ci, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.POINTER(ci), ctypes.POINTER(sz), dp, dp, sz]
    L.epr_sample_generate.restype = ctypes.c_int
    L.epr_energies.argtypes = [vp, ci, dp]
    L.epr_solve.argtypes = [vp, dp]
    L.epr_num_functions.argtypes = [vp, ci]
    L.epr_num_functions.restype = sz
    L.epr_get_coeffs.argtypes = [vp, ci, dp, dp]
    L.epr_set_coeffs.argtypes = [vp, ci, dp, dp]
    L.epr_quadratic_fits.argtypes = [vp, ci, dp]
    L.epr_apply.argtypes = [vp, ci, ctypes.c_char, dp, dp]
    L.epr_get_array.argtypes = [vp, ctypes.c_char_p, dp, sz, ctypes.POINTER(sz)]
    L.epr_write_tree.argtypes = [vp, ctypes.c_char_p, ctypes.c_ulonglong]
    L.epr_write_grid.argtypes = [vp, ctypes.c_char_p]
    L.epr_write_density.argtypes = [vp, ctypes.c_char_p]
    L.epr_build_batch.argtypes = [sz, ctypes.POINTER(vp), ci]
    L.epr_solve_batch.argtypes = [sz, ctypes.POINTER(vp), dp]
    L.epr_density_grid.argtypes = [ctypes.c_char_p, ci, ci, ctypes.c_char_p]
    L.epr_autotune.argtypes = [vp, ci, ctypes.c_char_p, sz]
    L.epr_autotune_batch.argtypes = [sz, ctypes.POINTER(vp), ci, ctypes.c_char_p, sz]
    L.epr_tune_get.argtypes = [ctypes.c_char_p, sz, ctypes.POINTER(sz)]
    L.epr_tune_set.argtypes = [ctypes.c_char_p]
    L.epr_time_op.argtypes = [vp, ctypes.c_char_p, ci, ci, dp, dp]
    L.epr_time_batch_sweeps.argtypes = [sz, ctypes.POINTER(vp), ci, ci, ci, dp, ctypes.POINTER(ci), ctypes.POINTER(ctypes.c_longlong)]
    L.epr_time_gs_sweeps.argtypes = [vp, ci, ci, dp, ctypes.POINTER(ci)]
    L.epr_time_gs_sweeps.restype = ctypes.c_int
    L.epr_time_gs_sweeps_fp32.argtypes = [vp, ci, ci, dp, ctypes.POINTER(ci)]
    L.epr_time_gs_sweeps_fp32.restype = ctypes.c_int
    up = ctypes.POINTER(ctypes.c_uint)
    L.epr_subgrid_triangulation.argtypes = [ci, sz, up, ci, ctypes.POINTER(vp)]
    L.epr_subgrid_triangulation.restype = ctypes.c_int
    L.epr_level_set.argtypes = [ci, ci, ci, sz, dp, sz, up, dp, dp, ctypes.POINTER(vp)]
    L.epr_level_set.restype = ctypes.c_int
    L.epr_mesh_info.argtypes = [vp, ctypes.POINTER(sz), ctypes.POINTER(ci), ctypes.POINTER(sz), ctypes.POINTER(ci)]
    L.epr_iso_open.argtypes = [ctypes.c_char_p, ctypes.POINTER(vp)]
    L.epr_iso_last_error.argtypes = [vp]
    L.epr_iso_last_error.restype = ctypes.c_char_p
    L.epr_iso_info.argtypes = [vp, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(ci), dp]
    L.epr_iso_values.argtypes = [vp, dp]
    L.epr_iso_level_set_nodes.argtypes = [vp, ci, sz, up, ctypes.POINTER(sz)]
    L.epr_iso_extract.argtypes = [vp, ci, ctypes.POINTER(vp)]
    L.epr_iso_level_set.argtypes = [vp, ci, ctypes.POINTER(vp)]
    L.epr_iso_write_grid.argtypes = [vp, ctypes.c_char_p]
    L.epr_iso_seconds.argtypes = [vp, ci]
    L.epr_iso_seconds.restype = ctypes.c_double
    L.epr_iso_destroy.argtypes = [vp]
    L.epr_iso_read_density.argtypes = [vp, ctypes.c_char_p]
    L.epr_iso_extract_trimmed.argtypes = [vp, ci, ctypes.c_double, ctypes.POINTER(vp)]
    L.epr_iso_component_densities.argtypes = [vp, sz, dp, ctypes.POINTER(sz)]
    L.epr_indexer_hash.argtypes = [ci, ci]
    L.epr_orient.argtypes = [ci, sz, up]
    L.epr_indexer_hash.restype = ctypes.c_ulonglong
    L.epr_mesh_info.restype = ctypes.c_int
    L.epr_mesh_copy.argtypes = [vp, dp, up]
    L.epr_mesh_copy.restype = ctypes.c_int
    L.epr_mesh_destroy.argtypes = [vp]
    L.epr_mg_create.argtypes = [sz, up, up, dp, ci, ctypes.POINTER(sz), ctypes.POINTER(sz), ctypes.POINTER(up),
                                ctypes.POINTER(up), ctypes.POINTER(dp), ctypes.POINTER(vp)]
    L.epr_mg_create.restype = ctypes.c_int
    L.epr_mg_num_levels.argtypes = [vp]
    L.epr_mg_num_levels.restype = ctypes.c_int
    L.epr_mg_level_size.argtypes = [vp, ci]
    L.epr_mg_level_size.restype = sz
    L.epr_mg_set_colours.argtypes = [vp, ci, ci, up, up]
    L.epr_mg_set_colours.restype = ctypes.c_int
    L.epr_mg_vcycle.argtypes = [vp, dp, dp, ctypes.c_uint]
    L.epr_mg_vcycle.restype = ctypes.c_int
    L.epr_mg_relax.argtypes = [vp, ci, dp, dp, ctypes.c_uint]
    L.epr_mg_relax.restype = ctypes.c_int
    L.epr_mg_destroy.argtypes = [vp]
    L.epr_time_copy.argtypes = [sz, ci, dp]
    L.epr_time_copy.restype = ctypes.c_int
    L.epr_adapted_indexer_write.argtypes = [ci, ci, ci, ctypes.c_double, sz, dp, ctypes.c_char_p]
    L.epr_adapted_indexer_write.restype = ctypes.c_int
    L.epr_last_seconds.argtypes = [vp, ci]
    L.epr_last_seconds.restype = ctypes.c_double
    for name in ["epr_create", "epr_set_samples", "epr_build", "epr_init", "epr_gs_sweeps",
                 "epr_prolong_add", "epr_energies", "epr_solve", "epr_get_coeffs", "epr_set_coeffs",
                 "epr_quadratic_fits", "epr_apply", "epr_get_array", "epr_write_tree", "epr_write_grid",
                 "epr_gs_colour", "epr_energies_noalloc", "epr_unet_create", "epr_unet_set_params",
                 "epr_unet_get_params", "epr_unet_forward", "epr_unet_loss_gradients", "epr_unet_set_dataset",
                 "epr_unet_train", "epr_learned_features", "epr_init_learned", "epr_solve_learned",
                 "epr_autotune", "epr_autotune_batch", "epr_tune_get", "epr_tune_set", "epr_time_op",
                 "epr_time_batch_sweeps", "epr_solve_schedule"]:
        getattr(L, name).restype = ctypes.c_int
    _lib = L
    return L


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def default_params(**kw):
    p = Params()
    load().epr_default_params(ctypes.byref(p))
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError(f"unknown parameter {k}")
        setattr(p, k, v)
    return p
