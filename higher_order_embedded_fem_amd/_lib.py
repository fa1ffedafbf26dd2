"""ctypes bindings of libhoefem.so (the C ABI declared in include/hoefem.h).

The HIP library is required: importing works without it, but every call raises
``HoefemLibraryMissing`` until ``build.py`` has produced ``libhoefem.so``. There is no CPU
fallback.
"""

from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhoefem.so")

EXPORTED = [
    "hoefem_last_error",
    "hoefem_nodes_per_element",
    "hoefem_lame_from_young_poisson",
    "hoefem_material_eval",
    "hoefem_material_energy",
    "hoefem_material_eval_kind",
    "hoefem_rotation_svd",
    "hoefem_mesh_reorder",
    "hoefem_model_create",
    "hoefem_model_destroy",
    "hoefem_model_set_stream",
    "hoefem_model_set_material",
    "hoefem_model_set_pointer_mode",
    "hoefem_model_ndof",
    "hoefem_model_nnz",
    "hoefem_model_pattern",
    "hoefem_model_num_colors",
    "hoefem_model_colors",
    "hoefem_model_forces_add",
    "hoefem_model_stiffness_add",
    "hoefem_model_stiffness_add_prec",
    "hoefem_model_mass",
    "hoefem_model_dirichlet_csr",
    "hoefem_model_spmv",
    "hoefem_schwarz_create",
    "hoefem_schwarz_destroy",
    "hoefem_schwarz_nnz",
    "hoefem_schwarz_pattern",
    "hoefem_schwarz_update",
    "hoefem_schwarz_values",
    "hoefem_schwarz_apply",
    "hoefem_pcg",
    "hoefem_pcg_noalloc",
    "hoefem_model_autotune",
    "hoefem_schwarz_autotune",
    "hoefem_sim_autotune",
    "hoefem_profile_enable",
    "hoefem_profile_read",
    "hoefem_profile_kernel",
    "hoefem_sim_create",
    "hoefem_sim_destroy",
    "hoefem_sim_representative_force",
    "hoefem_sim_mass_values",
    "hoefem_sim_jacobian_values",
    "hoefem_sim_damping_values",
    "hoefem_sim_residual",
    "hoefem_sim_step",
    "hoefem_sim_set_dirichlet_values",
    "hoefem_batch_create",
    "hoefem_batch_destroy",
    "hoefem_batch_step",
    "hoefem_sim_update_damping",
    "hoefem_sim_eval_f",
    "hoefem_sim_apply_mass_add",
    "hoefem_sim_set_state",
    "hoefem_sim_init_solve_jacobian_combination",
    "hoefem_sim_solve_jacobian_combination",
    "hoefem_sim_init_apply_jacobian_combination",
    "hoefem_sim_apply_jacobian_combination",
    "hoefem_learned_param_count",
    "hoefem_learned_create",
    "hoefem_learned_destroy",
    "hoefem_learned_harvest",
    "hoefem_learned_info",
    "hoefem_learned_train",
    "hoefem_learned_enable",
    "hoefem_learned_get_params",
    "hoefem_learned_set_params",
    "hoefem_learned_mlp_forward",
    "hoefem_learned_mlp_grad",
]


class HoefemLibraryMissing(RuntimeError):
    pass


class HoefemError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"hoefem status {code}: {message}")
        self.code = code


STATUS = {
    0: "OK",
    1: "INVALID_ARGUMENT",
    2: "DEVICE",
    3: "SINGULAR",
    4: "PATTERN",
    5: "CG_MAX_ITERATIONS",
    6: "CG_INDEFINITE_OPERATOR",
    7: "CG_INDEFINITE_PRECONDITIONER",
    8: "NEWTON_MAX_ITERATIONS",
    9: "LINE_SEARCH",
    10: "SUBSTEPPING",
    11: "NOT_FINITE",
    12: "JACOBIAN",
}


class Quadrature(C.Structure):
    _fields_ = [
        ("interior_num_points", C.c_int32),
        ("interior_weights", C.c_void_p),
        ("interior_points", C.c_void_p),
        ("interface_offsets", C.c_void_p),
        ("interface_weights", C.c_void_p),
        ("interface_points", C.c_void_p),
    ]


class ModelDesc(C.Structure):
    _fields_ = [
        ("element_type", C.c_int32),
        ("num_vertices", C.c_int64),
        ("vertices", C.c_void_p),
        ("num_elements", C.c_int64),
        ("connectivity", C.c_void_p),
        ("num_interior", C.c_int64),
        ("mass_quadrature", Quadrature),
        ("stiffness_quadrature", Quadrature),
        ("elliptic_quadrature", Quadrature),
        ("mass_regularization_factor", C.c_double),
    ]


class LearnedDesc(C.Structure):
    _fields_ = [("num_aggregates", C.c_int32), ("seed", C.c_uint64)]


class LearnedStats(C.Structure):
    _fields_ = [
        ("num_aggregates", C.c_int32),
        ("max_aggregate_dofs", C.c_int32),
        ("harvested_pairs", C.c_int64),
        ("stored_pairs", C.c_int64),
        ("trained", C.c_int32),
        ("enabled", C.c_int32),
        ("learned_solves", C.c_int64),
        ("fallbacks", C.c_int64),
        ("learned_iterations", C.c_int64),
        ("last_status", C.c_int32),
    ]


class SimDesc(C.Structure):
    _fields_ = [
        ("mu", C.c_double),
        ("lam", C.c_double),
        ("density", C.c_double),
        ("gravity", C.c_double),
        ("gravity_enabled", C.c_int32),
        ("num_dirichlet", C.c_int64),
        ("dirichlet_nodes", C.c_void_p),
        ("num_subdomains", C.c_int64),
        ("subdomain_offsets", C.c_void_p),
        ("subdomain_nodes", C.c_void_p),
        ("num_untreated", C.c_int64),
        ("untreated_nodes", C.c_void_p),
        ("newton_tolerance_factor", C.c_double),
        ("max_newton_iter", C.c_int64),
        ("cg_tolerance", C.c_double),
        ("cg_max_iter", C.c_int64),
        ("mixed_precision", C.c_int32),
        ("mass_damping_coefficient", C.c_double),
        ("stiffness_damping_coefficient", C.c_double),
    ]


class StepStats(C.Structure):
    _fields_ = [
        ("newton_iterations", C.c_int64),
        ("cg_iterations", C.c_int64),
        ("substeps", C.c_int64),
        ("residual_evaluations", C.c_int64),
        ("final_residual", C.c_double),
        ("tolerance", C.c_double),
    ]


_lib = None

vp, i64, i32, f64 = C.c_void_p, C.c_int64, C.c_int32, C.c_double


def load(path: str = LIB_PATH):
    """Load libhoefem.so (raises HoefemLibraryMissing when it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise HoefemLibraryMissing(
            f"{path} not found: build it with `python higher_order_embedded_fem_amd/build.py` (no CPU fallback exists)")
    L = C.CDLL(path)
    L.hoefem_last_error.restype = C.c_char_p
    L.hoefem_nodes_per_element.argtypes = [i32]
    L.hoefem_nodes_per_element.restype = i32
    L.hoefem_lame_from_young_poisson.argtypes = [f64, f64, C.POINTER(f64), C.POINTER(f64)]
    L.hoefem_lame_from_young_poisson.restype = None
    L.hoefem_material_eval.argtypes = [f64, f64, i64, vp, vp, vp]
    L.hoefem_material_eval_kind.argtypes = [i32, f64, f64, f64, i64, vp, vp, vp]
    L.hoefem_material_energy.argtypes = [f64, f64, i64, vp, vp]
    L.hoefem_rotation_svd.argtypes = [i64, vp, vp, vp, vp]
    L.hoefem_mesh_reorder.argtypes = [i32, i64, i64, vp, vp, vp]
    L.hoefem_model_create.argtypes = [C.POINTER(ModelDesc), C.POINTER(vp)]
    L.hoefem_model_destroy.argtypes = [vp]
    L.hoefem_model_destroy.restype = None
    L.hoefem_model_set_stream.argtypes = [vp, vp]
    L.hoefem_model_set_material.argtypes = [vp, i32, f64]
    L.hoefem_model_ndof.argtypes = [vp]
    L.hoefem_model_ndof.restype = i64
    L.hoefem_model_nnz.argtypes = [vp]
    L.hoefem_model_nnz.restype = i64
    L.hoefem_model_pattern.argtypes = [vp, vp, vp]
    L.hoefem_model_num_colors.argtypes = [vp, i32]
    L.hoefem_model_num_colors.restype = i32
    L.hoefem_model_colors.argtypes = [vp, i32, vp, vp]
    L.hoefem_model_forces_add.argtypes = [vp, f64, f64, vp, vp]
    L.hoefem_model_stiffness_add.argtypes = [vp, f64, f64, vp, vp]
    L.hoefem_model_stiffness_add_prec.argtypes = [vp, f64, f64, vp, vp, i32]
    L.hoefem_model_set_pointer_mode.argtypes = [vp, i32]
    L.hoefem_model_mass.argtypes = [vp, f64, vp]
    L.hoefem_model_dirichlet_csr.argtypes = [vp, i64, vp, vp]
    L.hoefem_model_spmv.argtypes = [vp, vp, f64, vp, f64, vp]
    L.hoefem_schwarz_create.argtypes = [vp, i64, vp, vp, i64, vp, C.POINTER(vp)]
    L.hoefem_schwarz_destroy.argtypes = [vp]
    L.hoefem_schwarz_destroy.restype = None
    L.hoefem_schwarz_nnz.argtypes = [vp]
    L.hoefem_schwarz_nnz.restype = i64
    L.hoefem_schwarz_pattern.argtypes = [vp, vp, vp]
    L.hoefem_schwarz_update.argtypes = [vp, vp]
    L.hoefem_schwarz_values.argtypes = [vp, vp]
    L.hoefem_schwarz_apply.argtypes = [vp, vp, vp]
    L.hoefem_pcg.argtypes = [vp, vp, vp, vp, vp, f64, i64, C.POINTER(i64)]
    L.hoefem_pcg_noalloc.argtypes = [vp, vp, vp, vp, vp, f64, i64, C.POINTER(i64)]
    L.hoefem_model_autotune.argtypes = [vp, C.POINTER(i32), vp]
    L.hoefem_schwarz_autotune.argtypes = [vp, C.POINTER(i32), vp]
    L.hoefem_sim_autotune.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), vp]
    L.hoefem_profile_enable.argtypes = [vp, i32]
    L.hoefem_profile_read.argtypes = [vp, C.POINTER(i64), C.POINTER(f64), C.POINTER(i64), C.POINTER(f64)]
    L.hoefem_profile_kernel.argtypes = [vp, i32, C.POINTER(i64), C.POINTER(f64)]
    L.hoefem_sim_create.argtypes = [vp, C.POINTER(SimDesc), C.POINTER(vp)]
    L.hoefem_sim_destroy.argtypes = [vp]
    L.hoefem_sim_destroy.restype = None
    L.hoefem_sim_representative_force.argtypes = [vp]
    L.hoefem_sim_representative_force.restype = f64
    L.hoefem_sim_mass_values.argtypes = [vp, vp]
    L.hoefem_sim_jacobian_values.argtypes = [vp, vp]
    L.hoefem_sim_damping_values.argtypes = [vp, vp]
    L.hoefem_sim_residual.argtypes = [vp, f64, f64, vp, vp, vp, vp]
    L.hoefem_sim_step.argtypes = [vp, f64, f64, vp, vp, C.POINTER(StepStats)]
    L.hoefem_sim_set_dirichlet_values.argtypes = [vp, vp, vp]
    L.hoefem_batch_create.argtypes = [C.POINTER(vp), i32, C.POINTER(vp)]
    L.hoefem_batch_destroy.argtypes = [vp]
    L.hoefem_batch_destroy.restype = None
    L.hoefem_batch_step.argtypes = [vp, C.POINTER(f64), C.POINTER(vp), C.POINTER(vp), C.POINTER(StepStats),
                                    C.POINTER(f64), C.POINTER(i32)]
    L.hoefem_sim_update_damping.argtypes = [vp, vp]
    L.hoefem_sim_eval_f.argtypes = [vp, f64, vp, vp, vp]
    L.hoefem_sim_apply_mass_add.argtypes = [vp, vp, vp]
    L.hoefem_sim_set_state.argtypes = [vp, f64, vp, vp]
    L.hoefem_sim_init_solve_jacobian_combination.argtypes = [vp, C.POINTER(f64), C.POINTER(f64)]
    L.hoefem_sim_solve_jacobian_combination.argtypes = [vp, vp, vp, C.POINTER(i64)]
    L.hoefem_sim_init_apply_jacobian_combination.argtypes = [vp, C.POINTER(f64), C.POINTER(f64), C.POINTER(f64)]
    L.hoefem_sim_apply_jacobian_combination.argtypes = [vp, vp, vp]
    L.hoefem_learned_param_count.argtypes = []
    L.hoefem_learned_param_count.restype = i64
    L.hoefem_learned_create.argtypes = [vp, C.POINTER(LearnedDesc), C.POINTER(vp)]
    L.hoefem_learned_destroy.argtypes = [vp]
    L.hoefem_learned_harvest.argtypes = [vp, i32, i32, i32]
    L.hoefem_learned_info.argtypes = [vp, C.POINTER(LearnedStats)]
    L.hoefem_learned_train.argtypes = [vp, i32, f64, i32, C.POINTER(f64), C.POINTER(f64)]
    L.hoefem_learned_enable.argtypes = [vp, i32]
    L.hoefem_learned_get_params.argtypes = [vp, vp]
    L.hoefem_learned_set_params.argtypes = [vp, vp]
    L.hoefem_learned_mlp_forward.argtypes = [vp, i64, vp, vp]
    L.hoefem_learned_mlp_grad.argtypes = [vp, i64, vp, vp, vp]
    for name in EXPORTED:
        if name not in ("hoefem_last_error", "hoefem_nodes_per_element", "hoefem_lame_from_young_poisson",
                        "hoefem_model_destroy", "hoefem_model_ndof", "hoefem_model_nnz", "hoefem_model_num_colors",
                        "hoefem_schwarz_destroy", "hoefem_schwarz_nnz", "hoefem_sim_destroy", "hoefem_batch_destroy",
                        "hoefem_sim_representative_force", "hoefem_learned_param_count"):
            getattr(L, name).restype = C.c_int
    _lib = L
    return L


def check(code: int):
    if code != 0:
        raise HoefemError(code, load().hoefem_last_error().decode())
