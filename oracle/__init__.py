"""ctypes bindings of the CPU FP64 oracle (oracle/oracle.cpp).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always to verify the HIP library, never in its place.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

f64p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i32p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
i64p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")


class OrcQuadrature(C.Structure):
    _fields_ = [
        ("interior_nq", C.c_int32),
        ("interior_w", C.c_void_p),
        ("interior_xi", C.c_void_p),
        ("interface_offsets", C.c_void_p),
        ("interface_w", C.c_void_p),
        ("interface_xi", C.c_void_p),
    ]


class OrcModelDesc(C.Structure):
    _fields_ = [
        ("element_type", C.c_int32),
        ("num_vertices", C.c_int64),
        ("vertices", C.c_void_p),
        ("num_elements", C.c_int64),
        ("connectivity", C.c_void_p),
        ("num_interior", C.c_int64),
        ("mass_quadrature", OrcQuadrature),
        ("stiffness_quadrature", OrcQuadrature),
        ("elliptic_quadrature", OrcQuadrature),
        ("mass_regularization_factor", C.c_double),
    ]


class OrcSimDesc(C.Structure):
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
        ("num_threads", C.c_int32),
        ("mass_damping_coefficient", C.c_double),
        ("stiffness_damping_coefficient", C.c_double),
    ]


class OrcStepStats(C.Structure):
    _fields_ = [
        ("newton_iterations", C.c_int64),
        ("cg_iterations", C.c_int64),
        ("substeps", C.c_int64),
        ("line_search_evals", C.c_int64),
        ("final_residual", C.c_double),
        ("tolerance", C.c_double),
        ("t_forces", C.c_double),
        ("t_stiffness", C.c_double),
        ("t_jacobian", C.c_double),
        ("t_precond", C.c_double),
        ("t_cg", C.c_double),
    ]


EVAL_FN = C.CFUNCTYPE(None, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))
SOLVE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double))
_D = C.POINTER(C.c_double)
DYN_EVAL_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, _D, _D, _D)
DYN_MASS_FN = C.CFUNCTYPE(None, C.c_void_p, _D, _D)
DYN_SOLVE_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_double, _D, _D, C.c_double, C.c_double, _D, _D)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        L.orc_last_error.restype = C.c_char_p
        L.orc_snh_stress.argtypes = [C.c_double, C.c_double, C.c_int64, f64p, f64p]
        L.orc_snh_dpdf.argtypes = [C.c_double, C.c_double, C.c_int64, f64p, f64p]
        L.orc_snh_energy.argtypes = [C.c_double, C.c_double, C.c_int64, f64p, f64p]
        L.orc_snh_contract_with.argtypes = [C.c_double, C.c_double, f64p, f64p, f64p, f64p]
        L.orc_snh_contract_multiple.argtypes = [C.c_double, C.c_double, f64p, C.c_int, f64p, f64p]
        L.orc_svd3.argtypes = [C.c_int64, f64p, f64p, f64p, f64p]
        L.orc_rotation_svd.argtypes = [C.c_int64, f64p, f64p, f64p, f64p]
        L.orc_project_deformation_gradient.argtypes = [C.c_double, C.c_int64, f64p, f64p]
        L.orc_snh_dpdf_eigen.argtypes = [C.c_double, C.c_double, f64p, f64p, f64p]
        L.orc_material_stress.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int64, f64p, f64p]
        L.orc_material_dpdf.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int64, f64p, f64p]
        L.orc_material_contract_multiple.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, f64p, C.c_int,
                                                     f64p, f64p]
        L.orc_model_set_material.argtypes = [C.c_void_p, C.c_int, C.c_double]
        L.orc_element_basis.argtypes = [C.c_int, C.c_int64, f64p, f64p]
        L.orc_element_gradients.argtypes = [C.c_int, C.c_int64, f64p, f64p]
        L.orc_element_reference_jacobian.argtypes = [C.c_int, f64p, f64p, f64p]
        L.orc_model_create.argtypes = [C.POINTER(OrcModelDesc), C.POINTER(C.c_void_p)]
        L.orc_model_destroy.argtypes = [C.c_void_p]
        L.orc_model_ndof.argtypes = [C.c_void_p]
        L.orc_model_ndof.restype = C.c_int64
        L.orc_model_num_colors.argtypes = [C.c_void_p, C.c_int]
        L.orc_model_colors.argtypes = [C.c_void_p, C.c_int, i32p, i32p]
        L.orc_model_nnz.argtypes = [C.c_void_p]
        L.orc_model_nnz.restype = C.c_int64
        L.orc_model_pattern.argtypes = [C.c_void_p, i32p, i32p]
        L.orc_model_assemble_mass.argtypes = [C.c_void_p, C.c_double, f64p]
        L.orc_model_forces_add.argtypes = [C.c_void_p, C.c_double, C.c_double, f64p, f64p]
        L.orc_model_stiffness_add.argtypes = [C.c_void_p, C.c_double, C.c_double, f64p, f64p]
        L.orc_model_stiffness_add_projected.argtypes = [C.c_void_p, C.c_double, C.c_double, f64p, f64p]
        L.orc_psd_project.argtypes = [C.c_int, f64p, f64p]
        L.orc_cuthill_mckee.argtypes = [C.c_int64, i32p, i32p, i64p]
        L.orc_mesh_reorder.argtypes = [C.c_int, C.c_int64, C.c_int64, i32p, i64p, i64p]
        L.orc_sim_damping_values.argtypes = [C.c_void_p, f64p]
        L.orc_greedy_coloring.argtypes = [C.c_int64, i64p, i32p, i32p, C.POINTER(C.c_int32)]
        L.orc_coo_pattern.argtypes = [C.c_int64, C.c_int64, i64p, i32p, C.c_int, C.POINTER(C.c_int64), C.c_void_p,
                                      C.c_void_p]
        L.orc_spmv.argtypes = [C.c_int64, i32p, i32p, f64p, C.c_double, f64p, C.c_double, f64p]
        L.orc_dirichlet_csr.argtypes = [C.c_int64, i32p, i32p, f64p, C.c_int64, i32p, C.c_int]
        L.orc_dirichlet_rhs.argtypes = [f64p, C.c_int64, i32p]
        L.orc_schwarz_create.argtypes = [C.c_int64, C.c_int64, i64p, i32p, C.c_int64, i32p, C.POINTER(C.c_void_p)]
        L.orc_schwarz_destroy.argtypes = [C.c_void_p]
        L.orc_schwarz_nnz.argtypes = [C.c_void_p]
        L.orc_schwarz_nnz.restype = C.c_int64
        L.orc_schwarz_pattern.argtypes = [C.c_void_p, i32p, i32p]
        L.orc_schwarz_update.argtypes = [C.c_void_p, i32p, i32p, f64p]
        L.orc_schwarz_values.argtypes = [C.c_void_p, f64p]
        L.orc_invert_block.argtypes = [C.c_int, f64p, f64p]
        f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
        L.orc_mlp_forward.argtypes = [f32p, C.c_int64, f32p, C.c_int, f64p]
        L.orc_mlp_grad.argtypes = [f32p, C.c_int64, f32p, f32p, C.c_int, f64p]
        L.orc_symmetric_eigen.argtypes = [C.c_int, f64p, f64p, f64p]
        L.orc_pcg_csr.argtypes = [C.c_int64, i32p, i32p, f64p, C.c_void_p, C.c_void_p, C.c_void_p, f64p, f64p,
                                  C.c_double, C.c_int64, C.POINTER(C.c_int64)]
        L.orc_backward_euler_step.argtypes = [C.c_int64, f64p, f64p, C.c_double, C.c_double, C.c_void_p, DYN_EVAL_FN,
                                              DYN_MASS_FN, DYN_SOLVE_FN, C.c_double, C.c_int64, C.POINTER(C.c_int64)]
        L.orc_integrate_with_substeps.argtypes = [C.c_int64, f64p, f64p, C.c_double, C.c_double, C.c_void_p,
                                                  DYN_EVAL_FN, DYN_MASS_FN, DYN_SOLVE_FN, C.c_double, C.c_int64,
                                                  C.POINTER(C.c_int64), C.POINTER(C.c_int64)]
        L.orc_newton.argtypes = [C.c_int64, f64p, C.c_void_p, EVAL_FN, SOLVE_FN, C.c_double, C.c_int64, C.c_int,
                                 C.POINTER(C.c_int64)]
        L.orc_sim_create.argtypes = [C.c_void_p, C.POINTER(OrcSimDesc), C.POINTER(C.c_void_p)]
        L.orc_sim_destroy.argtypes = [C.c_void_p]
        L.orc_sim_representative_force.argtypes = [C.c_void_p]
        L.orc_sim_representative_force.restype = C.c_double
        L.orc_sim_set_dirichlet_values.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_sim_mass_values.argtypes = [C.c_void_p, f64p]
        L.orc_sim_jacobian_values.argtypes = [C.c_void_p, f64p]
        L.orc_sim_step.argtypes = [C.c_void_p, C.c_double, C.c_double, f64p, f64p, C.POINTER(OrcStepStats)]
        L.orc_sim_residual.argtypes = [C.c_void_p, C.c_double, C.c_double, f64p, f64p, f64p, f64p]
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"oracle status {code}: {msg}")
        self.code = code


def check(code):
    if code != 0:
        raise OracleError(code, lib().orc_last_error().decode())


def _ptr(a):
    return None if a is None else a.ctypes.data


def lame(young, poisson):
    mu = 0.5 * young / (1.0 + poisson)
    return mu, 0.5 * mu * poisson / (1.0 - 2.0 * poisson)


def snh_stress(mu, lam, F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 3, 3)
    Fc = np.ascontiguousarray(F.transpose(0, 2, 1))  # column-major storage
    P = np.empty_like(Fc)
    check(lib().orc_snh_stress(mu, lam, Fc.shape[0], Fc.ravel(), P.reshape(-1)))
    return P.transpose(0, 2, 1)


def snh_dpdf(mu, lam, F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 3, 3)
    Fc = np.ascontiguousarray(F.transpose(0, 2, 1))
    out = np.empty((F.shape[0], 81))
    check(lib().orc_snh_dpdf(mu, lam, F.shape[0], Fc.ravel(), out.reshape(-1)))
    return out.reshape(-1, 9, 9).transpose(0, 2, 1)  # [q, row, col]


def snh_energy(mu, lam, F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 3, 3)
    Fc = np.ascontiguousarray(F.transpose(0, 2, 1))
    out = np.empty(F.shape[0])
    check(lib().orc_snh_energy(mu, lam, F.shape[0], Fc.ravel(), out))
    return out


def snh_contract_with(mu, lam, F, a, b):
    Fc = np.ascontiguousarray(np.asarray(F, dtype=np.float64).T)
    out = np.empty(9)
    check(lib().orc_snh_contract_with(mu, lam, Fc.ravel(), np.ascontiguousarray(a, dtype=np.float64),
                                      np.ascontiguousarray(b, dtype=np.float64), out))
    return out.reshape(3, 3).T


def snh_contract_multiple(mu, lam, F, a, out=None):
    """a: (3, n) columns a_I; returns/adds into the (3n, 3n) matrix."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[1]
    Fc = np.ascontiguousarray(np.asarray(F, dtype=np.float64).T)
    if out is None:
        out = np.zeros((3 * n, 3 * n))
    buf = np.ascontiguousarray(out.T)
    check(lib().orc_snh_contract_multiple(mu, lam, Fc.ravel(), n, np.ascontiguousarray(a.T).ravel(), buf.ravel()))
    return buf.T.copy()


# elastic model kinds (simulation_toolbox/src/fem/fe_model.rs:35-45)
MAT_SNH, MAT_INVERTIBLE_SNH, MAT_PROJECTED_SNH = 0, 1, 2


def _colmajor(F):
    F = np.ascontiguousarray(F, dtype=np.float64).reshape(-1, 3, 3)
    return F.shape[0], np.ascontiguousarray(F.transpose(0, 2, 1)).ravel()


def _rowmajor(buf, nq):
    return buf.reshape(nq, 3, 3).transpose(0, 2, 1).copy()


def cuthill_mckee(rowptr, col):
    """cuthill_mckee (fenris/src/reorder.rs:169-235): permutation, target index -> source index."""
    rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
    col = np.ascontiguousarray(col, dtype=np.int32)
    perm = np.zeros(len(rowptr) - 1, np.int64)
    check(lib().orc_cuthill_mckee(len(rowptr) - 1, rowptr, col, perm))
    return perm


def mesh_reorder(connectivity, num_vertices):
    """reorder_mesh_par (fenris/src/reorder.rs:52-88): (vertex permutation, element permutation)."""
    conn = np.ascontiguousarray(connectivity, dtype=np.int32)
    vp = np.zeros(num_vertices, np.int64)
    ep = np.zeros(conn.shape[0], np.int64)
    check(lib().orc_mesh_reorder(conn.shape[1], num_vertices, conn.shape[0], conn.ravel(), vp, ep))
    return vp, ep


def psd_project(A):
    """DefaultSemidefiniteProjection (fenris/src/assembly.rs:552-564) of a symmetric matrix."""
    A = np.asarray(A, dtype=np.float64)
    n = A.shape[0]
    out = np.empty(n * n)
    check(lib().orc_psd_project(n, np.ascontiguousarray(A.T).ravel(), out))
    return out.reshape(n, n).T.copy()


def svd3(F):
    """Oracle SVD (sigma >= 0, U, V orthogonal): returns U, S, V with F = U diag(S) V^T."""
    nq, Fc = _colmajor(F)
    U, S, V = np.empty(9 * nq), np.empty(3 * nq), np.empty(9 * nq)
    check(lib().orc_svd3(nq, Fc, U, S, V))
    return _rowmajor(U, nq), S.reshape(nq, 3), _rowmajor(V, nq)


def rotation_svd(F):
    """rotation_svd (fenris/src/util.rs:68-94): returns U, S, Vt with U, Vt rotations."""
    nq, Fc = _colmajor(F)
    U, S, Vt = np.empty(9 * nq), np.empty(3 * nq), np.empty(9 * nq)
    check(lib().orc_rotation_svd(nq, Fc, U, S, Vt))
    return _rowmajor(U, nq), S.reshape(nq, 3), _rowmajor(Vt, nq)


def project_deformation_gradient(F, threshold):
    """project_deformation_gradient (fenris/src/solid/materials.rs:876-915)."""
    nq, Fc = _colmajor(F)
    out = np.empty(9 * nq)
    check(lib().orc_project_deformation_gradient(threshold, nq, Fc, out))
    return _rowmajor(out, nq)


def snh_dpdf_eigen(mu, lam, F):
    """build_stable_neohookean_dp_df_eigen_3d (materials.rs:561-687): eigenvalues (9,) and
    eigenvectors as the columns of a (9, 9) matrix, vec() ordering."""
    Fc = np.ascontiguousarray(np.asarray(F, dtype=np.float64).T).ravel()
    ev, Q = np.empty(9), np.empty(81)
    check(lib().orc_snh_dpdf_eigen(mu, lam, Fc, ev, Q))
    return ev, Q.reshape(9, 9).T.copy()


def material_stress(kind, threshold, mu, lam, F):
    nq, Fc = _colmajor(F)
    out = np.empty(9 * nq)
    check(lib().orc_material_stress(kind, threshold, mu, lam, nq, Fc, out))
    return _rowmajor(out, nq)


def material_dpdf(kind, threshold, mu, lam, F):
    nq, Fc = _colmajor(F)
    out = np.empty(81 * nq)
    check(lib().orc_material_dpdf(kind, threshold, mu, lam, nq, Fc, out))
    return out.reshape(-1, 9, 9).transpose(0, 2, 1).copy()


def material_contract_multiple(kind, threshold, mu, lam, F, a):
    """The (3n, 3n) contraction of the elastic model kind with the columns a_I of a (3, n)."""
    a = np.asarray(a, dtype=np.float64)
    n = a.shape[1]
    Fc = np.ascontiguousarray(np.asarray(F, dtype=np.float64).T).ravel()
    buf = np.zeros(9 * n * n)
    check(lib().orc_material_contract_multiple(kind, threshold, mu, lam, Fc, n, np.ascontiguousarray(a.T).ravel(),
                                               buf))
    return buf.reshape(3 * n, 3 * n).T.copy()


def element_basis(etype, xi):
    xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, 3)
    n = lib().orc_nodes_per_element(etype)
    out = np.empty((xi.shape[0], n))
    check(lib().orc_element_basis(etype, xi.shape[0], xi.ravel(), out.reshape(-1)))
    return out


def element_gradients(etype, xi):
    """returns (npts, 3, n): d phi_I / d xi_k at [p, k, I]"""
    xi = np.ascontiguousarray(xi, dtype=np.float64).reshape(-1, 3)
    n = lib().orc_nodes_per_element(etype)
    out = np.empty((xi.shape[0], n, 3))
    check(lib().orc_element_gradients(etype, xi.shape[0], xi.ravel(), out.reshape(-1)))
    return out.transpose(0, 2, 1)


def _quad_struct(q, keep):
    keep.extend([q.interior_w, q.interior_xi, q.interface_offsets, q.interface_w, q.interface_xi])
    s = OrcQuadrature()
    s.interior_nq = len(q.interior_w)
    s.interior_w = _ptr(np.ascontiguousarray(q.interior_w, dtype=np.float64))
    keep.append(np.ascontiguousarray(q.interior_w, dtype=np.float64))
    iw = np.ascontiguousarray(q.interior_w, dtype=np.float64)
    ixi = np.ascontiguousarray(q.interior_xi, dtype=np.float64)
    keep.extend([iw, ixi])
    s.interior_w, s.interior_xi = _ptr(iw), _ptr(ixi)
    if q.interface_offsets is not None:
        off = np.ascontiguousarray(q.interface_offsets, dtype=np.int64)
        fw = np.ascontiguousarray(q.interface_w, dtype=np.float64)
        fxi = np.ascontiguousarray(q.interface_xi, dtype=np.float64)
        keep.extend([off, fw, fxi])
        s.interface_offsets, s.interface_w, s.interface_xi = _ptr(off), _ptr(fw), _ptr(fxi)
    return s


class Model:
    """Oracle model (NodalModel / EmbeddedModel semantics)."""

    def __init__(self, scene):
        self.scene = scene
        keep = []
        d = OrcModelDesc()
        d.element_type = scene.element_type
        v = np.ascontiguousarray(scene.vertices, dtype=np.float64)
        c = np.ascontiguousarray(scene.connectivity, dtype=np.int32)
        keep += [v, c]
        d.num_vertices = v.shape[0]
        d.vertices = _ptr(v)
        d.num_elements = c.shape[0]
        d.connectivity = _ptr(c)
        d.num_interior = scene.num_interior
        d.mass_quadrature = _quad_struct(scene.mass_quadrature, keep)
        d.stiffness_quadrature = _quad_struct(scene.stiffness_quadrature, keep)
        d.elliptic_quadrature = _quad_struct(scene.elliptic_quadrature, keep)
        d.mass_regularization_factor = scene.mass_regularization
        h = C.c_void_p()
        check(lib().orc_model_create(C.byref(d), C.byref(h)))
        self.h = h
        self._keep = keep
        self.ndof = lib().orc_model_ndof(h)
        self.nnz = lib().orc_model_nnz(h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_model_destroy(self.h)
            self.h = None

    def colors(self, p):
        L = lib()
        nc = L.orc_model_num_colors(self.h, p)
        off = np.zeros(nc + 1, np.int32)
        cnt = self.scene.num_interior if p == 0 else self.scene.num_elements - self.scene.num_interior
        labels = np.zeros(max(cnt, 1), np.int32)
        check(L.orc_model_colors(self.h, p, off, labels))
        return off, labels[:cnt]

    def pattern(self):
        rp = np.zeros(self.ndof + 1, np.int32)
        col = np.zeros(self.nnz, np.int32)
        check(lib().orc_model_pattern(self.h, rp, col))
        return rp, col

    def mass(self, density=None):
        vals = np.zeros(self.nnz)
        check(lib().orc_model_assemble_mass(self.h, self.scene.density if density is None else density, vals))
        return vals

    def set_material(self, kind, threshold=0.0):
        check(lib().orc_model_set_material(self.h, kind, threshold))

    def forces_add(self, mu, lam, u, f):
        check(lib().orc_model_forces_add(self.h, mu, lam, np.ascontiguousarray(u, dtype=np.float64), f))
        return f

    def stiffness_add_projected(self, mu, lam, u, vals):
        """values += the stiffness with every element matrix projected onto the PSD cone."""
        check(lib().orc_model_stiffness_add_projected(self.h, mu, lam, np.ascontiguousarray(u, dtype=np.float64),
                                                      vals))
        return vals

    def stiffness_add(self, mu, lam, u, vals):
        check(lib().orc_model_stiffness_add(self.h, mu, lam, np.ascontiguousarray(u, dtype=np.float64), vals))
        return vals


def greedy_coloring(offsets, nodes):
    offsets = np.ascontiguousarray(offsets, dtype=np.int64)
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    nsets = offsets.size - 1
    out = np.zeros(max(nsets, 1), np.int32)
    nc = C.c_int32()
    check(lib().orc_greedy_coloring(nsets, offsets, nodes, out, C.byref(nc)))
    return out[:nsets], nc.value


def spmv(rp, col, vals, x, y=None, alpha=1.0, beta=0.0):
    n = rp.size - 1
    y = np.zeros(n) if y is None else y
    lib().orc_spmv(n, rp, col, vals, beta, y, alpha, np.ascontiguousarray(x, dtype=np.float64))
    return y


def dirichlet_csr(rp, col, vals, nodes, sdim=3):
    nodes = np.ascontiguousarray(nodes, dtype=np.int32)
    if nodes.size == 0:
        nodes = np.zeros(1, np.int32)
        lib().orc_dirichlet_csr(rp.size - 1, rp, col, vals, 0, nodes, sdim)
    else:
        lib().orc_dirichlet_csr(rp.size - 1, rp, col, vals, nodes.size, nodes, sdim)
    return vals


def coo_pattern(num_nodes, sets, sdim):
    """CsrParAssembler::assemble_pattern on the given node sets (list of lists)."""
    off = np.zeros(len(sets) + 1, np.int64)
    for k, st in enumerate(sets):
        off[k + 1] = off[k] + len(st)
    nodes = np.ascontiguousarray(np.concatenate([np.asarray(st, np.int32) for st in sets] + [np.zeros(0, np.int32)]),
                                 dtype=np.int32)
    if nodes.size == 0:
        nodes = np.zeros(1, np.int32)
    nnz = C.c_int64()
    check(lib().orc_coo_pattern(num_nodes, len(sets), off, nodes, sdim, C.byref(nnz), None, None))
    rp = np.zeros(num_nodes * sdim + 1, np.int32)
    col = np.zeros(max(nnz.value, 1), np.int32)
    check(lib().orc_coo_pattern(num_nodes, len(sets), off, nodes, sdim, C.byref(nnz), rp.ctypes.data, col.ctypes.data))
    return rp, col[:nnz.value]


def backward_euler_step(u, v, t0, dt, eval_f, apply_mass, solve_jac, tol, max_newton_iter=100):
    """orc_backward_euler_step driven by Python callbacks; returns (iterations, status)."""
    n = len(u)

    def _ev(ctx, t, up, vp, fp):
        fv = np.ctypeslib.as_array(fp, shape=(n,))
        fv[:] = eval_f(t, np.ctypeslib.as_array(up, shape=(n,)).copy(), np.ctypeslib.as_array(vp, shape=(n,)).copy())
        return 0

    def _ms(ctx, yp, xp):
        yv = np.ctypeslib.as_array(yp, shape=(n,))
        yv += apply_mass(np.ctypeslib.as_array(xp, shape=(n,)).copy())

    def _so(ctx, t, up, vp, a, b, rp, sp):
        sv = np.ctypeslib.as_array(sp, shape=(n,))
        sv[:] = solve_jac(t, np.ctypeslib.as_array(up, shape=(n,)).copy(), np.ctypeslib.as_array(vp, shape=(n,)).copy(),
                          a, b, np.ctypeslib.as_array(rp, shape=(n,)).copy())
        return 0

    e, m, s = DYN_EVAL_FN(_ev), DYN_MASS_FN(_ms), DYN_SOLVE_FN(_so)
    it = C.c_int64()
    st = lib().orc_backward_euler_step(n, u, v, t0, dt, None, e, m, s, tol, max_newton_iter, C.byref(it))
    return it.value, st


def _dyn_callbacks(n, eval_f, apply_mass, solve_jac):
    """ctypes callbacks of a Python dynamic system; a solve_jac that raises reports a Jacobian error."""

    def _ev(ctx, t, up, vp, fp):
        fv = np.ctypeslib.as_array(fp, shape=(n,))
        fv[:] = eval_f(t, np.ctypeslib.as_array(up, shape=(n,)).copy(), np.ctypeslib.as_array(vp, shape=(n,)).copy())
        return 0

    def _ms(ctx, yp, xp):
        yv = np.ctypeslib.as_array(yp, shape=(n,))
        yv += apply_mass(np.ctypeslib.as_array(xp, shape=(n,)).copy())

    def _so(ctx, t, up, vp, a, b, rp, sp):
        sv = np.ctypeslib.as_array(sp, shape=(n,))
        try:
            sv[:] = solve_jac(t, np.ctypeslib.as_array(up, shape=(n,)).copy(),
                              np.ctypeslib.as_array(vp, shape=(n,)).copy(), a, b,
                              np.ctypeslib.as_array(rp, shape=(n,)).copy())
        except Exception:
            return 12
        return 0

    return DYN_EVAL_FN(_ev), DYN_MASS_FN(_ms), DYN_SOLVE_FN(_so)


def integrate_with_substeps(u, v, t0, dt, eval_f, apply_mass, solve_jac, tol, max_newton_iter=100):
    """orc_integrate_with_substeps driven by Python callbacks (u, v updated in place on success);
    returns (substeps, newton_iterations, status)."""
    e, m, s = _dyn_callbacks(len(u), eval_f, apply_mass, solve_jac)
    ns, it = C.c_int64(), C.c_int64()
    st = lib().orc_integrate_with_substeps(len(u), u, v, t0, dt, None, e, m, s, tol, max_newton_iter, C.byref(ns),
                                           C.byref(it))
    return ns.value, it.value, st


class Schwarz:
    def __init__(self, num_nodes, offsets, nodes, untreated):
        self._keep = [np.ascontiguousarray(offsets, dtype=np.int64), np.ascontiguousarray(nodes, dtype=np.int32),
                      np.ascontiguousarray(untreated, dtype=np.int32)]
        off, nd, un = self._keep
        if un.size == 0:
            un = np.zeros(1, np.int32)
            self._keep.append(un)
        h = C.c_void_p()
        check(lib().orc_schwarz_create(num_nodes, off.size - 1, off, nd, len(untreated), un, C.byref(h)))
        self.h = h
        self.ndof = 3 * num_nodes
        self.nnz = lib().orc_schwarz_nnz(h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_schwarz_destroy(self.h)
            self.h = None

    def pattern(self):
        rp = np.zeros(self.ndof + 1, np.int32)
        col = np.zeros(self.nnz, np.int32)
        check(lib().orc_schwarz_pattern(self.h, rp, col))
        return rp, col

    def update(self, rp, col, vals):
        check(lib().orc_schwarz_update(self.h, rp, col, vals))
        out = np.zeros(self.nnz)
        check(lib().orc_schwarz_values(self.h, out))
        return out


def invert_block(B):
    B = np.ascontiguousarray(np.asarray(B, dtype=np.float64).T)
    out = np.empty_like(B)
    check(lib().orc_invert_block(B.shape[0], B.ravel(), out.ravel()))
    return out.T.copy()


MLP_PARAMS = 128 * 64 + 128 + 128 * 128 + 128 + 64 * 128 + 64


def mlp_forward(params, x, emulate_bf16=True):
    """The learned correction's network on samples x[n][64] (FP64 arithmetic; with
    emulate_bf16, BF16 rounding of inputs, weights and hidden activations as on the GPU)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.zeros((x.shape[0], 64))
    check(lib().orc_mlp_forward(p, x.shape[0], x.ravel(), 1 if emulate_bf16 else 0, y.ravel()))
    return y


def mlp_grad(params, x, dy, emulate_bf16=True):
    """d(sum_s dy[s] . y[s]) / d params."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    x = np.ascontiguousarray(x, dtype=np.float32)
    dy = np.ascontiguousarray(dy, dtype=np.float32)
    g = np.zeros(MLP_PARAMS)
    check(lib().orc_mlp_grad(p, x.shape[0], x.ravel(), dy.ravel(), 1 if emulate_bf16 else 0, g))
    return g


def pcg(rp, col, vals, b, x0=None, prec=None, tol=1e-5, max_iter=500):
    """Returns (x, iterations, status)."""
    n = rp.size - 1
    x = np.zeros(n) if x0 is None else np.array(x0, dtype=np.float64)
    it = C.c_int64()
    keep = []
    if prec is not None:
        prp, pcol, pv = [np.ascontiguousarray(a) for a in prec]
        keep = [prp, pcol, pv]
        args = (prp.ctypes.data, pcol.ctypes.data, pv.ctypes.data)
    else:
        args = (None, None, None)
    st = lib().orc_pcg_csr(n, rp, col, vals, *args, np.ascontiguousarray(b, dtype=np.float64), x, tol, max_iter,
                           C.byref(it))
    del keep
    return x, it.value, st


def newton(n, x0, eval_fn, solve_fn, tol, max_iter, line_search=True):
    """Drive orc_newton with Python callbacks; returns (x, iterations, status)."""
    x = np.array(x0, dtype=np.float64)

    def _eval(ctx, xp, fp):
        xv = np.ctypeslib.as_array(xp, shape=(n,))
        fv = np.ctypeslib.as_array(fp, shape=(n,))
        fv[:] = eval_fn(xv.copy())

    def _solve(ctx, xp, rp, sp):
        xv = np.ctypeslib.as_array(xp, shape=(n,))
        rv = np.ctypeslib.as_array(rp, shape=(n,))
        sv = np.ctypeslib.as_array(sp, shape=(n,))
        try:
            sv[:] = solve_fn(xv.copy(), rv.copy())
            return 0
        except Exception:
            return 12

    ef, sf = EVAL_FN(_eval), SOLVE_FN(_solve)
    it = C.c_int64()
    st = lib().orc_newton(n, x, None, ef, sf, tol, max_iter, 1 if line_search else 0, C.byref(it))
    return x, it.value, st


class Sim:
    """Oracle FE dynamic system: NewtonCgIntegrator3d semantics on the CPU."""

    def __init__(self, model, newton_tolerance_factor=1e-5, cg_tolerance=1e-5, cg_max_iter=500,
                 max_newton_iter=100, schwarz=True, num_threads=0, mass_damping=0.0, stiffness_damping=0.0):
        sc = model.scene
        mu, lam = lame(sc.young, sc.poisson)
        d = OrcSimDesc()
        d.mu, d.lam, d.density = mu, lam, sc.density
        d.gravity = sc.gravity
        d.gravity_enabled = 1
        keep = [np.ascontiguousarray(sc.dirichlet_nodes, dtype=np.int32),
                np.ascontiguousarray(sc.subdomain_offsets, dtype=np.int64),
                np.ascontiguousarray(sc.subdomain_nodes, dtype=np.int32),
                np.ascontiguousarray(sc.untreated_nodes, dtype=np.int32)]
        d.num_dirichlet = keep[0].size
        d.dirichlet_nodes = _ptr(keep[0])
        if schwarz:
            d.num_subdomains = keep[1].size - 1
            d.subdomain_offsets = _ptr(keep[1])
            d.subdomain_nodes = _ptr(keep[2])
            d.num_untreated = keep[3].size
            d.untreated_nodes = _ptr(keep[3])
        d.newton_tolerance_factor = newton_tolerance_factor
        d.max_newton_iter = max_newton_iter
        d.cg_tolerance = cg_tolerance
        d.cg_max_iter = cg_max_iter
        d.num_threads = num_threads
        d.mass_damping_coefficient = float(mass_damping)
        d.stiffness_damping_coefficient = float(stiffness_damping)
        h = C.c_void_p()
        check(lib().orc_sim_create(model.h, C.byref(d), C.byref(h)))
        self.h = h
        self.model = model
        self._keep = keep
        self.ndof = model.ndof

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_sim_destroy(self.h)
            self.h = None

    @property
    def representative_force(self):
        return lib().orc_sim_representative_force(self.h)

    def mass(self):
        out = np.zeros(self.model.nnz)
        check(lib().orc_sim_mass_values(self.h, out))
        return out

    def damping(self):
        out = np.zeros(self.model.nnz)
        check(lib().orc_sim_damping_values(self.h, out))
        return out

    def jacobian(self):
        out = np.zeros(self.model.nnz)
        check(lib().orc_sim_jacobian_values(self.h, out))
        return out

    def set_dirichlet_values(self, u_values=None, v_values=None):
        """values of the Dirichlet condition for the sim's Dirichlet nodes ([num_dirichlet, 3] or flat;
        None: zeros)"""
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (u_values, v_values)]
        m = 3 * len(self.model.scene.dirichlet_nodes)
        for a in keep:
            if a is not None and a.size != m:
                raise ValueError(f"expected {m} Dirichlet values, got {a.size}")
        check(lib().orc_sim_set_dirichlet_values(self.h, *[None if a is None else a.ctypes.data for a in keep]))

    def step(self, t, dt, u, v):
        st = OrcStepStats()
        check(lib().orc_sim_step(self.h, t, dt, u, v, C.byref(st)))
        return {k: getattr(st, k) for k, _ in OrcStepStats._fields_}

    def residual(self, t, dt, u, v, w):
        F = np.zeros(self.ndof)
        check(lib().orc_sim_residual(self.h, t, dt, u, v, np.ascontiguousarray(w, dtype=np.float64), F))
        return F
