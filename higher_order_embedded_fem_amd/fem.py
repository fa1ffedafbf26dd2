"""Python mirror of the reference's hot-path interface over libhoefem.so.

Names follow the reference: ``StableNeoHookeanMaterial`` (fenris/src/solid/materials.rs),
``ElasticityModel`` (the NodalModel / EmbeddedModel trait surface of fenris/src/solid/mod.rs),
``SchwarzPreconditioner`` (simulation_toolbox/src/fem/schwarz_precond.rs),
``conjugate_gradient`` (fenris/src/cg.rs) and ``NewtonCgIntegrator`` (the Backward-Euler
Newton-CG step of simulation_toolbox/src/fem/newton_cg.rs). Hot-path arrays are torch CUDA
(HIP) tensors; every call runs the HIP library, there is no CPU fallback.
"""

from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np
import torch

from . import _lib
from ._lib import check
from .scenes import Scene, lame_from_young_poisson


def _vec(x, dtype=np.float64) -> int:
    """Pointer of a caller vector: a device tensor (device pointer mode) or a contiguous numpy array
    (host pointer mode, hoefem_model_set_pointer_mode)."""
    if isinstance(x, np.ndarray):
        if x.dtype != dtype or not x.flags["C_CONTIGUOUS"]:
            raise ValueError(f"expected a contiguous {np.dtype(dtype)} array")
        return x.ctypes.data
    return _dev(x, torch.float64 if dtype == np.float64 else torch.float32)


def _opt(x):
    return None if x is None else C.byref(C.c_double(float(x)))


def _dev(t: torch.Tensor, dtype=torch.float64) -> int:
    if not t.is_cuda:
        raise ValueError("expected a device (cuda/hip) tensor")
    if t.dtype != dtype:
        raise ValueError(f"expected {dtype}")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


def _np(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)


class StableNeoHookeanMaterial:
    """StableNeoHookeanMaterial with Lame parameters from YoungPoisson (reference convention)."""

    kind = 0  # HOEFEM_STABLE_NEO_HOOKEAN
    threshold = 0.0

    def __init__(self, young: float = None, poisson: float = None, mu: float = None, lam: float = None):
        if mu is None:
            mu, lam = lame_from_young_poisson(young, poisson)
        self.mu, self.lam = float(mu), float(lam)

    def evaluate(self, F: torch.Tensor, stress: bool = True, tangent: bool = True):
        """F: (nq, 3, 3) device float64 (row-major torch layout). Returns (P, dP/dF[9x9 vec])."""
        L = _lib.load()
        nq = F.shape[0]
        Fc = F.transpose(1, 2).contiguous()  # column-major storage
        P = torch.empty_like(Fc) if stress else None
        D = torch.empty((nq, 9, 9), dtype=torch.float64, device=F.device) if tangent else None
        check(L.hoefem_material_eval_kind(self.kind, self.threshold, self.mu, self.lam, nq, Fc.data_ptr(),
                                          P.data_ptr() if P is not None else None,
                                          D.data_ptr() if D is not None else None))
        torch.cuda.synchronize()
        Pout = P.transpose(1, 2) if P is not None else None
        Dout = D.transpose(1, 2) if D is not None else None  # [q, row, col]
        return Pout, Dout

    def energy(self, F: torch.Tensor) -> torch.Tensor:
        """Strain energy density of the plain Stable Neo-Hookean material at each F (nq, 3, 3), device
        float64 (compute_strain_energy_density, fenris/src/solid/materials.rs:269-297)."""
        if self.kind != 0:
            raise ValueError("the energy density is provided for the plain Stable Neo-Hookean material")
        L = _lib.load()
        nq = F.shape[0]
        Fc = F.transpose(1, 2).contiguous()
        psi = torch.empty(nq, dtype=torch.float64, device=F.device)
        check(L.hoefem_material_energy(self.mu, self.lam, nq, Fc.data_ptr(), psi.data_ptr()))
        return psi


class InvertibleMaterial(StableNeoHookeanMaterial):
    """InvertibleMaterial::new(StableNeoHookeanMaterial, threshold) (fenris/src/solid/materials.rs:862-945):
    F is projected (signed singular values raised to the threshold) before the stress and the tangent.
    The reference's default Material is this with threshold 1e-6 (simulation_toolbox/src/fem/fe_model.rs:145-158)."""

    kind = 1  # HOEFEM_INVERTIBLE_STABLE_NEO_HOOKEAN

    def __init__(self, material: StableNeoHookeanMaterial, threshold: float = 1e-6):
        super().__init__(mu=material.mu, lam=material.lam)
        self.threshold = float(threshold)


class ProjectedStableNeoHookeanMaterial(StableNeoHookeanMaterial):
    """ProjectedStableNeoHookeanMaterial (fenris/src/solid/materials.rs:387-441): Stable Neo-Hookean stress,
    tangent projected onto the positive semi-definite cone."""

    kind = 2  # HOEFEM_PROJECTED_STABLE_NEO_HOOKEAN


def rotation_svd(F: torch.Tensor):
    """rotation_svd (fenris/src/util.rs:68-94) on (nq, 3, 3) device float64: returns U, S, Vt."""
    L = _lib.load()
    nq = F.shape[0]
    Fc = F.transpose(1, 2).contiguous()
    U, Vt = torch.empty_like(Fc), torch.empty_like(Fc)
    S = torch.empty((nq, 3), dtype=torch.float64, device=F.device)
    check(L.hoefem_rotation_svd(nq, Fc.data_ptr(), U.data_ptr(), S.data_ptr(), Vt.data_ptr()))
    return U.transpose(1, 2), S, Vt.transpose(1, 2)


def _quad_struct(q, keep):
    s = _lib.Quadrature()
    iw, ixi = _np(q.interior_w, np.float64), _np(q.interior_xi, np.float64)
    keep += [iw, ixi]
    s.interior_num_points = iw.size
    s.interior_weights, s.interior_points = iw.ctypes.data, ixi.ctypes.data
    if q.interface_offsets is not None:
        off, fw, fxi = _np(q.interface_offsets, np.int64), _np(q.interface_w, np.float64), _np(q.interface_xi, np.float64)
        keep += [off, fw, fxi]
        s.interface_offsets, s.interface_weights, s.interface_points = off.ctypes.data, fw.ctypes.data, fxi.ctypes.data
    return s


class ElasticityModel:
    """NodalModel / EmbeddedModel on the GPU (ElasticityModelParallel, fenris/src/solid/mod.rs:124-172)."""

    def __init__(self, scene: Scene):
        L = _lib.load()
        self.scene = scene
        keep = []
        d = _lib.ModelDesc()
        v, c = _np(scene.vertices, np.float64), _np(scene.connectivity, np.int32)
        keep += [v, c]
        d.element_type = scene.element_type
        d.num_vertices, d.vertices = v.shape[0], v.ctypes.data
        d.num_elements, d.connectivity = c.shape[0], c.ctypes.data
        d.num_interior = scene.num_interior
        d.mass_quadrature = _quad_struct(scene.mass_quadrature, keep)
        d.stiffness_quadrature = _quad_struct(scene.stiffness_quadrature, keep)
        d.elliptic_quadrature = _quad_struct(scene.elliptic_quadrature, keep)
        d.mass_regularization_factor = scene.mass_regularization
        h = C.c_void_p()
        check(L.hoefem_model_create(C.byref(d), C.byref(h)))
        self.h = h
        self.ndof = L.hoefem_model_ndof(h)
        self.nnz = L.hoefem_model_nnz(h)
        self.device = torch.device("cuda", torch.cuda.current_device())

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib._lib is not None:
            _lib._lib.hoefem_model_destroy(self.h)
            self.h = None

    def pattern(self):
        rp = np.zeros(self.ndof + 1, np.int32)
        col = np.zeros(self.nnz, np.int32)
        check(_lib.load().hoefem_model_pattern(self.h, rp.ctypes.data, col.ctypes.data))
        return rp, col

    def colors(self, p: int):
        L = _lib.load()
        nc = L.hoefem_model_num_colors(self.h, p)
        cnt = self.scene.num_interior if p == 0 else self.scene.num_elements - self.scene.num_interior
        off = np.zeros(nc + 1, np.int32)
        labels = np.zeros(max(cnt, 1), np.int32)
        check(L.hoefem_model_colors(self.h, p, off.ctypes.data, labels.ctypes.data))
        return off, labels[:cnt]

    def zeros_vector(self):
        return torch.zeros(self.ndof, dtype=torch.float64, device=self.device)

    def zeros_matrix(self):
        return torch.zeros(self.nnz, dtype=torch.float64, device=self.device)

    def set_pointer_mode(self, host: bool):
        """Hot-path calls on this model (and on the simulations built on it) take host numpy arrays
        (True) or device tensors (False, the default)."""
        check(_lib.load().hoefem_model_set_pointer_mode(self.h, 1 if host else 0))

    def set_material(self, material: StableNeoHookeanMaterial):
        """The elastic model used by this model's assembly and by simulations built on it."""
        check(_lib.load().hoefem_model_set_material(self.h, material.kind, material.threshold))

    def assemble_elastic_pseudo_forces_into(self, f: torch.Tensor, u: torch.Tensor, material: StableNeoHookeanMaterial):
        """f += f_int(u) (assemble_elastic_pseudo_forces_into_par)."""
        self.set_material(material)
        check(_lib.load().hoefem_model_forces_add(self.h, material.mu, material.lam, _dev(u), _dev(f)))
        return f

    def assemble_stiffness_into(self, values: torch.Tensor, u: torch.Tensor, material: StableNeoHookeanMaterial):
        """values += K(u) on the model pattern (assemble_transformed_stiffness_into_par); FP64 values,
        or FP32 values (float32 tensor) from FP32 element tangents."""
        self.set_material(material)
        prec = 1 if values.dtype == torch.float32 else 0
        check(_lib.load().hoefem_model_stiffness_add_prec(self.h, material.mu, material.lam, _dev(u),
                                                          _dev(values, values.dtype), prec))
        return values

    def assemble_mass(self, density: float) -> torch.Tensor:
        vals = self.zeros_matrix()
        check(_lib.load().hoefem_model_mass(self.h, density, _dev(vals)))
        return vals

    def apply_homogeneous_dirichlet_bc_csr(self, values: torch.Tensor, nodes) -> torch.Tensor:
        nodes = _np(nodes, np.int32)
        check(_lib.load().hoefem_model_dirichlet_csr(self.h, nodes.size, nodes.ctypes.data if nodes.size else None,
                                                     _dev(values)))
        return values

    def spmv(self, values: torch.Tensor, x: torch.Tensor, y: Optional[torch.Tensor] = None, alpha=1.0, beta=0.0):
        y = self.zeros_vector() if y is None else y
        check(_lib.load().hoefem_model_spmv(self.h, _dev(values), beta, _dev(y), alpha, _dev(x)))
        torch.cuda.synchronize()
        return y


class SchwarzPreconditioner:
    """SchwarzPreconditioner (simulation_toolbox/src/fem/newton_cg.rs:297-370) over subdomain node sets."""

    def __init__(self, model: ElasticityModel, offsets, nodes, untreated):
        L = _lib.load()
        self.model = model
        off, nd, un = _np(offsets, np.int64), _np(nodes, np.int32), _np(untreated, np.int32)
        self._keep = [off, nd, un]
        h = C.c_void_p()
        check(L.hoefem_schwarz_create(model.h, off.size - 1, off.ctypes.data, nd.ctypes.data if nd.size else None,
                                      un.size, un.ctypes.data if un.size else None, C.byref(h)))
        self.h = h
        self.nnz = L.hoefem_schwarz_nnz(h)

    @classmethod
    def from_scene(cls, model: ElasticityModel, scene: Scene):
        return cls(model, scene.subdomain_offsets, scene.subdomain_nodes, scene.untreated_nodes)

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib._lib is not None:
            _lib._lib.hoefem_schwarz_destroy(self.h)
            self.h = None

    def pattern(self):
        rp = np.zeros(self.model.ndof + 1, np.int32)
        col = np.zeros(self.nnz, np.int32)
        check(_lib.load().hoefem_schwarz_pattern(self.h, rp.ctypes.data, col.ctypes.data))
        return rp, col

    def update(self, J: torch.Tensor):
        check(_lib.load().hoefem_schwarz_update(self.h, _dev(J)))

    def values(self) -> torch.Tensor:
        out = torch.zeros(self.nnz, dtype=torch.float64, device=self.model.device)
        check(_lib.load().hoefem_schwarz_values(self.h, _dev(out)))
        return out

    def apply(self, x: torch.Tensor) -> torch.Tensor:
        y = self.model.zeros_vector()
        check(_lib.load().hoefem_schwarz_apply(self.h, _dev(x), _dev(y)))
        torch.cuda.synchronize()
        return y


def conjugate_gradient(model: ElasticityModel, A: torch.Tensor, b: torch.Tensor, x: torch.Tensor,
                       preconditioner: Optional[SchwarzPreconditioner] = None, tol: float = 1e-5,
                       max_iter: int = 500):
    """ConjugateGradient::solve_with_guess with RelativeResidualCriterion. Returns (iterations, status)."""
    it = C.c_int64()
    st = _lib.load().hoefem_pcg(model.h, _dev(A), preconditioner.h if preconditioner is not None else None, _dev(b),
                                _dev(x), tol, max_iter, C.byref(it))
    if st not in (0, 5, 6, 7):
        check(st)
    return it.value, st


class NewtonCgIntegrator:
    """One Backward-Euler Newton-CG step per call (NewtonCgIntegrator3d::run)."""

    def __init__(self, model: ElasticityModel, scene: Scene, newton_tolerance_factor: float = 1e-5,
                 cg_tolerance: float = 1e-5, cg_max_iter: int = 500, max_newton_iter: int = 100,
                 schwarz: bool = True, gravity_enabled: bool = True, mixed_precision: bool = False,
                 material: Optional[StableNeoHookeanMaterial] = None, mass_damping: float = 0.0,
                 stiffness_damping: float = 0.0):
        L = _lib.load()
        self.model, self.scene = model, scene
        mu, lam = lame_from_young_poisson(scene.young, scene.poisson)
        if material is not None:
            # the scene's elastic model (fe_model.rs:35-45); plain Stable Neo-Hookean otherwise
            mu, lam = material.mu, material.lam
            model.set_material(material)
        d = _lib.SimDesc()
        d.mu, d.lam, d.density = mu, lam, scene.density
        d.gravity, d.gravity_enabled = scene.gravity, 1 if gravity_enabled else 0
        dn = _np(scene.dirichlet_nodes, np.int32)
        so, sn, su = (_np(scene.subdomain_offsets, np.int64), _np(scene.subdomain_nodes, np.int32),
                      _np(scene.untreated_nodes, np.int32))
        self._keep = [dn, so, sn, su]
        d.num_dirichlet, d.dirichlet_nodes = dn.size, dn.ctypes.data if dn.size else None
        if schwarz and so.size > 1:
            d.num_subdomains, d.subdomain_offsets, d.subdomain_nodes = so.size - 1, so.ctypes.data, sn.ctypes.data
            d.num_untreated, d.untreated_nodes = su.size, su.ctypes.data if su.size else None
        d.newton_tolerance_factor = newton_tolerance_factor
        d.max_newton_iter = max_newton_iter
        d.cg_tolerance = cg_tolerance
        d.cg_max_iter = cg_max_iter
        d.mixed_precision = int(mixed_precision)  # False/0, True/1, 2 or 3 (see hoefem_sim_desc)
        # Material::mass_damping_coefficient / stiffness_damping_coefficient (fe_model.rs:131-136); 0 = None
        d.mass_damping_coefficient = float(mass_damping)
        d.stiffness_damping_coefficient = float(stiffness_damping)
        h = C.c_void_p()
        check(L.hoefem_sim_create(model.h, C.byref(d), C.byref(h)))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib._lib is not None:
            _lib._lib.hoefem_sim_destroy(self.h)
            self.h = None

    def autotune(self):
        """Pick the SpMV lanes-per-node of the Jacobian and Schwarz kernels by timing (opt-in:
        a different choice changes the reduction order, so runs stay bitwise reproducible only
        with a fixed choice)."""
        a, b = C.c_int32(), C.c_int32()
        t = np.zeros(9)
        check(_lib.load().hoefem_sim_autotune(self.h, C.byref(a), C.byref(b), t.ctypes.data))
        self.autotune_us = t.reshape(3, 3)  # [lanes of J in (8, 16, 32)][lanes of P in (8, 16, 32)]
        return a.value, b.value

    @property
    def representative_force(self) -> float:
        return _lib.load().hoefem_sim_representative_force(self.h)

    def mass(self) -> torch.Tensor:
        out = self.model.zeros_matrix()
        check(_lib.load().hoefem_sim_mass_values(self.h, _dev(out)))
        return out

    def damping(self) -> torch.Tensor:
        """The damping matrix of the last (sub)step on the model pattern (zeros without damping)."""
        out = torch.empty(self.model.nnz, dtype=torch.float64, device=self.model.device)
        check(_lib.load().hoefem_sim_damping_values(self.h, out.data_ptr()))
        return out

    def jacobian(self) -> torch.Tensor:
        out = self.model.zeros_matrix()
        check(_lib.load().hoefem_sim_jacobian_values(self.h, _dev(out)))
        return out

    def residual(self, t, dt, u, v, w) -> torch.Tensor:
        F = self.model.zeros_vector()
        check(_lib.load().hoefem_sim_residual(self.h, t, dt, _dev(u), _dev(v), _dev(w), _dev(F)))
        return F

    def set_dirichlet_values(self, u_values=None, v_values=None):
        """Prescribed values of the Dirichlet nodes (DirichletBoundaryConditions::apply_displacement_bcs /
        apply_velocity_bcs, simulation_toolbox/src/fem/bcs.rs:13-27): host arrays of three values per
        Dirichlet node in the scene's order, None for zeros (the Homogeneous condition)."""
        keep = [None if a is None else np.ascontiguousarray(a, dtype=np.float64).ravel() for a in (u_values, v_values)]
        m = 3 * len(self.scene.dirichlet_nodes)
        for a in keep:
            if a is not None and a.size != m:
                raise ValueError(f"expected {m} Dirichlet values, got {a.size}")
        check(_lib.load().hoefem_sim_set_dirichlet_values(self.h, *[None if a is None else a.ctypes.data
                                                                   for a in keep]))

    def step(self, t: float, dt: float, u, v) -> Dict:
        st = _lib.StepStats()
        check(_lib.load().hoefem_sim_step(self.h, t, dt, _vec(u), _vec(v), C.byref(st)))
        return {k: getattr(st, k) for k, _ in _lib.StepStats._fields_}

    # ---- DynamicSystem / DifferentiableDynamicSystem (hamilton2/src/dynamic_system/mutable.rs:14-76) of
    # NewtonCgDynamicSystem (simulation_toolbox/src/fem/newton_cg.rs:413-820). Vectors: device tensors,
    # or numpy arrays when the model is in host pointer mode.
    def update_damping(self, u):
        """compute_damping_matrix_into at the step's initial u (integrate_model, newton_cg.rs:256-265)."""
        check(_lib.load().hoefem_sim_update_damping(self.h, _vec(u)))

    def eval_f(self, f, t: float, u, v):
        """f = M g + f_int(u) - D v with Dirichlet rows zeroed (overwrites f)."""
        check(_lib.load().hoefem_sim_eval_f(self.h, t, _vec(u), _vec(v), _vec(f)))
        return f

    def apply_mass_matrix(self, y, x):
        """y += M x."""
        check(_lib.load().hoefem_sim_apply_mass_add(self.h, _vec(y), _vec(x)))
        return y

    def set_state(self, t: float, u, v):
        check(_lib.load().hoefem_sim_set_state(self.h, t, _vec(u), _vec(v)))

    def init_solve_jacobian_combination(self, alpha: Optional[float], beta: Optional[float]):
        check(_lib.load().hoefem_sim_init_solve_jacobian_combination(self.h, _opt(alpha), _opt(beta)))

    def solve_jacobian_combination(self, sol, rhs) -> int:
        """Solves (M + alpha df/du + beta df/dv) sol = rhs; returns the CG iteration count."""
        it = C.c_int64()
        check(_lib.load().hoefem_sim_solve_jacobian_combination(self.h, _vec(rhs), _vec(sol), C.byref(it)))
        return it.value

    def init_apply_jacobian_combination(self, alpha: Optional[float], beta: Optional[float],
                                        gamma: Optional[float]):
        check(_lib.load().hoefem_sim_init_apply_jacobian_combination(self.h, _opt(alpha), _opt(beta), _opt(gamma)))

    def apply_jacobian_combination(self, y, x):
        """y += (alpha df/du + beta df/dv + gamma M) x at the state of set_state."""
        check(_lib.load().hoefem_sim_apply_jacobian_combination(self.h, _vec(y), _vec(x)))
        return y


class LearnedCoarseCorrection:
    """Learned correction on top of the Schwarz preconditioner of a NewtonCgIntegrator
    (no counterpart in the reference; hoefem_learned_* in include/hoefem.h).

    Workflow: ``harvest(True)``, run plain steps (their CG solves store error / residual
    snapshots), ``train(...)``, ``enable(True)``; later solves use flexible PCG with the
    correction (under the plain path's rules: the CG iteration limit is tolerated and the Newton
    tolerance is unchanged) and fall back to the plain Schwarz-CG on a breakdown.
    The correction holds a reference to the integrator, which must outlive it.
    """

    PARAMS = None

    def __init__(self, integrator: "NewtonCgIntegrator", num_aggregates: int = 512, seed: int = 0):
        L = _lib.load()
        d = _lib.LearnedDesc()
        d.num_aggregates, d.seed = num_aggregates, seed
        h = C.c_void_p()
        check(L.hoefem_learned_create(integrator.h, C.byref(d), C.byref(h)))
        self.h = h
        self.integrator = integrator
        LearnedCoarseCorrection.PARAMS = int(L.hoefem_learned_param_count())

    def __del__(self):
        if getattr(self, "h", None) is not None and _lib._lib is not None:
            _lib._lib.hoefem_learned_destroy(self.h)
            self.h = None

    def harvest(self, enable: bool = True, capacity: int = 256, every: int = 32):
        check(_lib.load().hoefem_learned_harvest(self.h, 1 if enable else 0, capacity, every))

    def info(self) -> dict:
        s = _lib.LearnedStats()
        check(_lib.load().hoefem_learned_info(self.h, C.byref(s)))
        return {name: getattr(s, name) for name, _ in s._fields_}

    def train(self, steps: int = 300, learning_rate: float = 1e-3, batch: int = 8):
        a, b = C.c_double(), C.c_double()
        check(_lib.load().hoefem_learned_train(self.h, steps, learning_rate, batch, C.byref(a), C.byref(b)))
        return a.value, b.value

    def enable(self, on: bool = True):
        check(_lib.load().hoefem_learned_enable(self.h, 1 if on else 0))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.PARAMS, dtype=np.float32)
        check(_lib.load().hoefem_learned_get_params(self.h, p.ctypes.data))
        return p

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.PARAMS
        check(_lib.load().hoefem_learned_set_params(self.h, p.ctypes.data))

    def mlp_forward(self, x: torch.Tensor) -> torch.Tensor:
        x = x.contiguous()
        assert x.dtype == torch.float32 and x.dim() == 2 and x.shape[1] == 64 and x.is_cuda
        y = torch.empty_like(x)
        check(_lib.load().hoefem_learned_mlp_forward(self.h, x.shape[0], _dev(x, torch.float32), _dev(y, torch.float32)))
        return y

    def mlp_grad(self, x: torch.Tensor, dy: torch.Tensor) -> torch.Tensor:
        x, dy = x.contiguous(), dy.contiguous()
        assert x.dtype == torch.float32 and dy.dtype == torch.float32 and x.shape == dy.shape
        g = torch.zeros(self.PARAMS, dtype=torch.float32, device=x.device)
        check(_lib.load().hoefem_learned_mlp_grad(self.h, x.shape[0], _dev(x, torch.float32), _dev(dy, torch.float32),
                                                   _dev(g, torch.float32)))
        return g
