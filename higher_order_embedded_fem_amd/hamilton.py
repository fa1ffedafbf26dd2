"""Host-side Backward-Euler / Newton driver over the dynamic-system entry points.

This is the caller side of the hamilton2 boundary: what ``hamilton2::backward_euler_step``
(hamilton2/src/integrators/euler.rs:155-199) with ``newton_line_search`` and
``BacktrackingLineSearch`` (hamilton2/src/newton.rs:76-131, :171-246) do with a
``DifferentiableDynamicSystem``, written against ``NewtonCgIntegrator``'s ``eval_f``,
``apply_mass_matrix``, ``set_state``, ``init_solve_jacobian_combination`` and
``solve_jacobian_combination`` with host numpy vectors (the model in host pointer mode), and
``integrate_model`` (simulation_toolbox/src/fem/newton_cg.rs:230-295) around it.

The vector updates round every operation separately, as nalgebra does; the library's own
``hoefem_sim_step`` does the same on the device, so both produce the same iterates.
"""

from __future__ import annotations

from typing import Callable, Dict

import numpy as np


class NewtonError(RuntimeError):
    """NewtonError (hamilton2/src/newton.rs:26-33): kind is "max_iterations", "jacobian" or
    "line_search"."""

    def __init__(self, kind: str, message: str):
        super().__init__(message)
        self.kind = kind


# first trial step lengths, then a geometric tail 0.25 * 0.25^k (newton.rs:212-218)
_FIRST_ALPHAS = (1.0, 0.75, 0.5)


def _alphas():
    yield from _FIRST_ALPHAS
    a = 0.25
    while True:
        yield a
        a *= 0.25


def backtracking_step(eval_into: Callable[[np.ndarray, np.ndarray], None], f: np.ndarray, x: np.ndarray,
                      p: np.ndarray, stats: Dict) -> float:
    """One backtracking line search along p on g = |F|^2 / 2 with the Armijo test
    g(x + a p) <= (1 - c a) g(x), c = 1e-4; fails below a = 1e-6. x and f are updated in place
    (x moves by the difference of consecutive trial lengths). Returns the accepted length."""
    c, alpha_min = 1e-4, 1e-6
    g0 = 0.5 * float(f @ f)
    prev = 0.0
    for alpha in _alphas():
        x += (alpha - prev) * p
        eval_into(f, x)
        stats["evaluations"] += 1
        stats["alphas"].append(alpha)
        if 0.5 * float(f @ f) <= (1.0 - c * alpha) * g0:
            return alpha
        if alpha < alpha_min:
            raise NewtonError("line_search", f"step length {alpha} below the minimum {alpha_min}")
        prev = alpha
    raise AssertionError("unreachable")


def newton_line_search(eval_into, solve_into, x: np.ndarray, tolerance: float, max_iterations,
                       stats: Dict) -> int:
    """Newton's method with backtracking on F(x) = 0: solve J(-dx) = F, then line search along dx.
    Raises NewtonError when max_iterations is reached (None: no limit) or a solve fails."""
    f = np.zeros_like(x)
    eval_into(f, x)
    stats["evaluations"] += 1
    it = 0
    while np.sqrt(float(f @ f)) > tolerance:
        if max_iterations is not None and it == max_iterations:
            raise NewtonError("max_iterations", f"Newton reached {it} iterations")
        minus_dx = np.zeros_like(x)
        try:
            solve_into(minus_dx, x, f)
        except Exception as e:  # the system's solve error becomes NewtonError::JacobianError
            raise NewtonError("jacobian", str(e)) from e
        p = minus_dx * -1.0
        backtracking_step(eval_into, f, x, p, stats)
        it += 1
    stats["final_residual"] = float(np.sqrt(float(f @ f)))
    return it


def backward_euler_step(system, u: np.ndarray, v: np.ndarray, t0: float, dt: float, tolerance: float,
                        max_newton_iter=100, stats: Dict = None) -> int:
    """Backward Euler on M dv/dt = f(t, u, v), du/dt = v: with u_p = u + dt v, solve
    F(w) = M w - dt f(t + dt, u_p + dt w, v + w) = 0 by Newton (tolerance dt * tolerance), then
    v += w, u += dt v. u, v (host arrays) are updated in place; returns the Newton iterations."""
    stats = stats if stats is not None else {}
    stats.setdefault("evaluations", 0)
    stats.setdefault("alphas", [])
    stats.setdefault("cg_iterations", 0)
    t_next = t0 + dt
    v_n = v.copy()
    u_p = u + v * dt
    fbuf = np.zeros_like(u)

    def eval_into(F, w):
        uu = u_p + w * dt
        vv = v_n + w
        system.eval_f(fbuf, t_next, uu, vv)
        F[:] = fbuf * (dt * -1.0)
        system.apply_mass_matrix(F, w)

    def solve_into(sol, w, rhs):
        uu = u_p + w * dt
        vv = v_n + w
        system.set_state(t_next, uu, vv)
        system.init_solve_jacobian_combination(-dt * dt, -dt)
        stats["cg_iterations"] += system.solve_jacobian_combination(sol, np.ascontiguousarray(rhs))

    w = np.zeros_like(u)
    it = newton_line_search(eval_into, solve_into, w, dt * tolerance, max_newton_iter, stats)
    v += w
    u += v * dt
    return it


def integrate_model(system, u: np.ndarray, v: np.ndarray, t: float, dt: float, dirichlet_nodes,
                    tolerance_factor: float = 1e-5, max_newton_iter=100, stats: Dict = None) -> int:
    """One step of integrate_model (simulation_toolbox/src/fem/newton_cg.rs:230-295): damping at the
    step's initial u, backward_euler_step with tolerance factor * |M a_rep|, then the Dirichlet
    values (zero) on u and v. The Dirichlet values of the initial state are applied first, as the
    library's step does."""
    nodes = np.asarray(dirichlet_nodes, dtype=np.int64)
    dofs = (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()
    u[dofs] = 0.0
    v[dofs] = 0.0
    system.update_damping(u)
    it = backward_euler_step(system, u, v, t, dt, tolerance_factor * system.representative_force,
                             max_newton_iter, stats)
    u[dofs] = 0.0
    v[dofs] = 0.0
    return it
