"""Oracle checks of the damping path (CPU).

compute_damping_matrix_into (simulation_toolbox/src/fem/system_assembly.rs:145-194): D = a_M M +
a_K K+(u_n), K+ assembled from element matrices projected onto the positive semi-definite cone
(DefaultSemidefiniteProjection, fenris/src/assembly.rs:552-564); eval_f adds -D v
(newton_cg.rs:469-471) and the Newton matrix is M + dt D + dt^2 K (system_assembly.rs:197-235 with
BackwardEulerImplicitFunction's (alpha, beta) = (-dt^2, -dt), hamilton2/src/integrators/euler.rs:122-140).

Pinned by: the projection against numpy's eigh; a one-element projected stiffness against the eigh
projection of that element's assembled stiffness; a rigid translation under mass damping, where
Backward Euler has the closed form v1 = v0 / (1 + a_M dt).
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from higher_order_embedded_fem_amd import scenes


def _dense(rp, col, vals):
    n = len(rp) - 1
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = vals[rp[i]:rp[i + 1]]
    return A


def _eigh_projection(A):
    w, V = np.linalg.eigh(0.5 * (A + A.T))
    return (V * np.maximum(w, 0.0)) @ V.T


@pytest.mark.parametrize("n", [6, 30, 61])
def test_psd_projection_matches_eigh(n):
    rng = np.random.default_rng(n)
    B = rng.standard_normal((n, n))
    A = 0.5 * (B + B.T)
    got = O.psd_project(A)
    ref = _eigh_projection(A)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(A).max()
    # positive semi-definite input: unchanged up to rounding; negative definite: zero
    P = B @ B.T
    assert np.abs(O.psd_project(P) - P).max() <= 1e-12 * np.abs(P).max()
    assert np.abs(O.psd_project(-P)).max() <= 1e-12 * np.abs(P).max()


def _one_element_scene(element_type):
    sc = scenes.cube_scene(1, element_type)
    verts, conn = scenes.keep_cells(sc.vertices, sc.connectivity, np.array([2]))
    so, sn, su = scenes._per_element_subdomains(conn)
    return dataclasses.replace(sc, vertices=np.ascontiguousarray(verts), connectivity=np.ascontiguousarray(conn),
                               num_interior=1, dirichlet_nodes=np.zeros(0, np.int32), subdomain_offsets=so,
                               subdomain_nodes=sn, untreated_nodes=su)


@pytest.mark.parametrize("etype", [scenes.TET10, scenes.TET20])
def test_projected_stiffness_is_element_projection(etype):
    sc = _one_element_scene(etype)
    om = O.Model(sc)
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    rng = np.random.default_rng(5)
    u = 0.25 * rng.standard_normal(sc.ndof)  # large enough for an indefinite element tangent
    rp, col = om.pattern()
    K = _dense(rp, col, om.stiffness_add(mu, lam, u, np.zeros(om.nnz)))
    w = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert w.min() < -1e-6 * np.abs(w).max()  # the state exercises the clamp
    Kp = _dense(rp, col, om.stiffness_add_projected(mu, lam, u, np.zeros(om.nnz)))
    assert np.abs(Kp - _eigh_projection(K)).max() <= 1e-10 * np.abs(K).max()


@pytest.mark.parametrize("a_m, a_k", [(0.4, 0.0), (0.4, 0.05), (0.0, 0.05)])
def test_rigid_translation_closed_form(a_m, a_k):
    """No supports, no gravity, uniform initial velocity: the elastic force and K v vanish, so
    M (v1 - v0) / dt = -a_M M v1 and v1 = v0 / (1 + a_M dt)."""
    sc = dataclasses.replace(scenes.cube_scene(2, scenes.TET10), dirichlet_nodes=np.zeros(0, np.int32),
                             gravity=0.0)
    om = O.Model(sc)
    sim = O.Sim(om, newton_tolerance_factor=1e-10, cg_tolerance=1e-13, cg_max_iter=5000, mass_damping=a_m,
                stiffness_damping=a_k)
    v0 = np.tile([0.3, -0.2, 0.5], sc.num_vertices)
    u = np.zeros(sc.ndof)
    v = v0.copy()
    sim.step(0.0, sc.dt, u, v)
    v1 = v0 / (1.0 + a_m * sc.dt)
    assert np.abs(v - v1).max() <= 1e-9 * np.abs(v0).max()
    assert np.abs(u - sc.dt * v1).max() <= 1e-9 * sc.dt * np.abs(v0).max()
    D = sim.damping()
    assert np.abs(D).max() > 0


def test_damping_dissipates():
    """With damping the kinetic energy after a few steps from a deforming initial velocity is
    lower than without (a sign error in -D v or dt D would raise it)."""
    sc = dataclasses.replace(scenes.cube_scene(2, scenes.TET10), gravity=0.0)
    X = sc.vertices
    v0 = (np.stack([X[:, 2], -X[:, 2], 0.5 * X[:, 2]], 1) * 2.0).ravel()
    energies = []
    for a_m, a_k in [(0.0, 0.0), (0.5, 0.05)]:
        om = O.Model(sc)
        sim = O.Sim(om, newton_tolerance_factor=1e-9, cg_tolerance=1e-12, cg_max_iter=5000, mass_damping=a_m,
                    stiffness_damping=a_k)
        u, v = np.zeros(sc.ndof), v0.copy()
        for k in range(3):
            sim.step(k * sc.dt, sc.dt, u, v)
        M = sim.mass()
        rp, col = om.pattern()
        Md = _dense(rp, col, M)
        energies.append(0.5 * v @ Md @ v)
    assert energies[1] < energies[0]


def test_damping_off_matches_plain():
    sc = scenes.cube_scene(2, scenes.TET10)
    runs = []
    for kw in ({}, {"mass_damping": 0.0, "stiffness_damping": 0.0}):
        om = O.Model(sc)
        sim = O.Sim(om, **kw)
        u, v = np.zeros(sc.ndof), np.zeros(sc.ndof)
        sim.step(0.0, sc.dt, u, v)
        runs.append(u)
        assert np.abs(sim.damping()).max() == 0.0
    assert np.array_equal(runs[0], runs[1])
