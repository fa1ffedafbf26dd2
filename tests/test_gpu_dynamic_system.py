"""The hamilton2 dynamic-system boundary on the GPU, and the Newton / substepping / tolerance paths
against the oracle.

* A Backward-Euler step composed on the host from hoefem_sim_eval_f, hoefem_sim_apply_mass_add,
  hoefem_sim_set_state, hoefem_sim_init_solve_jacobian_combination and
  hoefem_sim_solve_jacobian_combination (higher_order_embedded_fem_amd/hamilton.py, the caller side
  of hamilton2/src/integrators/euler.rs:155-199) reproduces hoefem_sim_step bit for bit.
* eval_f, apply_mass and apply_jacobian_combination against the oracle and the assembled Jacobian.
* The FP32 stiffness variant against the FP64 oracle; host pointer mode against device mode.
* Substepping (simulation_toolbox/src/fem/newton_cg.rs:146-228) forced by a one-iteration Newton
  limit, line-search backtracking (hamilton2/src/newton.rs:171-246) on a large step, and the
  representative force (simulation_toolbox/src/fem/fe_model.rs:423-431): GPU and oracle take the
  same decisions (substeps, Newton iterations, residual evaluations) and reach the same state.
"""

import math

import numpy as np
import pytest

import oracle as O
from higher_order_embedded_fem_amd import hamilton, scenes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def dev(x):
    return torch.as_tensor(np.ascontiguousarray(x), dtype=torch.float64, device="cuda")


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.fixture(scope="module")
def fem():
    from higher_order_embedded_fem_amd import fem as F

    return F


@pytest.mark.parametrize("name,make", [("cyl8_hex20", lambda: scenes.cylinder_scene(8, scenes.HEX20, 8)),
                                       ("cube3_tet10", lambda: scenes.cube_scene(3, scenes.TET10))])
def test_composed_backward_euler_matches_sim_step_bitwise(fem, name, make):
    sc = make()
    # the library's own step, device vectors
    gm_d = fem.ElasticityModel(sc)
    sim_d = fem.NewtonCgIntegrator(gm_d, sc, newton_tolerance_factor=1e-6)
    # the same system driven from the host through the dynamic-system entry points (host pointers)
    gm_h = fem.ElasticityModel(sc)
    gm_h.set_pointer_mode(True)
    sim_h = fem.NewtonCgIntegrator(gm_h, sc, newton_tolerance_factor=1e-6)
    assert sim_h.representative_force == sim_d.representative_force
    u_d = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
    v_d = torch.zeros_like(u_d)
    u_h, v_h = np.zeros(sc.ndof), np.zeros(sc.ndof)
    for k in range(3):
        st = sim_d.step(k * sc.dt, sc.dt, u_d, v_d)
        stats = {}
        it = hamilton.integrate_model(sim_h, u_h, v_h, k * sc.dt, sc.dt, sc.dirichlet_nodes, tolerance_factor=1e-6,
                                      stats=stats)
        print(f"{name} step {k}: newton {st['newton_iterations']} / {it}, cg {st['cg_iterations']} / "
              f"{stats['cg_iterations']}, evaluations {st['residual_evaluations']} / {stats['evaluations']}")
        assert st["substeps"] == 1
        assert it == st["newton_iterations"]
        assert stats["cg_iterations"] == st["cg_iterations"]
        assert stats["evaluations"] == st["residual_evaluations"]
        assert np.array_equal(host(u_d), u_h), (k, rel(host(u_d), u_h))
        assert np.array_equal(host(v_d), v_h), (k, rel(host(v_d), v_h))
    assert np.linalg.norm(u_h) > 0


def test_eval_f_apply_mass_and_jacobian_combination(fem):
    sc = scenes.cube_scene(3, scenes.TET10)
    om = O.Model(sc)
    osim = O.Sim(om)
    gm = fem.ElasticityModel(sc)
    sim = fem.NewtonCgIntegrator(gm, sc, cg_tolerance=1e-13, cg_max_iter=5000)
    rng = np.random.default_rng(3)
    u = 0.01 * rng.standard_normal(sc.ndof)
    dt = sc.dt
    # eval_f(u, v = 0) against the oracle's residual F(w = 0) = -dt f(u, 0) at (u - 0 dt, 0)
    f = sim.eval_f(gm.zeros_vector(), 0.0, dev(u), gm.zeros_vector())
    F_o = osim.residual(0.0, dt, u, np.zeros(sc.ndof), np.zeros(sc.ndof))
    assert rel(host(f) * (dt * -1.0), F_o) < 1e-10
    # apply_mass: y += M x with the Dirichlet-constrained mass matrix
    rp, col = om.pattern()
    x = rng.standard_normal(sc.ndof)
    y0 = rng.standard_normal(sc.ndof)
    y = sim.apply_mass_matrix(dev(y0), dev(x))
    assert rel(host(y), y0 + O.spmv(rp, col, osim.mass(), x)) < 1e-13
    # solve then apply: H (H^-1 b) = b, and apply equals a product with the assembled matrix
    sim.set_state(0.0, dev(u), gm.zeros_vector())
    sim.init_solve_jacobian_combination(-dt * dt, -dt)
    b = rng.standard_normal(sc.ndof)
    for v_ in sc.dirichlet_nodes:
        b[3 * v_:3 * v_ + 3] = 0.0
    sol = gm.zeros_vector()
    its = sim.solve_jacobian_combination(sol, dev(b))
    assert its > 0
    J = sim.jacobian()  # the matrix of the last solve
    sim.init_apply_jacobian_combination(-dt * dt, -dt, 1.0)
    Hx = sim.apply_jacobian_combination(gm.zeros_vector(), sol)
    assert np.array_equal(host(Hx), host(gm.spmv(J, sol)))
    assert rel(host(Hx), b) < 1e-11
    # the combination itself: M + dt^2 K(u) with the Dirichlet treatment, against the oracle
    K = om.stiffness_add(*scenes.lame_from_young_poisson(sc.young, sc.poisson),
                         np.where(np.repeat(np.isin(np.arange(sc.num_vertices), sc.dirichlet_nodes), 3), 0.0, u),
                         np.zeros(om.nnz))
    J_o = O.dirichlet_csr(rp, col, (0.0 + dt * dt * K) + 1.0 * osim.mass(), sc.dirichlet_nodes)
    assert rel(host(J), J_o) < 1e-12
    # None coefficients drop their terms: gamma M alone
    sim.init_apply_jacobian_combination(None, None, 2.0)
    y = sim.apply_jacobian_combination(gm.zeros_vector(), dev(x))
    assert rel(host(y), 2.0 * O.spmv(rp, col, osim.mass(), x)) < 1e-13


@pytest.mark.parametrize("make", [lambda: scenes.cube_scene(3, scenes.TET10),
                                  lambda: scenes.cylinder_scene(6, scenes.HEX20, 8)])
def test_stiffness_fp32_and_host_pointers(fem, make):
    sc = make()
    om = O.Model(sc)
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    mat = fem.StableNeoHookeanMaterial(mu=mu, lam=lam)
    u = 0.005 * np.random.default_rng(7).standard_normal(sc.ndof)
    K_o = om.stiffness_add(mu, lam, u, np.zeros(om.nnz))
    gm = fem.ElasticityModel(sc)
    K64 = host(gm.assemble_stiffness_into(gm.zeros_matrix(), dev(u), mat))
    assert rel(K64, K_o) < 1e-10
    K32 = torch.zeros(gm.nnz, dtype=torch.float32, device="cuda")
    gm.assemble_stiffness_into(K32, dev(u), mat)
    err32 = rel(host(K32).astype(np.float64), K_o)
    print(f"FP32 stiffness rel L2 vs FP64 oracle: {err32:.2e}")
    assert err32 < 1e-5  # FP32 element tangents: FP32 rounding of every product and sum
    # host pointers: the same calls on numpy arrays give the device results bit for bit
    f_d = host(gm.assemble_elastic_pseudo_forces_into(gm.zeros_vector(), dev(u), mat))
    gm.set_pointer_mode(True)
    f_h = np.zeros(sc.ndof)
    from higher_order_embedded_fem_amd import _lib

    L = _lib.load()
    _lib.check(L.hoefem_model_forces_add(gm.h, mu, lam, u.ctypes.data, f_h.ctypes.data))
    K_h = np.zeros(gm.nnz)
    _lib.check(L.hoefem_model_stiffness_add(gm.h, mu, lam, u.ctypes.data, K_h.ctypes.data))
    assert np.array_equal(f_h, f_d) and np.array_equal(K_h, K64)
    gm.set_pointer_mode(False)


@pytest.mark.parametrize("factor,substeps,newton", [(3e-5, 5, 6), (1e-6, 25, 27)])
def test_substepping_matches_oracle(fem, factor, substeps, newton):
    # a one-iteration Newton limit: the full step fails (the residual after one exact Newton step
    # is above the tolerance), the substeps succeed (the nonlinearity shrinks with the step)
    sc = scenes.cube_scene(3, scenes.TET10)
    kw = dict(newton_tolerance_factor=factor, cg_tolerance=1e-12, cg_max_iter=5000, max_newton_iter=1)
    osim = O.Sim(O.Model(sc), **kw)
    gm = fem.ElasticityModel(sc)
    gsim = fem.NewtonCgIntegrator(gm, sc, **kw)
    u_o, v_o = np.zeros(sc.ndof), np.zeros(sc.ndof)
    u_g = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
    v_g = torch.zeros_like(u_g)
    so = osim.step(0.0, sc.dt, u_o, v_o)
    sg = gsim.step(0.0, sc.dt, u_g, v_g)
    print(f"substeps {sg['substeps']} / {so['substeps']}, newton {sg['newton_iterations']} / {so['newton_iterations']}")
    assert so["substeps"] == sg["substeps"] == substeps
    assert so["newton_iterations"] == sg["newton_iterations"] == newton
    assert rel(host(u_g), u_o) < 1e-8 and rel(host(v_g), v_o) < 1e-8
    # the substepped step equals `substeps` plain steps of dt / substeps from the same state
    gm2 = fem.ElasticityModel(sc)
    gsim2 = fem.NewtonCgIntegrator(gm2, sc, newton_tolerance_factor=factor, cg_tolerance=1e-12, cg_max_iter=5000)
    u2 = torch.zeros_like(u_g)
    v2 = torch.zeros_like(u_g)
    for k in range(substeps):
        s2 = gsim2.step(k * sc.dt / substeps, sc.dt / substeps, u2, v2)
        assert s2["substeps"] == 1 and s2["newton_iterations"] == 1
    assert np.array_equal(host(u2), host(u_g)) and np.array_equal(host(v2), host(v_g))


def test_substepping_gives_up(fem):
    sc = scenes.cube_scene(2, scenes.TET10)
    from higher_order_embedded_fem_amd import _lib

    gm = fem.ElasticityModel(sc)
    # zero Newton iterations allowed and a tolerance no residual meets: every (sub)step fails
    gsim = fem.NewtonCgIntegrator(gm, sc, newton_tolerance_factor=1e-30, max_newton_iter=0)
    u = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
    v = torch.zeros_like(u)
    with pytest.raises(_lib.HoefemError) as e:
        gsim.step(0.0, sc.dt, u, v)
    assert e.value.code == 10  # HOEFEM_ERROR_SUBSTEPPING
    assert float(u.abs().max()) == 0.0 and float(v.abs().max()) == 0.0


def test_line_search_backtracking_matches_oracle(fem):
    # soft material and a long step: Newton needs the backtracking line search in the third step
    sc = scenes.cube_scene(3, scenes.TET10, young=1e4, dt=0.1)
    kw = dict(newton_tolerance_factor=1e-9, cg_tolerance=1e-12, cg_max_iter=5000)
    osim = O.Sim(O.Model(sc), **kw)
    gm = fem.ElasticityModel(sc)
    gsim = fem.NewtonCgIntegrator(gm, sc, **kw)
    u_o, v_o = np.zeros(sc.ndof), np.zeros(sc.ndof)
    u_g = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
    v_g = torch.zeros_like(u_g)
    backtracked = 0
    for k in range(3):
        so = osim.step(k * sc.dt, sc.dt, u_o, v_o)
        sg = gsim.step(k * sc.dt, sc.dt, u_g, v_g)
        print(f"step {k}: newton {sg['newton_iterations']} / {so['newton_iterations']}, evaluations "
              f"{sg['residual_evaluations']} / {so['line_search_evals']}")
        assert sg["newton_iterations"] == so["newton_iterations"]
        assert sg["residual_evaluations"] == so["line_search_evals"]
        backtracked += so["line_search_evals"] - (1 + so["newton_iterations"])
        assert rel(host(u_g), u_o) < 1e-8, (k, rel(host(u_g), u_o))
    assert backtracked >= 1  # at least one rejected trial step length


@pytest.mark.parametrize("make", [lambda: scenes.cube_scene(3, scenes.TET10),
                                  lambda: scenes.cylinder_scene(8, scenes.HEX20, 8)])
def test_representative_force(fem, make):
    sc = make()
    gm = fem.ElasticityModel(sc)
    gsim = fem.NewtonCgIntegrator(gm, sc)
    rp, col = gm.pattern()
    M = host(gsim.mass())
    expected = math.sqrt(100.0 / 3.0) * np.linalg.norm(O.spmv(rp, col, M, np.ones(sc.ndof)))
    assert gsim.representative_force == pytest.approx(expected, rel=1e-13)
    assert gsim.representative_force == pytest.approx(O.Sim(O.Model(sc)).representative_force, rel=1e-12)
