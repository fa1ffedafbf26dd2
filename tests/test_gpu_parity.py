"""GPU parity: the HIP library against the CPU oracle on the same seeded inputs.

Tolerances: forces / stiffness / mass within 1e-10 relative L2 (stated bar); CSR pattern and
element coloring bit-exact; Schwarz operator within 1e-9 (the two sides use different
Jacobi orderings); per-step displacements within 1e-8 relative L2 with tightened solver
tolerances (so both converge to the same Newton solution).
"""

import numpy as np
import pytest

import oracle as O
from higher_order_embedded_fem_amd import scenes

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


SMALL_SCENES = {
    "cube3_tet10": lambda: scenes.cube_scene(3, scenes.TET10),
    "cube2_tet4": lambda: scenes.cube_scene(2, scenes.TET4),
    "cube2_tet20": lambda: scenes.cube_scene(2, scenes.TET20),
    "cyl8_hex20": lambda: scenes.cylinder_scene(8, scenes.HEX20, 8),
    "cyl6_hex27": lambda: scenes.cylinder_scene(6, scenes.HEX27, 27),
    "cyl6_hex8": lambda: scenes.cylinder_scene(6, scenes.HEX8, 8),
    "cyl8_hex20_emb": lambda: scenes.cylinder_scene(8, scenes.HEX20, 8, quadrature="embedded"),
}


@pytest.fixture(scope="module")
def fem():
    from higher_order_embedded_fem_amd import fem as F

    return F


def random_displacement(scene, scale, seed):
    rng = np.random.default_rng(seed)
    u = scale * rng.standard_normal(scene.ndof)
    return u


def schwarz_lapack(rp, col, J, prp, pcol, sc):
    """P = sum_s R_s^T B_s^+ R_s + 1/J_ii on untreated nodes, with B_s^+ from numpy's eigh
    (LAPACK ?syevd) and the clamp max(|lambda|, 1e-12 max|lambda|)
    (simulation_toolbox/src/fem/schwarz_precond.rs:171-216, 329-342). A third computation of
    the Schwarz operator, independent of both Jacobi eigen-solvers."""
    n = len(rp) - 1
    keys = np.repeat(np.arange(n, dtype=np.int64), np.diff(prp)) * n + pcol
    P = np.zeros(len(pcol))

    def entries(rows, cols):
        k = rows[:, None].astype(np.int64) * n + cols[None, :]
        pos = np.searchsorted(keys, k)
        assert np.all(keys[pos] == k)
        return pos

    for s in range(len(sc.subdomain_offsets) - 1):
        nodes = sc.subdomain_nodes[sc.subdomain_offsets[s]:sc.subdomain_offsets[s + 1]]
        dofs = (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()
        B = np.zeros((len(dofs), len(dofs)))
        for a, r in enumerate(dofs):
            cols = col[rp[r]:rp[r + 1]]
            pos = np.searchsorted(cols, dofs)
            assert np.all(cols[pos] == dofs)
            B[a] = J[rp[r] + pos]
        w, V = np.linalg.eigh(B)
        lower = 1e-12 * np.abs(w).max()
        f = np.where(np.abs(w) >= lower, 1.0 / np.abs(w), 1.0 / lower)
        np.add.at(P, entries(dofs, dofs), (V * f) @ V.T)
    for v_ in sc.untreated_nodes:
        for i in range(3):
            r = 3 * v_ + i
            k = rp[r] + np.searchsorted(col[rp[r]:rp[r + 1]], r)
            P[prp[r]] = 1.0 / J[k]
    return P


def test_material_eval_matches_oracle(fem):
    rng = np.random.default_rng(1)
    F = np.eye(3) + 0.2 * rng.standard_normal((64, 3, 3))
    mu, lam = scenes.lame_from_young_poisson(1e6, 0.4)
    mat = fem.StableNeoHookeanMaterial(mu=mu, lam=lam)
    P, D = mat.evaluate(dev(F))
    assert rel(host(P), O.snh_stress(mu, lam, F)) < 1e-13
    assert rel(host(D), O.snh_dpdf(mu, lam, F)) < 1e-13
    # the energy density (a diagnostic off the Newton path), also at F = I where it vanishes
    G = np.concatenate([F, np.eye(3)[None]])
    psi = host(mat.energy(dev(G)))
    assert rel(psi, O.snh_energy(mu, lam, G)) < 1e-13
    assert abs(psi[-1]) <= 1e-9 * mu


@pytest.mark.parametrize("name", list(SMALL_SCENES))
def test_pattern_and_colors_bit_exact(fem, name):
    sc = SMALL_SCENES[name]()
    om = O.Model(sc)
    gm = fem.ElasticityModel(sc)
    rp_o, col_o = om.pattern()
    rp_g, col_g = gm.pattern()
    assert np.array_equal(rp_o, rp_g)
    assert np.array_equal(col_o, col_g)
    for p in (0, 1):
        off_o, lab_o = om.colors(p)
        off_g, lab_g = gm.colors(p)
        assert np.array_equal(off_o, off_g)
        assert np.array_equal(lab_o, lab_g)


@pytest.mark.parametrize("name", list(SMALL_SCENES))
def test_forces_stiffness_mass_match_oracle(fem, name):
    sc = SMALL_SCENES[name]()
    om = O.Model(sc)
    gm = fem.ElasticityModel(sc)
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    mat = fem.StableNeoHookeanMaterial(mu=mu, lam=lam)
    u = random_displacement(sc, 0.01, 7)
    f0 = np.random.default_rng(3).standard_normal(sc.ndof)
    # forces accumulate into f
    f_o = om.forces_add(mu, lam, u, f0.copy())
    f_g = host(gm.assemble_elastic_pseudo_forces_into(dev(f0), dev(u), mat))
    assert rel(f_g, f_o) < 1e-10
    # stiffness accumulates into the CSR values
    k0 = np.random.default_rng(4).standard_normal(om.nnz)
    k_o = om.stiffness_add(mu, lam, u, k0.copy())
    k_g = host(gm.assemble_stiffness_into(dev(k0), dev(u), mat))
    assert rel(k_g, k_o) < 1e-10
    m_o = om.mass(sc.density)
    m_g = host(gm.assemble_mass(sc.density))
    assert rel(m_g, m_o) < 1e-12


@pytest.mark.parametrize("name", ["cube3_tet10", "cyl8_hex20", "cyl6_hex27"])
def test_dirichlet_and_schwarz_match_oracle(fem, name):
    sc = SMALL_SCENES[name]()
    om = O.Model(sc)
    gm = fem.ElasticityModel(sc)
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    u = random_displacement(sc, 0.005, 11)
    rp, col = om.pattern()
    K = om.stiffness_add(mu, lam, u, np.zeros(om.nnz))
    M = om.mass(sc.density)
    J = (0.0 + 1e-4 * K) + 1.0 * M
    J_o = O.dirichlet_csr(rp, col, J.copy(), sc.dirichlet_nodes)
    J_g = host(gm.apply_homogeneous_dirichlet_bc_csr(dev(J), sc.dirichlet_nodes))
    assert np.array_equal(J_g, J_o)
    osw = O.Schwarz(sc.num_vertices, sc.subdomain_offsets, sc.subdomain_nodes, sc.untreated_nodes)
    P_o = osw.update(rp, col, J_o)
    gsw = fem.SchwarzPreconditioner.from_scene(gm, sc)
    prp_o, pcol_o = osw.pattern()
    prp_g, pcol_g = gsw.pattern()
    assert np.array_equal(prp_o, prp_g) and np.array_equal(pcol_o, pcol_g)
    gsw.update(dev(J_o))
    P_g = host(gsw.values())
    # The block pseudo-inverses are as accurate as the blocks' conditioning allows (eps * kappa);
    # Hex27 blocks of nearly empty cut cells reach kappa ~ 1e8. The bar: within 1e-9 of the
    # oracle, or no further from a LAPACK-based P than the oracle itself is (times 4).
    err_go = rel(P_g, P_o)
    if err_go >= 1e-9:
        P_l = schwarz_lapack(rp, col, J_o, prp_o, pcol_o, sc)
        err_gl, err_ol = rel(P_g, P_l), rel(P_o, P_l)
        print(f"{name}: |P_gpu - P_oracle| {err_go:.2e}, |P_gpu - P_lapack| {err_gl:.2e}, "
              f"|P_oracle - P_lapack| {err_ol:.2e}")
        assert err_gl <= 4.0 * err_ol and err_go < 1e-7
    # the apply kernel against a CPU SpMV with the same (GPU) operator values
    x = np.random.default_rng(5).standard_normal(sc.ndof)
    assert rel(host(gsw.apply(dev(x))), O.spmv(prp_o, pcol_o, P_g, x)) < 1e-13
    # PCG with the Schwarz preconditioner, tight tolerance, on the SPD rest-state Jacobian
    # (a deformed state can make M + c K indefinite where background nodes carry tiny mass)
    K0 = om.stiffness_add(mu, lam, np.zeros(sc.ndof), np.zeros(om.nnz))
    J0 = O.dirichlet_csr(rp, col, (0.0 + 1e-4 * K0) + 1.0 * M, sc.dirichlet_nodes)
    P0 = osw.update(rp, col, J0)
    gsw.update(dev(J0))
    b = np.random.default_rng(6).standard_normal(sc.ndof)
    x_o, it_o, st_o = O.pcg(rp, col, J0, b, prec=(prp_o, pcol_o, P0), tol=1e-12, max_iter=5000)
    xg = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
    it_g, st_g = fem.conjugate_gradient(gm, dev(J0), dev(b), xg, gsw, tol=1e-12, max_iter=5000)
    assert st_o == 0 and st_g == 0
    assert rel(host(xg), x_o) < 1e-9
    assert abs(it_g - it_o) <= max(2, it_o // 20)


@pytest.mark.parametrize("name", ["cube3_tet10", "cyl8_hex20", "cyl8_hex20_emb"])
def test_step_matches_oracle(fem, name):
    sc = SMALL_SCENES[name]()
    om = O.Model(sc)
    osim = O.Sim(om, newton_tolerance_factor=1e-9, cg_tolerance=1e-12, cg_max_iter=5000)
    gm = fem.ElasticityModel(sc)
    gsim = fem.NewtonCgIntegrator(gm, sc, newton_tolerance_factor=1e-9, cg_tolerance=1e-12, cg_max_iter=5000)
    assert abs(gsim.representative_force - osim.representative_force) <= 1e-12 * osim.representative_force
    assert rel(host(gsim.mass()), osim.mass()) < 1e-12
    u_o, v_o = np.zeros(sc.ndof), np.zeros(sc.ndof)
    u_g = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
    v_g = torch.zeros_like(u_g)
    for k in range(3):
        so = osim.step(k * sc.dt, sc.dt, u_o, v_o)
        sg = gsim.step(k * sc.dt, sc.dt, u_g, v_g)
        assert so["newton_iterations"] == sg["newton_iterations"]
        assert rel(host(u_g), u_o) < 1e-8, (k, rel(host(u_g), u_o))
        assert rel(host(v_g), v_o) < 1e-8
