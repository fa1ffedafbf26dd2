"""Known answers and edge cases on the GPU path, through the C ABI.

* The reference's CG unit tests (fenris/tests/unit_tests/cg.rs:5-58): the identity operator, an
  arbitrary SPD matrix, and an arbitrary SPD matrix with an arbitrary SPD preconditioner from a
  non-zero initial guess (exactly 3 iterations). They run on the node-0 block of a one-element
  Tet4 model, with the identity on the other nodes and zero right-hand side there, so the
  Krylov spaces are the reference's 3 x 3 ones.
* The reference's block preconditioner test (simulation_toolbox/src/fem/newton_cg.rs:827-876): the
  Schwarz operator with one subdomain per node is the 3 x 3 block-diagonal inverse of the known
  9 x 9 matrix.
* CG status cases: zero right-hand side, iteration limit, no iteration limit (the solve must end on
  the converged status alone).
* A model without elements, and interface elements with empty and single-point rules, against the
  oracle; a degenerate element reports the singular status, and the library stays usable after it.
* An unknown precision mode is rejected at sim creation.
"""

import dataclasses

import numpy as np
import pytest

import oracle as O
from higher_order_embedded_fem_amd import scenes
from test_gpu_parity import dev, host, rel

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

HOEFEM_ERROR_INVALID_ARGUMENT = 1
HOEFEM_ERROR_SINGULAR = 3
HOEFEM_ERROR_CG_MAX_ITERATIONS = 5


def _single_tet(subdomains=(0, 1, 2), untreated=(3,)):
    base = scenes.cube_scene(1, scenes.TET4)
    tet = base.connectivity[0]
    order = {int(a): i for i, a in enumerate(tet)}
    verts = base.vertices[tet]
    conn = np.array([[order[int(a)] for a in tet]], np.int32)
    so = np.arange(len(subdomains) + 1, dtype=np.int64)
    return dataclasses.replace(
        base, name="tet1", vertices=np.ascontiguousarray(verts), connectivity=conn, num_interior=1,
        dirichlet_nodes=np.zeros(0, np.int32), subdomain_offsets=so,
        subdomain_nodes=np.array(subdomains, np.int32), untreated_nodes=np.array(untreated, np.int32))


def _csr_values(rp, col, A):
    rows = np.repeat(np.arange(len(rp) - 1), np.diff(rp))
    return np.ascontiguousarray(A[rows, col])


@pytest.fixture(scope="module")
def fem():
    from higher_order_embedded_fem_amd import fem as F

    return F


@pytest.fixture(scope="module")
def tet1(fem):
    sc = _single_tet()
    gm = fem.ElasticityModel(sc)
    rp, col = gm.pattern()
    assert len(col) == 144  # one element: every DOF couples with every DOF
    return sc, gm, rp, col


def test_cg_identity(fem, tet1):
    sc, gm, rp, col = tet1
    A = _csr_values(rp, col, np.eye(12))
    b = np.full(12, 5.0)
    x = torch.zeros(12, dtype=torch.float64, device="cuda")
    it, st = fem.conjugate_gradient(gm, dev(A), dev(b), x, None, tol=1e-8, max_iter=100)
    assert st == 0 and it <= 1
    assert np.array_equal(host(x), b)


def test_cg_arbitrary_spd(fem, tet1):
    sc, gm, rp, col = tet1
    D = np.eye(12)
    D[:3, :3] = np.where(np.eye(3) > 0, 7.0, 2.0)
    x0 = np.zeros(12)
    x0[:3] = [1.0, 2.0, 3.0]
    b = D @ x0
    x = torch.zeros(12, dtype=torch.float64, device="cuda")
    it, st = fem.conjugate_gradient(gm, dev(_csr_values(rp, col, D)), dev(b), x, None, tol=1e-14, max_iter=100)
    assert st == 0 and 0 < it <= 3
    assert np.abs(host(x) - x0).max() <= 1e-12


def test_cg_arbitrary_preconditioned_three_iterations(fem):
    # Schwarz with the single subdomain {node 0}: its block of the update matrix is inv(p), so
    # the operator on node 0 is the reference's p; the other nodes are untreated with diagonal 1
    sc = _single_tet(subdomains=(0,), untreated=(1, 2, 3))
    gm = fem.ElasticityModel(sc)
    rp, col = gm.pattern()
    a = np.array([[21.0, -1.0, -5.0], [-1.0, 11.0, -4.0], [-5.0, -4.0, 26.0]])
    p = np.array([[17.0, 6.0, 3.0], [6.0, 14.0, 9.0], [3.0, 9.0, 10.0]])
    A = np.eye(12)
    A[:3, :3] = a
    Jp = np.eye(12)
    Jp[:3, :3] = np.linalg.inv(p)
    gsw = fem.SchwarzPreconditioner.from_scene(gm, sc)
    gsw.update(dev(_csr_values(rp, col, Jp)))
    prp, pcol = gsw.pattern()
    P_g = host(gsw.values())
    P_dense = np.zeros((12, 12))
    P_dense[np.repeat(np.arange(12), np.diff(prp)), pcol] = P_g
    assert np.abs(P_dense[:3, :3] - p).max() <= 1e-12 * np.abs(p).max()
    x0 = np.zeros(12)
    x0[:3] = [1.0, 3.0, 2.0]
    b = A @ x0
    xg = np.zeros(12)
    xg[:3] = [2.0, 1.0, 0.0]  # the reference's initial guess
    x = dev(xg)
    it, st = fem.conjugate_gradient(gm, dev(_csr_values(rp, col, A)), dev(b), x, gsw, tol=1e-14, max_iter=100)
    assert st == 0 and it == 3
    assert np.abs(host(x) - x0).max() <= 1e-12


def test_block_preconditioner_known_matrix(fem, tet1):
    sc, gm, rp, col = tet1
    m9 = np.array([
        26.0, 4.0, 12.0, -5.0, -7.0, 1.0, 2.0, 14.0, 11.0, 4.0, 23.0, 0.0, 2.0, -15.0, -11.0, 22.0, 0.0, 5.0,
        12.0, 0.0, 37.0, 6.0, -10.0, -13.0, 6.0, -4.0, 5.0, -5.0, 2.0, 6.0, 36.0, -8.0, 2.0, 5.0, -1.0, 6.0,
        -7.0, -15.0, -10.0, -8.0, 24.0, -4.0, -19.0, -2.0, -13.0, 1.0, -11.0, -13.0, 2.0, -4.0, 31.0, -10.0,
        7.0, 4.0, 2.0, 22.0, 6.0, 5.0, -19.0, -10.0, 33.0, -7.0, 0.0, 14.0, 0.0, -4.0, -1.0, -2.0, 7.0, -7.0,
        37.0, 9.0, 11.0, 5.0, 5.0, 6.0, -13.0, 4.0, 0.0, 9.0, 26.0]).reshape(9, 9)
    J = np.eye(12)
    J[:9, :9] = m9
    gsw = fem.SchwarzPreconditioner.from_scene(gm, sc)
    gsw.update(dev(_csr_values(rp, col, J)))
    x = np.concatenate([[8.0, 7.0, 6.0, 5.0, 4.0, 3.0, 2.0, 1.0, 0.0], [1.0, -2.0, 3.0]])
    y = host(gsw.apply(dev(x)))
    blocks = [np.linalg.inv(m9[3 * k:3 * k + 3, 3 * k:3 * k + 3]) for k in range(3)]
    y_exp = np.concatenate([blocks[k] @ x[3 * k:3 * k + 3] for k in range(3)] + [x[9:]])
    assert np.linalg.norm(y - y_exp) < 1e-13


def test_cg_zero_rhs_and_iteration_limit(fem, tet1):
    sc, gm, rp, col = tet1
    D = np.eye(12)
    D[:3, :3] = np.array([[21.0, -1.0, -5.0], [-1.0, 11.0, -4.0], [-5.0, -4.0, 26.0]])
    A = dev(_csr_values(rp, col, D))
    x = torch.zeros(12, dtype=torch.float64, device="cuda")
    it, st = fem.conjugate_gradient(gm, A, dev(np.zeros(12)), x, None, tol=1e-10, max_iter=100)
    assert st == 0 and it == 0 and not host(x).any()
    b = np.zeros(12)
    b[:3] = [1.0, -2.0, 0.5]
    x = torch.zeros(12, dtype=torch.float64, device="cuda")
    it, st = fem.conjugate_gradient(gm, A, dev(b), x, None, tol=1e-15, max_iter=1)
    assert st == HOEFEM_ERROR_CG_MAX_ITERATIONS and it == 1


def test_cg_without_iteration_limit_stops_on_convergence(fem, tet1):
    # max_iter < 0: no host-side batch limit, the solve ends only on the status the converging kernel
    # writes back (ConjugateGradient without with_max_iter, fenris/src/cg.rs:377-482)
    sc, gm, rp, col = tet1
    D = np.eye(12)
    D[:3, :3] = np.array([[21.0, -1.0, -5.0], [-1.0, 11.0, -4.0], [-5.0, -4.0, 26.0]])
    x0 = np.zeros(12)
    x0[:3] = [1.0, 3.0, 2.0]
    x = torch.zeros(12, dtype=torch.float64, device="cuda")
    it, st = fem.conjugate_gradient(gm, dev(_csr_values(rp, col, D)), dev(D @ x0), x, None, tol=1e-14, max_iter=-1)
    assert st == 0 and 0 < it <= 3
    assert np.abs(host(x) - x0).max() <= 1e-12


def _assembly_matches_oracle(fem, sc, tol=1e-10):
    om = O.Model(sc)
    gm = fem.ElasticityModel(sc)
    assert gm.nnz == om.nnz
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    mat = fem.StableNeoHookeanMaterial(mu=mu, lam=lam)
    u = 0.01 * np.random.default_rng(5).standard_normal(sc.ndof)
    f0 = np.random.default_rng(6).standard_normal(sc.ndof)
    f_o = om.forces_add(mu, lam, u, f0.copy())
    f_g = host(gm.assemble_elastic_pseudo_forces_into(dev(f0), dev(u), mat))
    assert rel(f_g, f_o) <= tol
    if om.nnz:
        k0 = np.random.default_rng(7).standard_normal(om.nnz)
        k_o = om.stiffness_add(mu, lam, u, k0.copy())
        k_g = host(gm.assemble_stiffness_into(dev(k0), dev(u), mat))
        assert rel(k_g, k_o) <= tol
        assert rel(host(gm.assemble_mass(sc.density)), om.mass(sc.density)) <= 1e-12
    return f_g


def test_model_without_elements(fem):
    sc = _single_tet()
    sc = dataclasses.replace(sc, connectivity=np.zeros((0, 4), np.int32), num_interior=0,
                             subdomain_offsets=np.zeros(1, np.int64), subdomain_nodes=np.zeros(0, np.int32),
                             untreated_nodes=np.arange(4, dtype=np.int32))
    f = _assembly_matches_oracle(fem, sc)
    assert np.array_equal(f, np.random.default_rng(6).standard_normal(12))  # f unchanged


def test_interface_elements_with_empty_and_single_point_rules(fem):
    hv, hc = scenes.uniform_hex_grid(2, 1.0)
    verts, conn = scenes.refine(hv, hc, scenes.HEX20)
    w, xi = scenes.gauss_legendre_hex(2)
    ragged = scenes.Quadrature(w, xi, np.array([0, 0, 1], np.int32), np.array([0.7]), np.array([[0.1, -0.2, 0.3]]))
    base = scenes.cylinder_scene(4, scenes.HEX20, 8)
    sc = dataclasses.replace(
        base, name="hex20_ragged", vertices=np.ascontiguousarray(verts), connectivity=np.ascontiguousarray(conn, np.int32),
        num_interior=6, mass_quadrature=ragged, stiffness_quadrature=ragged, elliptic_quadrature=ragged,
        dirichlet_nodes=np.zeros(0, np.int32), subdomain_offsets=np.zeros(1, np.int64),
        subdomain_nodes=np.zeros(0, np.int32), untreated_nodes=np.arange(verts.shape[0], dtype=np.int32),
        mass_regularization=0.0)
    _assembly_matches_oracle(fem, sc)


def test_degenerate_element_reports_singular(fem):
    from higher_order_embedded_fem_amd._lib import HoefemError

    sc = _single_tet()
    flat = sc.vertices.copy()
    flat[:, 2] = 0.0  # all four vertices in one plane: singular reference Jacobian
    bad = dataclasses.replace(sc, vertices=flat)
    gm = fem.ElasticityModel(bad)
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    mat = fem.StableNeoHookeanMaterial(mu=mu, lam=lam)
    with pytest.raises(HoefemError) as info:
        gm.assemble_elastic_pseudo_forces_into(dev(np.zeros(12)), dev(np.zeros(12)), mat)
        torch.cuda.synchronize()
    assert info.value.code == HOEFEM_ERROR_SINGULAR
    # the error is cleared: a valid model assembles normally afterwards
    _assembly_matches_oracle(fem, sc)


@pytest.mark.parametrize("mode", [-1, 4])
def test_unknown_precision_mode_is_rejected(fem, tet1, mode):
    """hoefem_sim_create accepts precision modes 0-3 only; any other value is an invalid argument
    and creates no sim."""
    from higher_order_embedded_fem_amd._lib import HoefemError

    sc, gm, _, _ = tet1
    with pytest.raises(HoefemError) as info:
        fem.NewtonCgIntegrator(gm, sc, mixed_precision=mode)
    assert info.value.code == HOEFEM_ERROR_INVALID_ARGUMENT
    assert "mixed_precision" in str(info.value)
