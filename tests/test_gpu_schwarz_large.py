"""Schwarz subdomains above 96 DOFs on the GPU: greedy 64-element patches (node pairs that J does
not couple enter the blocks as zeros) through the blocked, MFMA-tiled Gauss-Jordan path, against the
oracle (whose large blocks go through a tridiagonal QR eigensolver pinned against LAPACK) and
against a LAPACK-built operator (simulation_toolbox/src/fem/schwarz_precond.rs:134-216, :329-342)."""

import numpy as np
import pytest

import oracle as O
from higher_order_embedded_fem_amd import _lib, scenes

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


def lapack_P(rp, col, J, prp, pcol, so, sn):
    n = rp.size - 1
    Jd = np.zeros((n, n))
    Jd[np.repeat(np.arange(n), np.diff(rp)), col] = J
    Pd = np.zeros((n, n))
    for k in range(len(so) - 1):
        nodes = sn[so[k]:so[k + 1]]
        dofs = (3 * nodes[:, None] + np.arange(3)[None, :]).ravel()
        w, V = np.linalg.eigh(Jd[np.ix_(dofs, dofs)])
        lo = 1e-12 * np.abs(w).max()
        Pd[np.ix_(dofs, dofs)] += (V * np.where(np.abs(w) >= lo, 1.0 / np.abs(w), 1.0 / lo)) @ V.T
    return Pd[np.repeat(np.arange(n), np.diff(prp)), pcol]


def rest_jacobian(sc, om, c=1e-4, seed=None):
    rp, col = om.pattern()
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    u = np.zeros(sc.ndof) if seed is None else 0.01 * np.random.default_rng(seed).standard_normal(sc.ndof)
    K = om.stiffness_add(mu, lam, u, np.zeros(om.nnz))
    return rp, col, O.dirichlet_csr(rp, col, (0.0 + c * K) + 1.0 * om.mass(sc.density), sc.dirichlet_nodes)


@pytest.mark.parametrize("cells", [3, 4])
def test_patch_schwarz_matches_oracle_and_lapack(fem, cells):
    sc = scenes.cube_scene(cells, scenes.TET10, schwarz="patches")
    dims = 3 * np.diff(sc.subdomain_offsets)
    print(f"cube{cells} patches: {dims.size} blocks, dims {dims.min()}..{dims.max()}")
    assert dims.max() > 96
    om = O.Model(sc)
    gm = fem.ElasticityModel(sc)
    rp, col, J = rest_jacobian(sc, om, seed=4)
    osw = O.Schwarz(sc.num_vertices, sc.subdomain_offsets, sc.subdomain_nodes, sc.untreated_nodes)
    P_o = osw.update(rp, col, J)
    gsw = fem.SchwarzPreconditioner.from_scene(gm, sc)
    prp, pcol = osw.pattern()
    prp_g, pcol_g = gsw.pattern()
    assert np.array_equal(prp, prp_g) and np.array_equal(pcol, pcol_g)
    gsw.update(dev(J))
    P_g = host(gsw.values())
    P_l = lapack_P(rp, col, J, prp, pcol, sc.subdomain_offsets, sc.subdomain_nodes)
    e_go, e_gl, e_ol = rel(P_g, P_o), rel(P_g, P_l), rel(P_o, P_l)
    print(f"|P_gpu - P_oracle| {e_go:.2e}, |P_gpu - P_lapack| {e_gl:.2e}, |P_oracle - P_lapack| {e_ol:.2e}")
    assert e_go < 1e-9 and e_gl < 1e-9
    # apply, and PCG with it on the SPD rest-state Jacobian
    x = np.random.default_rng(5).standard_normal(sc.ndof)
    assert rel(host(gsw.apply(dev(x))), O.spmv(prp, pcol, P_g, x)) < 1e-13
    rp, col, J0 = rest_jacobian(sc, om)
    P0 = osw.update(rp, col, J0)
    gsw.update(dev(J0))
    b = np.random.default_rng(6).standard_normal(sc.ndof)
    x_o, it_o, st_o = O.pcg(rp, col, J0, b, prec=(prp, pcol, P0), tol=1e-12, max_iter=5000)
    xg = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
    it_g, st_g = fem.conjugate_gradient(gm, dev(J0), dev(b), xg, gsw, tol=1e-12, max_iter=5000)
    assert st_o == 0 and st_g == 0
    assert rel(host(xg), x_o) < 1e-9
    assert abs(it_g - it_o) <= max(2, it_o // 20)


def test_patch_schwarz_rejects_indefinite_blocks(fem):
    sc = scenes.cube_scene(3, scenes.TET10, schwarz="patches")
    om = O.Model(sc)
    gm = fem.ElasticityModel(sc)
    rp, col, J = rest_jacobian(sc, om)
    gsw = fem.SchwarzPreconditioner.from_scene(gm, sc)
    L = _lib.load()
    Jneg = dev(-J)  # every large block negative definite: the first pivot fails
    st = L.hoefem_schwarz_update(gsw.h, Jneg.data_ptr())
    assert st == 3  # HOEFEM_ERROR_SINGULAR
    assert b"not certified" in L.hoefem_last_error()
    # the operator recovers on the next update
    gsw.update(dev(J))
    osw = O.Schwarz(sc.num_vertices, sc.subdomain_offsets, sc.subdomain_nodes, sc.untreated_nodes)
    assert rel(host(gsw.values()), osw.update(rp, col, J)) < 1e-9


def test_patch_schwarz_steps_match_oracle(fem):
    sc = scenes.cube_scene(3, scenes.TET10, schwarz="patches")
    kw = dict(newton_tolerance_factor=1e-9, cg_tolerance=1e-12, cg_max_iter=5000)
    osim = O.Sim(O.Model(sc), **kw)
    gm = fem.ElasticityModel(sc)
    gsim = fem.NewtonCgIntegrator(gm, sc, **kw)
    u_o, v_o = np.zeros(sc.ndof), np.zeros(sc.ndof)
    u_g = torch.zeros(sc.ndof, dtype=torch.float64, device="cuda")
    v_g = torch.zeros_like(u_g)
    for k in range(3):
        so = osim.step(k * sc.dt, sc.dt, u_o, v_o)
        sg = gsim.step(k * sc.dt, sc.dt, u_g, v_g)
        print(f"step {k}: newton {sg['newton_iterations']} / {so['newton_iterations']}, cg {sg['cg_iterations']} / "
              f"{so['cg_iterations']}")
        assert sg["newton_iterations"] == so["newton_iterations"]
        assert abs(sg["cg_iterations"] - so["cg_iterations"]) <= max(4, so["cg_iterations"] // 20)
        assert rel(host(u_g), u_o) < 1e-8 and rel(host(v_g), v_o) < 1e-8
