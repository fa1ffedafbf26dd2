"""GPU parity of the SVD-based material variants against the CPU oracle.

InvertibleMaterial<StableNeoHookean> (fenris/src/solid/materials.rs:862-945, the reference's
default Material with threshold 1e-6, simulation_toolbox/src/fem/fe_model.rs:145-158),
ProjectedStableNeoHookean (materials.rs:387-441) and rotation_svd (fenris/src/util.rs:68-94).

The GPU computes the SVD from the eigenvectors of F^T F, the oracle by one-sided Jacobi
rotations; the GPU removes the negative eigenpairs from dP/dF where the oracle recomposes the
clamped analytic eigensystem, and the GPU contracts dP/dF at the projected F where the oracle
uses the reference's per-pair closed form. Tolerances: 1e-11 relative per deformation gradient
for stress and tangent; 1e-10 (forces) and 1e-9 (stiffness) relative L2 on assembled vectors
with inverted elements; 1e-7 relative L2 on displacements and velocities after each of three
Newton-converged steps.
"""

import numpy as np
import pytest

import oracle as O
from higher_order_embedded_fem_amd import scenes

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

THR = 1e-6


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


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def deformations(seed):
    rng = np.random.default_rng(seed)
    stretches = [(1.2, 0.9, 0.8), (2.0, 0.5, 0.25), (1.1, 0.7, -0.3), (0.9, -0.5, -0.6), (-0.8, -0.9, -1.1),
                 (1.3, 0.6, 1e-8), (1.0, 0.4, -1e-9), (1.5, 1.5, 0.7), (0.3, 0.3, 0.3), (0.05, 1.0, 3.0),
                 (1.0, 0.9, -1.1), (2.0, 1e-7, 1e-7)]
    out = [_rotation(rng) @ np.diag(s) @ _rotation(rng).T for s in stretches for _ in range(4)]
    out += list(np.eye(3) + 0.3 * rng.normal(size=(40, 3, 3)))
    return np.array(out)


def test_rotation_svd(fem):
    F = deformations(0)
    U, S, Vt = (host(t) for t in fem.rotation_svd(dev(F)))
    for k in range(len(F)):
        sv = np.linalg.svd(F[k], compute_uv=False)
        assert np.isclose(np.linalg.det(U[k]), 1.0, atol=1e-12)
        assert np.isclose(np.linalg.det(Vt[k]), 1.0, atol=1e-12)
        assert np.allclose(U[k] @ np.diag(S[k]) @ Vt[k], F[k], rtol=0, atol=1e-13 * sv[0]), k
        assert np.allclose(np.abs(S[k]), sv, rtol=0, atol=1e-13 * sv[0]), k  # sorted by magnitude
        d = np.linalg.det(F[k])
        if abs(d) > 1e-12:
            assert (S[k] < 0).sum() == (1 if d < 0 else 0)
            assert S[k][0] >= 0 and S[k][1] >= 0


def test_material_kinds_match_oracle(fem):
    mu, lam = scenes.lame_from_young_poisson(1e6, 0.4)
    F = deformations(1)
    base = fem.StableNeoHookeanMaterial(mu=mu, lam=lam)
    for mat in (fem.InvertibleMaterial(base, THR), fem.InvertibleMaterial(base, 0.45),
                fem.ProjectedStableNeoHookeanMaterial(mu=mu, lam=lam)):
        if mat.kind == 1:
            Fp = O.project_deformation_gradient(F, mat.threshold)
            keep = np.abs(np.linalg.det(Fp)) > 1e-20
        else:
            keep = np.abs(np.linalg.det(F)) > 1e-6
        G = F[keep]
        P, D = (host(t) for t in mat.evaluate(dev(G)))
        P_o = O.material_stress(mat.kind, mat.threshold, mu, lam, G)
        D_o = O.material_dpdf(mat.kind, mat.threshold, mu, lam, G)
        # Two or more raised singular values sigma_i, sigma_j: the projection keeps F's rotation
        # between their singular subspaces, known only to eps |F| / (sigma_i + sigma_j), and
        # scales it by the threshold; both sides carry that rounding.
        sv = np.linalg.svd(G, compute_uv=False)
        raised = (sv < mat.threshold).sum(1) if mat.kind == 1 else np.zeros(len(G))
        amp = np.where(raised >= 2, mat.threshold / (sv[:, 1] + sv[:, 2]), 1.0)
        for k in range(len(G)):
            tol = 1e-11 + 1e-14 * amp[k]
            assert np.abs(P[k] - P_o[k]).max() <= tol * np.abs(P_o[k]).max(), (mat.kind, mat.threshold, k)
            assert np.abs(D[k] - D_o[k]).max() <= tol * np.abs(D_o[k]).max(), (mat.kind, mat.threshold, k)


def _inverting_displacement(sc, seed):
    """A smooth displacement that inverts a share of the elements: a compression of the cube
    through its centre plane plus noise."""
    X = sc.vertices
    c = X.mean(0)
    rng = np.random.default_rng(seed)
    u = -1.6 * (X - c) * np.array([1.0, 0.2, 0.4]) + 0.02 * rng.standard_normal(X.shape)
    return u.ravel()


@pytest.mark.parametrize("kind", [1, 2])
def test_forces_stiffness_with_inverted_elements(fem, kind):
    sc = scenes.cube_scene(3, scenes.TET10)
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    base = fem.StableNeoHookeanMaterial(mu=mu, lam=lam)
    mat = fem.InvertibleMaterial(base, THR) if kind == 1 else fem.ProjectedStableNeoHookeanMaterial(mu=mu, lam=lam)
    u = _inverting_displacement(sc, 3)
    om = O.Model(sc)
    gm = fem.ElasticityModel(sc)
    om.set_material(kind, mat.threshold)
    f_o = om.forces_add(mu, lam, u, np.zeros(sc.ndof))
    k_o = om.stiffness_add(mu, lam, u, np.zeros(om.nnz))
    # the state exercises the variant: the plain model differs
    plain = O.Model(sc)
    if kind == 1:
        assert rel(plain.forces_add(mu, lam, u, np.zeros(sc.ndof)), f_o) > 1e-6
    else:
        assert rel(plain.stiffness_add(mu, lam, u, np.zeros(om.nnz)), k_o) > 1e-6
    f_g = host(gm.assemble_elastic_pseudo_forces_into(gm.zeros_vector(), dev(u), mat))
    k_g = host(gm.assemble_stiffness_into(gm.zeros_matrix(), dev(u), mat))
    assert rel(f_g, f_o) < 1e-10
    assert rel(k_g, k_o) < 1e-9
    # switching back to the plain model restores the plain assembly
    f_p = host(gm.assemble_elastic_pseudo_forces_into(gm.zeros_vector(), dev(u), base))
    assert rel(f_p, plain.forces_add(mu, lam, u, np.zeros(sc.ndof))) < 1e-10


@pytest.mark.parametrize("kind", [1, 2])
def test_step_with_material_variant(fem, kind):
    """Three Backward-Euler steps from a compressive initial velocity that inverts elements in
    the Newton iterates of the first step."""
    sc = scenes.cube_scene(3, scenes.TET10)
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    base = fem.StableNeoHookeanMaterial(mu=mu, lam=lam)
    mat = fem.InvertibleMaterial(base, THR) if kind == 1 else fem.ProjectedStableNeoHookeanMaterial(mu=mu, lam=lam)
    X = sc.vertices
    v0 = (-(0.5 / sc.dt) * (X - X.mean(0)) * np.array([1.0, 0.3, 0.6])).ravel()
    om = O.Model(sc)
    om.set_material(kind, mat.threshold)
    osim = O.Sim(om, newton_tolerance_factor=1e-9, cg_tolerance=1e-12, cg_max_iter=5000)
    gm = fem.ElasticityModel(sc)
    gsim = fem.NewtonCgIntegrator(gm, sc, newton_tolerance_factor=1e-9, cg_tolerance=1e-12, cg_max_iter=5000,
                                  material=mat)
    u_o, v_o = np.zeros(sc.ndof), v0.copy()
    u_g, v_g = dev(np.zeros(sc.ndof)), dev(v0)
    for k in range(3):
        so = osim.step(k * sc.dt, sc.dt, u_o, v_o)
        sg = gsim.step(k * sc.dt, sc.dt, u_g, v_g)
        assert rel(host(u_g), u_o) < 1e-7, (k, rel(host(u_g), u_o), so, sg)
        assert rel(host(v_g), v_o) < 1e-7
        # Newton iterates that pass through inverted states amplify rounding into the line
        # search's choices, so the two sides may take different paths to the same solution
        n_o, n_g = so["newton_iterations"], sg["newton_iterations"]
        assert abs(n_o - n_g) <= max(2, n_o // 4), (k, so, sg)
