"""Oracle checks of the SVD-based material variants (CPU).

InvertibleMaterial<StableNeoHookean> (fenris/src/solid/materials.rs:862-945), rotation_svd
(fenris/src/util.rs:68-94) and ProjectedStableNeoHookean (materials.rs:387-441, analytic
eigensystem :561-687). The oracle computes its SVD by one-sided Jacobi rotations; these tests
pin it against LAPACK's SVD (numpy), pin the projection against the reference's steps written
with that independent SVD, and pin the analytic eigensystem against the plain dP/dF (itself
pinned by the finite-difference and golden tests of the Stable Neo-Hookean model) and the
projected tangent against numpy's symmetric eigensolver.
"""
import numpy as np
import pytest

import oracle as O

YOUNG, POISSON = 1e5, 0.4
MU, LAM = O.lame(YOUNG, POISSON)
THR = 1e-6


def _rotation(rng):
    q, r = np.linalg.qr(rng.normal(size=(3, 3)))
    q = q * np.sign(np.diag(r))
    if np.linalg.det(q) < 0:
        q[:, 0] = -q[:, 0]
    return q


def _deformations(seed=0):
    """Rotations times signed diagonal stretches: ordinary, compressed below the threshold,
    inverted (one or three negative), repeated and rank-deficient ones. No inverted F has a
    tie for the smallest singular value: which triplet the reference flips then depends on the
    order its SVD returns them in."""
    rng = np.random.default_rng(seed)
    stretches = [
        (1.2, 0.9, 0.8), (1.0, 1.0, 1.0), (2.0, 0.5, 0.25), (1.1, 0.7, -0.3), (0.9, -0.5, -0.6),
        (-0.8, -0.9, -1.1), (1.3, 0.6, 1e-8), (1.0, 0.4, -1e-9), (1.5, 1.5, 0.7), (0.3, 0.3, 0.3),
        (1.0, 0.8, 0.0), (2.0, 1e-7, 1e-7), (0.05, 1.0, 3.0), (1.0, 0.9, -1.1),
    ]
    out = []
    for s in stretches:
        for _ in range(3):
            out.append(_rotation(rng) @ np.diag(s) @ _rotation(rng).T)
    out += list(np.eye(3) + 0.3 * rng.normal(size=(20, 3, 3)))
    return np.array(out)


def _reference_projection(F, threshold):
    """project_deformation_gradient's steps with LAPACK's SVD (numpy)."""
    u, s, vt = np.linalg.svd(F)
    u, s = u.copy(), s.copy()
    if np.sign(np.linalg.det(u)) != np.sign(np.linalg.det(vt)):
        k = int(np.argmin(s))
        u[:, k] *= -1.0
        s[k] *= -1.0
    if not np.any(s < threshold):
        return F.copy()
    s = np.maximum(s, threshold)
    return u @ np.diag(s) @ vt


def test_svd3_matches_lapack():
    F = _deformations()
    U, S, V = O.svd3(F)
    for k in range(len(F)):
        ref = np.linalg.svd(F[k], compute_uv=False)
        assert np.allclose(np.sort(S[k])[::-1], ref, rtol=0, atol=1e-14 * max(ref[0], 1.0))
        assert np.allclose(U[k] @ U[k].T, np.eye(3), atol=1e-13)
        assert np.allclose(V[k] @ V[k].T, np.eye(3), atol=1e-13)
        assert np.allclose(U[k] @ np.diag(S[k]) @ V[k].T, F[k], atol=1e-13)


def test_rotation_svd_signs():
    F = _deformations(1)
    U, S, Vt = O.rotation_svd(F)
    for k in range(len(F)):
        assert np.isclose(np.linalg.det(U[k]), 1.0, atol=1e-12)
        assert np.isclose(np.linalg.det(Vt[k]), 1.0, atol=1e-12)
        assert np.allclose(U[k] @ np.diag(S[k]) @ Vt[k], F[k], atol=1e-13)
        ref = np.linalg.svd(F[k], compute_uv=False)
        assert np.allclose(np.sort(np.abs(S[k]))[::-1], ref, atol=1e-14 * max(ref[0], 1.0))
        # at most one negative entry, on the smallest magnitude, iff det F < 0
        neg = S[k] < 0
        d = np.linalg.det(F[k])
        if abs(d) > 1e-12:
            assert neg.sum() == (1 if d < 0 else 0)
            if d < 0:
                assert np.argmin(np.abs(S[k])) == int(np.flatnonzero(neg)[0])


@pytest.mark.parametrize("threshold", [THR, 0.45])
def test_projection_matches_reference_steps(threshold):
    F = _deformations(2)
    P = O.project_deformation_gradient(F, threshold)
    changed = 0
    for k in range(len(F)):
        ref = _reference_projection(F[k], threshold)
        # a raised singular value scales the rounding of its singular vectors by threshold / sigma
        s = np.linalg.svd(F[k], compute_uv=False)
        amp = max(1.0, threshold / max(s[-1], 1e-300)) if s[-1] < threshold else 1.0
        assert np.allclose(P[k], ref, rtol=0, atol=1e-14 * amp * max(1.0, np.abs(F[k]).max())), k
        if not np.array_equal(ref, F[k]):
            changed += 1
            # the result has signed singular values >= threshold and positive determinant
            _, s, _ = O.rotation_svd(P[k][None])
            assert s.min() >= threshold * (1 - 1e-9)
            assert np.linalg.det(P[k]) > 0
            # idempotent up to rounding
            assert np.allclose(O.project_deformation_gradient(P[k][None], threshold)[0], P[k], atol=1e-13)
        else:
            assert np.array_equal(P[k], F[k])  # untouched F is returned bit for bit
    assert changed >= 10


def test_projection_leaves_admissible_f_untouched():
    rng = np.random.default_rng(3)
    F = np.eye(3) + 0.1 * rng.normal(size=(50, 3, 3))
    assert np.array_equal(O.project_deformation_gradient(F, THR), F)


def test_invertible_stress_is_snh_at_projected_f():
    F = _deformations(4)
    keep = [k for k in range(len(F)) if abs(np.linalg.det(O.project_deformation_gradient(F[k][None], THR)[0])) > 0]
    F = F[keep]
    P = O.material_stress(O.MAT_INVERTIBLE_SNH, THR, MU, LAM, F)
    Fp = O.project_deformation_gradient(F, THR)
    assert np.array_equal(P, O.snh_stress(MU, LAM, Fp))
    # admissible F: exactly the Stable Neo-Hookean stress
    rng = np.random.default_rng(5)
    G = np.eye(3) + 0.1 * rng.normal(size=(20, 3, 3))
    assert np.array_equal(O.material_stress(O.MAT_INVERTIBLE_SNH, THR, MU, LAM, G), O.snh_stress(MU, LAM, G))
    # ProjectedStableNeoHookean keeps the plain stress
    assert np.array_equal(O.material_stress(O.MAT_PROJECTED_SNH, 0.0, MU, LAM, G), O.snh_stress(MU, LAM, G))


def test_dpdf_eigensystem_recomposes_dpdf():
    F = _deformations(6)
    for k in range(len(F)):
        if abs(np.linalg.det(F[k])) < 1e-6:
            continue
        ev, Q = O.snh_dpdf_eigen(MU, LAM, F[k])
        D = O.snh_dpdf(MU, LAM, F[k][None])[0]
        scale = np.abs(D).max()
        assert np.allclose(Q.T @ Q, np.eye(9), atol=1e-12), k
        assert np.allclose(Q @ np.diag(ev) @ Q.T, D, rtol=0, atol=1e-11 * scale), k
        assert np.allclose(np.sort(ev), np.linalg.eigvalsh(0.5 * (D + D.T)), rtol=0, atol=1e-11 * scale), k


def test_projected_tangent_is_psd_projection():
    F = _deformations(7)
    negatives = 0
    for k in range(len(F)):
        if abs(np.linalg.det(F[k])) < 1e-6:
            continue
        D = O.snh_dpdf(MU, LAM, F[k][None])[0]
        w, V = np.linalg.eigh(0.5 * (D + D.T))
        negatives += int((w < 0).sum() > 0)
        ref = (V * np.maximum(w, 0.0)) @ V.T
        got = O.material_dpdf(O.MAT_PROJECTED_SNH, 0.0, MU, LAM, F[k][None])[0]
        assert np.allclose(got, ref, rtol=0, atol=1e-11 * np.abs(D).max()), k
        assert np.linalg.eigvalsh(0.5 * (got + got.T)).min() > -1e-9 * np.abs(D).max()
    assert negatives >= 10  # the set exercises the clamp


def test_invertible_contraction_is_pairwise_closed_form():
    """InvertibleMaterial inherits the default per-pair contraction (fenris/src/solid/mod.rs:56-86),
    the closed form of contract_stress_tensor_with (materials.rs:324-347) at the projected F. It
    equals the dP/dF contraction; the closed form goes through F^-1 twice, so on a projected F
    with condition number kappa it carries up to ~eps kappa^2 relative rounding."""
    rng = np.random.default_rng(8)
    a = rng.normal(size=(3, 6))
    for F in _deformations(9):
        Fp = O.project_deformation_gradient(F[None], THR)[0]
        s = np.linalg.svd(Fp, compute_uv=False)
        if s[-1] == 0.0:
            continue
        kappa = s[0] / s[-1]
        got = O.material_contract_multiple(O.MAT_INVERTIBLE_SNH, THR, MU, LAM, F, a)
        ref = O.snh_contract_multiple(MU, LAM, Fp, a)
        tol = 1e-14 + 20 * 2.2e-16 * kappa ** 2
        assert np.abs(got - ref).max() <= tol * np.abs(ref).max(), kappa


def test_unknown_material_kind_rejected():
    with pytest.raises(Exception):
        O.material_stress(7, 0.0, MU, LAM, np.eye(3)[None])
