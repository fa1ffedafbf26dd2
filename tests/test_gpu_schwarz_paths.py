"""The three ways a Schwarz block is inverted on the GPU, each against a LAPACK-based operator.

Blocks B_s of an SPD system matrix are inverted by Gauss-Jordan and accepted when a cheap
condition bound (|B|_F |B^-1|_F or |B|_inf |B^-1|_inf) stays below 5e11; when it does not, a
sharper bound (|B^8|_F^(1/8) |B^-8|_F^(1/8)) decides; the rest go to the Jacobi eigen path with the
reference's clamp max(|lambda|, 1e-12 max|lambda|) (simulation_toolbox/src/fem/schwarz_precond.rs:
329-342). Scaling the three DOFs of one node by sigma drives the condition numbers of the blocks
that hold it through all three cases (checked below with numpy, so the sweep covers each), and
the GPU operator must stay within what the conditioning allows of numpy's eigh-based one. Blocks up
to 64 x 64 and larger ones take different inversion kernels; both sizes are driven through.
"""

import numpy as np
import pytest

import oracle as O
from higher_order_embedded_fem_amd import scenes
from test_gpu_parity import dev, host, rel, schwarz_lapack

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

EPS = np.finfo(np.float64).eps


def _bound(X):
    s = np.linalg.norm(X)
    C = X / s
    for _ in range(3):
        C = C @ C
    return s * np.linalg.norm(C) ** 0.125


def _case(B):
    ev = np.linalg.eigvalsh(B)
    kappa = np.abs(ev).max() / np.abs(ev).min()
    if ev.min() <= 0:
        return "eigen", kappa
    Bi = np.linalg.inv(B)
    cheap = min(np.linalg.norm(B) * np.linalg.norm(Bi), np.abs(B).sum(1).max() * np.abs(Bi).sum(1).max())
    if cheap < 5e11:
        return "cheap", kappa
    return ("power" if _bound(B) * _bound(Bi) < 5e11 else "eigen"), kappa


def _tet10_cube():
    sc = scenes.cube_scene(3, scenes.TET10)
    return sc, int(sc.subdomain_nodes[0]), (1.0, 6e-6, 4e-6, 2e-6, 1e-7)


def _hex27_cylinder():
    # a node of a full 81 x 81 block; the cut cells already make these blocks ill-conditioned
    sc = scenes.cylinder_scene(6, scenes.HEX27, 27)
    off, nodes = np.asarray(sc.subdomain_offsets), np.asarray(sc.subdomain_nodes)
    dset = set(sc.dirichlet_nodes.tolist())
    v = next(int(n) for b in range(len(off) - 1) if off[b + 1] - off[b] == 27 for n in nodes[off[b]:off[b + 1]]
             if int(n) not in dset)
    return sc, v, (1.0, 0.04, 0.03, 0.02)


# 30 x 30 blocks (inverted on the matrix cores, k_spd_inverse_mf) and 81 x 81 blocks (in registers,
# k_spd_inverse): each kernel's accepted, re-certified and rejected blocks
@pytest.mark.parametrize("make", [_tet10_cube, _hex27_cylinder], ids=["tet10_30x30", "hex27_81x81"])
def test_block_inverse_paths(make):
    from higher_order_embedded_fem_amd import fem

    sc, v, sigmas = make()
    om = O.Model(sc)
    mu, lam = scenes.lame_from_young_poisson(sc.young, sc.poisson)
    rp, col = om.pattern()
    K0 = om.stiffness_add(mu, lam, np.zeros(sc.ndof), np.zeros(om.nnz))
    J0 = O.dirichlet_csr(rp, col, 1e-4 * K0 + om.mass(sc.density), sc.dirichlet_nodes)
    gm = fem.ElasticityModel(sc)
    gsw = fem.SchwarzPreconditioner.from_scene(gm, sc)
    prp, pcol = gsw.pattern()
    off, nodes = sc.subdomain_offsets, sc.subdomain_nodes
    assert v not in set(sc.dirichlet_nodes.tolist())
    held = [b for b in range(len(off) - 1) if v in nodes[off[b]:off[b + 1]]]
    rows = np.repeat(np.arange(sc.ndof), np.diff(rp))
    seen = set()
    for sigma in sigmas:
        s = np.ones(sc.ndof)
        s[3 * v:3 * v + 3] = sigma
        J = J0 * s[rows] * s[col]
        kmax = 0.0
        for b in held:
            g = nodes[off[b]:off[b + 1]]
            d = (3 * g[:, None] + np.arange(3)).ravel()
            B = np.zeros((d.size, d.size))
            for a, r in enumerate(d):
                cols = col[rp[r]:rp[r + 1]]
                pos = np.searchsorted(cols, d)
                ok = (pos < cols.size) & (cols[np.minimum(pos, cols.size - 1)] == d)
                B[a, ok] = J[rp[r] + pos[ok]]
            case, kappa = _case(B)
            seen.add(case)
            kmax = max(kmax, kappa)
        gsw.update(dev(J))
        P_g = host(gsw.values())
        P_l = schwarz_lapack(rp, col, J, prp, pcol, sc)
        err = rel(P_g, P_l)
        # inverses of blocks with condition kappa are accurate to ~kappa eps (LAPACK's too); the
        # clamp caps the effective condition at 1e12
        bar = 50.0 * min(kmax, 1e12) * EPS + 1e-12
        print(f"sigma {sigma:g}: kappa_max {kmax:.2e}, |P_gpu - P_lapack| {err:.2e} (bar {bar:.2e})")
        assert err <= bar
    assert seen == {"cheap", "power", "eigen"}
