"""Embedded quadrature (CPU, setup-time): tetrahedron rules, clipping, embedding, moment fitting.

Pinned by exact integrals (monomials over the unit tetrahedron, volumes of known shapes), by the
reference's own test of compute_element_embedded_quadrature (fenris/tests/unit_tests/embedding.rs:
93-151: the cube [0, 0.5]^3 made of hexahedra, embedded in the reference hexahedron, integrates
f = x^5 z^3 + x^3 y^3 z + x + y + z + 5 exactly and has total weight 0.125), and by the moment
conditions the reduced rules must meet.
"""
import math

import numpy as np
import pytest

from higher_order_embedded_fem_amd import embedding as E
from higher_order_embedded_fem_amd import scenes


def _monomial_integral(a, b, c):
    return math.factorial(a) * math.factorial(b) * math.factorial(c) / math.factorial(a + b + c + 3)


@pytest.mark.parametrize("strength", [1, 2, 3, 5, 8, 10])
def test_tet_rule_exact(strength):
    w, p = E.tet_rule(strength)
    assert np.all(w > 0)
    for a in range(strength + 1):
        for b in range(strength + 1 - a):
            for c in range(strength + 1 - a - b):
                got = np.sum(w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** c)
                assert abs(got - _monomial_integral(a, b, c)) <= 1e-14, (a, b, c)


def _box_tets(lo, hi, n):
    hv, hc = scenes.uniform_hex_grid(n, 1.0)
    t = scenes.hex_to_tets(hv, hc)
    v = lo + hv * (np.asarray(hi) - np.asarray(lo))
    return v[t]


def test_clip_volumes():
    T = _box_tets(np.zeros(3), np.ones(3), 3)  # unit cube as tetrahedra
    vol = lambda t: float(np.abs(E._signed_volumes(t)).sum())
    assert abs(vol(T) - 1.0) < 1e-13
    # a box partly outside: the intersection is [0.2, 1] x [0, 0.5] x [0.3, 1]
    sub = E.clip_to_box(T, np.array([0.2, -1.0, 0.3]), np.array([2.0, 0.5, 1.7]))
    assert abs(vol(sub) - 0.8 * 0.5 * 0.7) < 1e-13
    # a plane through the diagonal: half the cube
    half = E.clip_tets(T, np.array([1.0, 1.0, 1.0]), 1.5)
    assert abs(vol(half) - 0.5) < 1e-13
    # cells partitioning space: the pieces add up to the whole volume
    rng = np.random.default_rng(0)
    Tr = T + 0.013 * rng.standard_normal(3)
    tot = 0.0
    for i in range(4):
        for j in range(4):
            for k in range(4):
                lo = np.array([i, j, k]) * 0.35 - 0.1
                tot += vol(E.clip_to_box(Tr, lo, lo + 0.35))
    assert abs(tot - 1.0) < 1e-12


@pytest.mark.parametrize("res", [1, 2])
def test_reference_embedded_quadrature_example(res):
    pieces = _box_tets(np.zeros(3), 0.5 * np.ones(3), res)
    sub = E.clip_to_box(pieces, -np.ones(3), np.ones(3))
    w, xi = E.element_embedded_quadrature(sub, -np.ones(3), np.ones(3), E.tet_rule(10))
    assert abs(w.sum() - 0.125) < 1e-12
    x, y, z = xi.T
    f = x ** 5 * z ** 3 + x ** 3 * y ** 3 * z + x + y + z + 5.0
    a = 0.5
    exact = a ** 11 / 24 + a ** 10 / 32 + 3 * a ** 4 / 2 + 5 * a ** 3
    assert abs(np.sum(w * f) - exact) < 1e-12


def test_embedding_classification_and_volume():
    verts, tets = E.hollow_cylinder_tet_mesh(0.6, 0.8, 0.8, 2, 48, 8)
    exact = float(np.abs(E._signed_volumes(verts[tets])).sum())
    hv, hc = scenes.uniform_hex_grid(8, 2.0, -1.0)
    lo, hi = hv[hc[:, 0]], hv[hc[:, 6]]
    emb = E.embed_tet_mesh(lo, hi, verts, tets)
    h3 = float(np.prod(hi[0] - lo[0]))
    vol = (emb.volume_fractions[emb.interior_cells].sum() + emb.volume_fractions[emb.interface_cells].sum()) * h3
    lost = emb.volume_fractions[emb.exterior_cells].sum() * h3
    assert abs(vol + lost - exact) < 1e-12 * exact
    assert np.all(emb.volume_fractions[emb.interface_cells] < E.UPPER_VOLUME_THRESHOLD)
    assert np.all(emb.volume_fractions[emb.interior_cells] >= E.UPPER_VOLUME_THRESHOLD)
    assert emb.interface_cells.size > 0


@pytest.mark.parametrize("strength", [2, 3])
def test_optimized_quadrature_keeps_moments(strength):
    verts, tets = E.hollow_cylinder_tet_mesh(0.6, 0.8, 0.8, 2, 48, 8)
    hv, hc = scenes.uniform_hex_grid(8, 2.0, -1.0)
    lo, hi = hv[hc[:, 0]], hv[hc[:, 6]]
    emb = E.embed_tet_mesh(lo, hi, verts, tets)
    rule = E.tet_rule(strength + 1)
    dim = E.polynomial_space_dim(strength, 3)
    for k, c in enumerate(emb.interface_cells[:20]):
        w0, p0 = E.element_embedded_quadrature(emb.interface_pieces[k], lo[c], hi[c], rule)
        w, p = E.optimize_quadrature(w0, p0, strength)
        assert np.all(w > 0) and w.size <= dim
        # every monomial of degree <= strength keeps its integral
        for a in range(strength + 1):
            for b in range(strength + 1 - a):
                for cc in range(strength + 1 - a - b):
                    m0 = np.sum(w0 * p0[:, 0] ** a * p0[:, 1] ** b * p0[:, 2] ** cc)
                    m1 = np.sum(w * p[:, 0] ** a * p[:, 1] ** b * p[:, 2] ** cc)
                    assert abs(m1 - m0) <= 1e-10 * max(abs(w0).sum(), 1e-300), (k, a, b, cc)


def test_embedded_cylinder_scene_rules():
    sc = scenes.cylinder_scene(8, scenes.HEX20, 8, quadrature="embedded")
    q = sc.stiffness_quadrature
    off = q.interface_offsets
    cnt = np.diff(off)
    assert np.all(cnt <= E.polynomial_space_dim(2, 3)) and np.all(q.interface_w > 0)
    # the reduced rules keep the cut volume: sum of weights = 8 (reference cell volume) x fraction
    m = sc.mass_quadrature
    for k in range(len(cnt)):
        ws = q.interface_w[off[k]:off[k + 1]].sum()
        wm = m.interface_w[m.interface_offsets[k]:m.interface_offsets[k + 1]]
        # the mass rule carries the 1e-8 stabilization on top of the cut volume
        assert abs(ws - (wm.sum() - 1e-8 * 8.0)) <= 1e-12 * 8.0
        assert 0.0 < ws < 8.0 * E.UPPER_VOLUME_THRESHOLD + 1e-12
    assert sc.num_interior + len(cnt) == sc.num_elements
