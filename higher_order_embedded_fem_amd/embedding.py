"""Embedded quadrature of background cells cut by an embedded tetrahedral mesh (setup time).

Restates the reference's embedding pipeline:

* ``embed_mesh_3d`` (fenris/src/embedding/embedding3d.rs:158-206, ``embed_in_cell`` :93-131): every
  background cell is intersected with the embedded mesh and classified by the intersected volume:
  exterior below 1e-4 of the cell volume, interior from 0.999 on, interface in between
  (``EmbedOptions::default``, :147-155).
* ``compute_element_embedded_quadrature`` (:232-281): a tetrahedron rule on every tetrahedron of the
  tetrahedralized intersection, mapped to the background element's reference coordinates with
  w_bg = w_tet |det J_tet| / |det J_bg|, plus an optional stabilization (factor times a rule of the
  uncut cell).
* ``optimize_quadrature`` (fenris/src/embedding/quadrature_reduction.rs:157-227): moment fitting on
  the tensor Chebyshev polynomials of total degree <= strength, scaled to the points' bounding box
  (``build_moment_fitting_system``, :44-126); the constraints P w = P w0 are replaced by the rows
  of V^T of P's SVD above eps * sigma_max, and a basic feasible point of {V_r^T w = V_r^T w0, w >= 0}
  from a simplex method keeps at most rank(P) points.

The reference intersects with PolyMesh3d::intersect_convex_polyhedron and triangulates; here each
embedded tetrahedron is clipped by the six half-spaces of the axis-aligned background cell, which
yields a tetrahedralization of the same intersection directly. Tetrahedron rules are Stroud's
conical product rules (Gauss-Jacobi x Gauss-Jacobi x Gauss-Legendre collapsed onto the simplex,
positive weights, exact to total degree 2n - 1 with n points per direction); the reference's own
symmetric rules have other points but the same strengths. The simplex solver is scipy's HiGHS dual
simplex in place of OR-Tools GLOP (lp-bfp/cpp/lp-bfp.cpp:7-83), which is not available here.
"""

from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

LOWER_VOLUME_THRESHOLD = 1e-4
UPPER_VOLUME_THRESHOLD = 0.999


# --------------------------------------------------------------------------- tetrahedron rules
def tet_rule(strength: int) -> Tuple[np.ndarray, np.ndarray]:
    """Conical product rule on the unit tetrahedron {x, y, z >= 0, x + y + z <= 1} exact for
    polynomials of total degree <= strength. Returns (weights, points[q, 3])."""
    from scipy.special import roots_jacobi, roots_legendre

    n = max(1, math.ceil((strength + 1) / 2))
    tu, wu = roots_jacobi(n, 2.0, 0.0)  # weight (1 - t)^2 on [-1, 1]
    tv, wv = roots_jacobi(n, 1.0, 0.0)  # weight (1 - t)
    tw, ww = roots_legendre(n)
    u, v, w = (tu + 1) / 2, (tv + 1) / 2, (tw + 1) / 2
    wu, wv, ww = wu / 8.0, wv / 4.0, ww / 2.0
    U, V, W = np.meshgrid(u, v, w, indexing="ij")
    WU, WV, WW = np.meshgrid(wu, wv, ww, indexing="ij")
    x = U
    y = (1 - U) * V
    z = (1 - U) * (1 - V) * W
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], axis=1)
    wts = (WU * WV * WW).ravel()
    return wts, pts


# --------------------------------------------------------------------------- clipping
def _signed_volumes(t: np.ndarray) -> np.ndarray:
    a = t[:, 0]
    return np.einsum("ij,ij->i", t[:, 1] - a, np.cross(t[:, 2] - a, t[:, 3] - a)) / 6.0


def clip_tets(t: np.ndarray, normal: np.ndarray, offset: float) -> np.ndarray:
    """The part of each tetrahedron (t[m, 4, 3]) in the half-space normal . x <= offset, as
    tetrahedra: one vertex inside gives 1 tetrahedron, two or three give a prism split into 3."""
    if t.shape[0] == 0:
        return t
    s = t @ normal - offset  # (m, 4)
    inside = s <= 0.0
    cnt = inside.sum(axis=1)
    out = [t[cnt == 4]]
    # order the vertices of every tetrahedron: inside ones first (stable)
    order = np.argsort(~inside, axis=1, kind="stable")
    tt = np.take_along_axis(t, order[:, :, None], axis=1)
    ss = np.take_along_axis(s, order, axis=1)

    def cut(i, j, sel):
        si, sj = ss[sel, i], ss[sel, j]
        f = (si / (si - sj))[:, None]
        return tt[sel, i] + f * (tt[sel, j] - tt[sel, i])

    sel = cnt == 1
    if sel.any():
        A = tt[sel, 0]
        out.append(np.stack([A, cut(0, 1, sel), cut(0, 2, sel), cut(0, 3, sel)], axis=1))
    sel = cnt == 2
    if sel.any():
        A, B = tt[sel, 0], tt[sel, 1]
        pac, pad, pbc, pbd = cut(0, 2, sel), cut(0, 3, sel), cut(1, 2, sel), cut(1, 3, sel)
        # prism: bottom (A, pac, pad), top (B, pbc, pbd)
        out += [np.stack([A, pac, pad, B], 1), np.stack([pac, pad, B, pbc], 1), np.stack([pad, B, pbc, pbd], 1)]
    sel = cnt == 3
    if sel.any():
        A, B, Cc = tt[sel, 0], tt[sel, 1], tt[sel, 2]
        pad, pbd, pcd = cut(0, 3, sel), cut(1, 3, sel), cut(2, 3, sel)
        # prism: bottom (A, B, C), top (pad, pbd, pcd)
        out += [np.stack([A, B, Cc, pad], 1), np.stack([B, Cc, pad, pbd], 1), np.stack([Cc, pad, pbd, pcd], 1)]
    res = np.concatenate(out, axis=0)
    return res[np.abs(_signed_volumes(res)) > 0.0]


def clip_to_box(t: np.ndarray, lo: np.ndarray, hi: np.ndarray) -> np.ndarray:
    """Tetrahedralization of (union of tetrahedra t) intersected with the box [lo, hi]."""
    for axis in range(3):
        n = np.zeros(3)
        n[axis] = -1.0
        t = clip_tets(t, n, -lo[axis])
        n[axis] = 1.0
        t = clip_tets(t, n, hi[axis])
    return t


# --------------------------------------------------------------------------- embedding
class Embedding:
    """embed_mesh_3d's result for an axis-aligned background grid: cell classes and, for every
    interface cell, the tetrahedralized intersection (in physical coordinates)."""

    def __init__(self, interior: np.ndarray, interface: np.ndarray, exterior: np.ndarray,
                 pieces: List[np.ndarray], fractions: np.ndarray):
        self.interior_cells = interior
        self.interface_cells = interface
        self.exterior_cells = exterior
        self.interface_pieces = pieces  # per interface cell: (k, 4, 3) tetrahedra
        self.volume_fractions = fractions  # per background cell


def embed_tet_mesh(cell_lo: np.ndarray, cell_hi: np.ndarray, tet_vertices: np.ndarray,
                   tets: np.ndarray) -> Embedding:
    """embed_mesh_3d for axis-aligned background cells [cell_lo[c], cell_hi[c]] and an embedded
    Tet4 mesh. Candidate tetrahedra per cell come from their bounding boxes (the reference uses an
    R-tree of the embedded cells' bounding boxes)."""
    T = tet_vertices[tets]  # (m, 4, 3)
    tlo, thi = T.min(axis=1), T.max(axis=1)
    ncell = cell_lo.shape[0]
    # bucket tetrahedra by the cells their boxes overlap (uniform grids: index ranges per axis)
    origin = cell_lo.min(axis=0)
    h = (cell_hi - cell_lo)[0]
    dims = np.rint((cell_hi.max(axis=0) - origin) / h).astype(np.int64)
    ijk = np.rint((cell_lo - origin) / h).astype(np.int64)
    cell_at = -np.ones(tuple(dims), np.int64)
    cell_at[ijk[:, 0], ijk[:, 1], ijk[:, 2]] = np.arange(ncell)
    a = np.clip(np.floor((tlo - origin) / h).astype(np.int64), 0, dims - 1)
    b = np.clip(np.floor((thi - origin) / h).astype(np.int64), 0, dims - 1)
    buckets: List[List[int]] = [[] for _ in range(ncell)]
    for k in range(T.shape[0]):
        for i in range(a[k, 0], b[k, 0] + 1):
            for j in range(a[k, 1], b[k, 1] + 1):
                for l in range(a[k, 2], b[k, 2] + 1):
                    c = cell_at[i, j, l]
                    if c >= 0:
                        buckets[c].append(k)
    interior, interface, exterior, pieces = [], [], [], []
    frac = np.zeros(ncell)
    for c in range(ncell):
        vol_cell = float(np.prod(cell_hi[c] - cell_lo[c]))
        if not buckets[c]:
            exterior.append(c)
            continue
        cand = np.array(sorted(buckets[c]))
        sub = clip_to_box(T[cand], cell_lo[c], cell_hi[c])
        vol = float(np.abs(_signed_volumes(sub)).sum()) if sub.shape[0] else 0.0
        frac[c] = vol / vol_cell
        if vol < LOWER_VOLUME_THRESHOLD * vol_cell:
            exterior.append(c)
        elif vol < UPPER_VOLUME_THRESHOLD * vol_cell:
            interface.append(c)
            pieces.append(sub)
        else:
            interior.append(c)
    return Embedding(np.array(interior, np.int64), np.array(interface, np.int64), np.array(exterior, np.int64),
                     pieces, frac)


def element_embedded_quadrature(pieces: np.ndarray, lo: np.ndarray, hi: np.ndarray, rule, stabilization=None):
    """compute_element_embedded_quadrature for the axis-aligned cell [lo, hi] (reference coordinates
    xi = 2 (x - lo) / (hi - lo) - 1): the rule on every piece, mapped, w_bg = w |det J_tet| / det J_bg.
    stabilization = (factor, (weights, points)) appends factor * that rule of the uncut cell."""
    w_t, p_t = rule
    A = pieces[:, 0]
    J = np.stack([pieces[:, 1] - A, pieces[:, 2] - A, pieces[:, 3] - A], axis=2)  # (k, 3, 3) columns
    detj = np.abs(np.linalg.det(J))
    x = A[:, None, :] + np.einsum("kij,qj->kqi", J, p_t)  # (k, q, 3)
    scale = (hi - lo) / 2.0
    xi = ((x - lo) / scale - 1.0).reshape(-1, 3)
    w = (detj[:, None] * w_t[None, :]).reshape(-1) / float(np.prod(scale))
    if stabilization is not None:
        factor, (sw, sp) = stabilization
        w = np.concatenate([w, factor * sw])
        xi = np.concatenate([xi, sp], axis=0)
    return w, xi


# --------------------------------------------------------------------------- reduction
def _chebyshev(x: np.ndarray, n: int) -> np.ndarray:
    if n == 0:
        return np.ones_like(x)
    t0, t1 = np.ones_like(x), x
    for _ in range(n - 1):
        t0, t1 = t1, 2 * x * t1 - t0
    return t1


def moment_matrix(points: np.ndarray, strength: int) -> np.ndarray:
    """build_moment_fitting_system's matrix: rows = tensor Chebyshev polynomials T_i(a x + b)
    T_j(..) T_k(..) with i + j + k <= strength (i slowest), scaled so the points' bounding box maps
    to [-1, 1]^3 (a degenerate direction keeps only constant polynomials), columns = points."""
    lo, hi = points.min(axis=0), points.max(axis=0)
    center, half = (lo + hi) / 2, (hi - lo) / 2
    alpha = np.where(half != 0, 1.0 / np.where(half != 0, half, 1.0), 0.0)
    beta = 1.0 - alpha * (center + half)
    y = points * alpha + beta
    rows = []
    for i in range(strength + 1):
        for j in range(strength + 1):
            for k in range(strength + 1):
                if i + j + k <= strength:
                    rows.append(_chebyshev(y[:, 0], i) * _chebyshev(y[:, 1], j) * _chebyshev(y[:, 2], k))
    return np.array(rows)


def polynomial_space_dim(n: int, d: int) -> int:
    return sum(math.comb(k + d - 1, k) for k in range(n + 1))


def optimize_quadrature(weights: np.ndarray, points: np.ndarray, strength: int):
    """optimize_quadrature: a basic feasible point of the moment-fitting LP; returns the points with
    positive weight."""
    from scipy.optimize import linprog

    P = moment_matrix(points, strength)
    if weights.size <= P.shape[0]:
        return weights.copy(), points.copy()
    _, sv, vt = np.linalg.svd(P, full_matrices=False)
    keep = sv > np.finfo(float).eps * sv.max()
    vr = vt[keep]
    b = vr @ weights
    res = linprog(np.zeros(weights.size), A_eq=vr, b_eq=b, bounds=[(0.0, None)] * weights.size, method="highs-ds")
    if res.status != 0:
        raise RuntimeError(f"quadrature optimization failed: {res.message}")
    w = res.x
    sel = w > 0.0
    return w[sel], points[sel]


# --------------------------------------------------------------------------- embedded geometry
def hollow_cylinder_tet_mesh(r_in: float, r_out: float, half_h: float, nr: int, ntheta: int, nz: int):
    """A Tet4 mesh of the hollow cylinder r in [r_in, r_out], |z| <= half_h: a structured
    (r, theta, z) hex grid, 6 tetrahedra per hex (Nelson Max's split), mapped to Cartesian
    coordinates; the facets approximate the curved surfaces with ntheta segments."""
    from .scenes import hex_to_tets

    r = np.linspace(r_in, r_out, nr + 1)
    th = np.linspace(0.0, 2 * np.pi, ntheta + 1)
    z = np.linspace(-half_h, half_h, nz + 1)
    K, J, I = np.meshgrid(np.arange(nz + 1), np.arange(ntheta + 1), np.arange(nr + 1), indexing="ij")
    R, TH, Z = r[I.ravel()], th[J.ravel()], z[K.ravel()]
    verts = np.stack([R * np.cos(TH), R * np.sin(TH), Z], axis=1)
    param = np.stack([I.ravel(), J.ravel(), K.ravel()], axis=1).astype(np.float64)
    n0, n1 = nr + 1, ntheta + 1

    def idx(i, j, k):
        return k * n0 * n1 + j * n0 + i

    ck, cj, ci = np.meshgrid(np.arange(nz), np.arange(ntheta), np.arange(nr), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    hexes = np.stack([idx(ci, cj, ck), idx(ci + 1, cj, ck), idx(ci + 1, cj + 1, ck), idx(ci, cj + 1, ck),
                      idx(ci, cj, ck + 1), idx(ci + 1, cj, ck + 1), idx(ci + 1, cj + 1, ck + 1),
                      idx(ci, cj + 1, ck + 1)], axis=1)
    tets = hex_to_tets(param, hexes)  # split and orient in parameter space
    return verts, tets
