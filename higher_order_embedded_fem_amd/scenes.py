"""Seeded synthetic scenes for the implicit elastodynamics step.

The generators follow the reference's construction rules:

* ``create_rectangular_uniform_hex_mesh`` (fenris/src/geometry/procedural.rs:191-251):
  lexicographic vertex numbering ``(nx*ny)*k + nx*j + i`` and the Hex8 node order of
  fenris/src/element.rs:1039-1242.
* hex -> 6 tetrahedra following Nelson Max's consistent subdivision as implemented in
  ``PolyMesh::triangulate`` (fenris/src/geometry/polymesh.rs:402-500): every quad face is
  split from its smallest-index vertex, every cell is coned from its smallest-index
  vertex over the faces that do not contain it.
* higher-order refinement with first-seen node numbering keyed by (child index, sorted
  parents) (fenris/src/mesh_convert.rs:41-83, :163-298, :313-405).
* ``Mesh::keep_cells`` relabelling (fenris/src/mesh.rs:321-360): kept vertices are
  renumbered in increasing original index.
* embedded models store interior cells first, then interface cells
  (fenris/src/embedding/mod.rs:291-313).

The embedded quadrature is the seeded rejection-sampled reduced quadrature the
workload definition asks for (it replaces the reference's polyhedral cut + LP moment
fitting, fenris/src/embedding/embedding3d.rs:158-377, which is setup-time work).
"""

from __future__ import annotations

import dataclasses
import math
from typing import Dict, List, Optional

import numpy as np

TET4, TET10, TET20, HEX8, HEX20, HEX27 = 0, 1, 2, 3, 4, 5
NODES_PER_ELEMENT = {TET4: 4, TET10: 10, TET20: 20, HEX8: 8, HEX20: 20, HEX27: 27}
ELEMENT_NAMES = {TET4: "Tet4", TET10: "Tet10", TET20: "Tet20", HEX8: "Hex8", HEX20: "Hex20", HEX27: "Hex27"}


# --------------------------------------------------------------------------- quadrature
def tet_quadrature_strength_1():
    """fenris/src/quadrature.rs:573-592 (reference simplex with vertices at -1/+1)."""
    return np.array([4.0 / 3.0]), np.array([[-0.5, -0.5, -0.5]])


def tet_quadrature_strength_2():
    """4-point degree-2 rule (fenris/src/quadrature.rs:595-634).

    Barycentric points (b, b, b, a) with a = (5 + 3 sqrt 5)/20, b = (5 - sqrt 5)/20,
    mapped to the [-1, 1] reference simplex (x = 2 lambda - 1); weights = volume/4 = 1/3.
    """
    lo = -0.72360679774997896964091736687312762354
    hi = 0.17082039324993690892275210061938287063
    w = np.full(4, 1.0 / 3.0)
    p = np.array([[lo, lo, hi], [lo, hi, lo], [hi, lo, lo], [lo, lo, lo]])
    return w, p


def gauss_legendre_hex(k: int):
    """Tensor Gauss-Legendre rule with k points per axis on [-1, 1]^3 (x fastest)."""
    x, w = np.polynomial.legendre.leggauss(k)
    pts, wts = [], []
    for c in range(k):
        for b in range(k):
            for a in range(k):
                pts.append([x[a], x[b], x[c]])
                wts.append(w[a] * w[b] * w[c])
    return np.array(wts), np.array(pts)


# --------------------------------------------------------------------------- meshes
def uniform_hex_grid(cells: int, length: float = 1.0, origin: float = 0.0):
    """create_rectangular_uniform_hex_mesh(length, 1, 1, 1, cells) translated by origin."""
    nv = cells + 1
    cs = length / cells
    k, j, i = np.meshgrid(np.arange(nv), np.arange(nv), np.arange(nv), indexing="ij")
    verts = np.stack([i.ravel() * cs, j.ravel() * cs, k.ravel() * cs], axis=1) + origin

    def idx(i, j, k):
        return (nv * nv) * k + nv * j + i

    ck, cj, ci = np.meshgrid(np.arange(cells), np.arange(cells), np.arange(cells), indexing="ij")
    ci, cj, ck = ci.ravel(), cj.ravel(), ck.ravel()
    conn = np.stack(
        [
            idx(ci, cj, ck),
            idx(ci + 1, cj, ck),
            idx(ci + 1, cj + 1, ck),
            idx(ci, cj + 1, ck),
            idx(ci, cj, ck + 1),
            idx(ci + 1, cj, ck + 1),
            idx(ci + 1, cj + 1, ck + 1),
            idx(ci, cj + 1, ck + 1),
        ],
        axis=1,
    )
    return verts, conn.astype(np.int64)


_HEX_FACES = [(0, 3, 2, 1), (4, 5, 6, 7), (0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (3, 0, 4, 7)]


def hex_to_tets(verts: np.ndarray, hexes: np.ndarray) -> np.ndarray:
    """Split every hex into 6 tets by Nelson Max's smallest-index rule.

    Each tet is (a, b, c, v) with (a, b, c) a face triangle and v the cell's smallest
    vertex; the vertex order is flipped to positive orientation.
    """
    tets = []
    for h in hexes:
        v = int(h.min())
        for f in _HEX_FACES:
            face = [int(h[k]) for k in f]
            if v in face:
                continue
            m = int(np.argmin(face))
            for i in range(2):
                a, b, c = face[m], face[(i + 1 + m) % 4], face[(i + 2 + m) % 4]
                t = [a, b, c, v]
                x = verts[t]
                if np.linalg.det(np.stack([x[1] - x[0], x[2] - x[0], x[3] - x[0]], axis=1)) < 0:
                    t = [a, c, b, v]
                tets.append(t)
    return np.array(tets, dtype=np.int64)


_TET10_EDGES = [(0, 1), (1, 2), (0, 2), (0, 3), (2, 3), (1, 3)]
_TET20_EDGE_PAIRS = [(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
_TET20_FACES = [(0, 1, 2), (0, 1, 3), (0, 2, 3), (1, 2, 3)]
_HEX_EDGES = [(0, 1), (0, 3), (0, 4), (1, 2), (1, 5), (2, 3), (2, 6), (3, 7), (4, 5), (4, 7), (5, 6), (6, 7)]
_HEX_FACE_NODES = [(0, 1, 2, 3), (0, 1, 4, 5), (0, 3, 4, 7), (1, 2, 5, 6), (2, 3, 6, 7), (4, 5, 6, 7)]


def refine(verts: np.ndarray, conn: np.ndarray, target: int):
    """Refine Tet4 -> Tet10/Tet20 or Hex8 -> Hex20/Hex27 with first-seen numbering.

    Each new node is keyed by (child index, sorted parent vertices); the final label is
    the order of first appearance in the element-major traversal
    (fenris/src/mesh_convert.rs:313-405).
    """
    ne = conn.shape[0]
    keys: List[np.ndarray] = []
    coords: List[np.ndarray] = []
    pad = 8

    def key_of(parents: np.ndarray, child: int = 0) -> np.ndarray:
        p = np.sort(parents, axis=1)
        out = np.full((ne, pad + 1), -1, dtype=np.int64)
        out[:, 0] = child
        out[:, 1 : 1 + p.shape[1]] = p
        return out

    for k in range(conn.shape[1]):
        keys.append(key_of(conn[:, k : k + 1]))
        coords.append(verts[conn[:, k]])
    if target == TET10:
        for a, b in _TET10_EDGES:
            keys.append(key_of(conn[:, [a, b]]))
            coords.append(verts[conn[:, a]] * 0.5 + verts[conn[:, b]] * 0.5)
    elif target == TET20:
        for a, b in _TET20_EDGE_PAIRS:
            # two nodes per edge; child index 0 is the node nearest the smaller global vertex
            pa, pb = verts[conn[:, a]], verts[conn[:, b]]
            lower_first = conn[:, a] < conn[:, b]
            k0 = key_of(conn[:, [a, b]])
            k0[:, 0] = np.where(lower_first, 0, 1)
            keys.append(k0)
            coords.append(pa * (2.0 / 3.0) + pb * (1.0 / 3.0))
            k1 = key_of(conn[:, [a, b]])
            k1[:, 0] = np.where(lower_first, 1, 0)
            keys.append(k1)
            coords.append(pa * (1.0 / 3.0) + pb * (2.0 / 3.0))
        for a, b, c in _TET20_FACES:
            keys.append(key_of(conn[:, [a, b, c]]))
            coords.append((verts[conn[:, a]] + verts[conn[:, b]] + verts[conn[:, c]]) / 3.0)
    elif target in (HEX20, HEX27):
        for a, b in _HEX_EDGES:
            keys.append(key_of(conn[:, [a, b]]))
            coords.append(verts[conn[:, a]] * 0.5 + verts[conn[:, b]] * 0.5)
        if target == HEX27:
            for f in _HEX_FACE_NODES:
                keys.append(key_of(conn[:, list(f)]))
                coords.append(sum(verts[conn[:, k]] for k in f) * 0.25)
            keys.append(key_of(conn[:, :8]))
            coords.append(sum(verts[conn[:, k]] for k in range(8)) * 0.125)
    elif target not in (TET4, HEX8):  # linear targets keep the corner nodes only
        raise ValueError("unsupported refinement target")

    nloc = len(keys)
    allkeys = np.stack(keys, axis=1).reshape(ne * nloc, pad + 1)
    allcoords = np.stack(coords, axis=1).reshape(ne * nloc, 3)
    _, first, inverse = np.unique(allkeys, axis=0, return_index=True, return_inverse=True)
    inverse = inverse.ravel()
    order = np.argsort(first, kind="stable")
    label_of_unique = np.empty_like(order)
    label_of_unique[order] = np.arange(order.size)
    labels = label_of_unique[inverse]
    new_verts = allcoords[first[order]]
    new_conn = labels.reshape(ne, nloc)
    return new_verts, new_conn


def keep_cells(verts: np.ndarray, conn: np.ndarray, cells: np.ndarray):
    """Mesh::keep_cells (fenris/src/mesh.rs:321-360)."""
    kept = conn[cells]
    keep = np.zeros(verts.shape[0], dtype=bool)
    keep[kept.ravel()] = True
    new_index = np.cumsum(keep) - 1
    return verts[keep], new_index[kept]


# --------------------------------------------------------------------------- scenes
@dataclasses.dataclass
class Quadrature:
    """EmbeddedQuadrature (fenris/src/embedding/mod.rs:67-99): one interior rule shared by
    all interior elements plus one rule per interface element (CSR offsets)."""

    interior_w: np.ndarray
    interior_xi: np.ndarray
    interface_offsets: Optional[np.ndarray] = None
    interface_w: Optional[np.ndarray] = None
    interface_xi: Optional[np.ndarray] = None

    def for_element(self, e: int, num_interior: int):
        if e < num_interior:
            return self.interior_w, self.interior_xi
        k = e - num_interior
        a, b = self.interface_offsets[k], self.interface_offsets[k + 1]
        return self.interface_w[a:b], self.interface_xi[a:b]

    def total_points(self, num_interior: int, num_elements: int) -> int:
        n = num_interior * len(self.interior_w)
        if self.interface_offsets is not None:
            n += int(self.interface_offsets[-1])
        return n


@dataclasses.dataclass
class Scene:
    name: str
    element_type: int
    vertices: np.ndarray  # (nv, 3) float64
    connectivity: np.ndarray  # (ne, n) int32
    num_interior: int
    mass_quadrature: Quadrature
    stiffness_quadrature: Quadrature
    elliptic_quadrature: Quadrature
    young: float
    poisson: float
    density: float
    dirichlet_nodes: np.ndarray  # int32, sorted
    subdomain_offsets: np.ndarray  # int64 (nsub+1)
    subdomain_nodes: np.ndarray  # int32
    untreated_nodes: np.ndarray  # int32, sorted
    dt: float
    steps: int
    gravity: float = -9.81
    mass_regularization: float = 0.0
    info: Dict = dataclasses.field(default_factory=dict)

    @property
    def num_vertices(self) -> int:
        return int(self.vertices.shape[0])

    @property
    def num_elements(self) -> int:
        return int(self.connectivity.shape[0])

    @property
    def ndof(self) -> int:
        return 3 * self.num_vertices

    @property
    def nodes_per_element(self) -> int:
        return int(self.connectivity.shape[1])


def _per_element_subdomains(conn: np.ndarray):
    """SchwarzPreconditionerComponent::from_space (simulation_toolbox/src/fem/schwarz_precond.rs:45-63)."""
    ne, n = conn.shape
    return (np.arange(ne + 1, dtype=np.int64) * n), conn.astype(np.int32).ravel().copy(), np.zeros(0, np.int32)


def patch_of_elements(conn: np.ndarray, patch_size: int = 64) -> np.ndarray:
    """Patch index of every element: seeds in element order, each patch grown breadth first over
    element adjacency (elements sharing a node, neighbours in increasing element index) until it
    holds patch_size elements or its front is exhausted."""
    return _grow_patches(np.asarray(conn), patch_size)[0]


def greedy_patches(conn: np.ndarray, patch_size: int = 64):
    """Schwarz subdomains of greedy patches of about patch_size elements (patch_of_elements):
    subdomain k is the sorted union of its elements' nodes (a NestedVec of node sets, as
    SchwarzPreconditionerComponent takes them, simulation_toolbox/src/fem/schwarz_precond.rs:26-42).
    Returns (offsets, nodes, untreated) with no untreated nodes."""
    return _grow_patches(np.asarray(conn), patch_size)[1]


def _grow_patches(conn: np.ndarray, patch_size: int):
    from collections import deque

    ne = conn.shape[0]
    order = np.argsort(conn.ravel(), kind="stable")
    node_of = conn.ravel()[order]
    elem_of = order // conn.shape[1]
    starts = np.searchsorted(node_of, np.arange(int(conn.max()) + 2))
    patch = np.full(ne, -1, dtype=np.int64)
    offsets, nodes = [0], []
    for seed in range(ne):
        if patch[seed] >= 0:
            continue
        pid = len(offsets) - 1
        members = [seed]
        patch[seed] = pid
        front = deque([seed])
        while front and len(members) < patch_size:
            x = front.popleft()
            nb = np.unique(np.concatenate([elem_of[starts[v]:starts[v + 1]] for v in conn[x]]))
            for y in nb:
                if len(members) >= patch_size:
                    break
                if patch[y] < 0:
                    patch[y] = pid
                    members.append(int(y))
                    front.append(int(y))
        union = np.unique(conn[members].ravel())
        nodes.append(union)
        offsets.append(offsets[-1] + union.size)
    return patch, (np.asarray(offsets, dtype=np.int64), np.concatenate(nodes).astype(np.int32), np.zeros(0, np.int32))


def cube_scene(cells: int = 8, element_type: int = TET10, young: float = 1e6, poisson: float = 0.4,
               density: float = 1000.0, dt: float = 0.01, steps: int = 10, schwarz: str = "element") -> Scene:
    """Unit cube of cells^3 hexes, 6 tets per hex, refined to Tet10 (or Tet4/Tet20).

    Tet quadrature strength 2 for mass, stiffness and elliptic terms; nodes on z = 0 fixed.
    Schwarz subdomains: "element" (every element, SchwarzPreconditionerComponent::from_space) or
    "patches" (greedy 64-element patches, greedy_patches).
    """
    hv, hc = uniform_hex_grid(cells, 1.0)
    tets = hex_to_tets(hv, hc)
    if element_type == TET4:
        verts, conn = hv, tets
        # drop nothing: every hex vertex is used by the tets
    else:
        verts, conn = refine(hv, tets, element_type)
    w, xi = tet_quadrature_strength_2()
    q = Quadrature(w, xi)
    dir_nodes = np.nonzero(np.abs(verts[:, 2]) <= 1e-12)[0].astype(np.int32)
    if schwarz == "patches":
        so, sn, su = greedy_patches(conn, 64)
    elif schwarz == "element":
        so, sn, su = _per_element_subdomains(conn)
    else:
        raise ValueError(f"unknown Schwarz subdomains {schwarz!r}")
    return Scene(
        name=f"cube{cells}_{ELEMENT_NAMES[element_type].lower()}" + ("_patches" if schwarz == "patches" else ""),
        element_type=element_type,
        vertices=np.ascontiguousarray(verts, dtype=np.float64),
        connectivity=np.ascontiguousarray(conn, dtype=np.int32),
        num_interior=int(conn.shape[0]),
        mass_quadrature=q,
        stiffness_quadrature=q,
        elliptic_quadrature=q,
        young=young,
        poisson=poisson,
        density=density,
        dirichlet_nodes=dir_nodes,
        subdomain_offsets=so,
        subdomain_nodes=sn,
        untreated_nodes=su,
        dt=dt,
        steps=steps,
    )


def _hollow_cylinder_inside(p: np.ndarray, r_in: float, r_out: float, half_h: float) -> np.ndarray:
    r = np.hypot(p[..., 0], p[..., 1])
    return (r >= r_in) & (r <= r_out) & (np.abs(p[..., 2]) <= half_h)


def _classify_cells(hv, hc, r_in, r_out, half_h, strata=16):
    lo = hv[hc[:, 0]]
    hi = hv[hc[:, 6]]
    x0, y0, z0 = lo.T
    x1, y1, z1 = hi.T
    dx = np.maximum.reduce([np.zeros_like(x0), x0, -x1])
    dy = np.maximum.reduce([np.zeros_like(y0), y0, -y1])
    rmin = np.hypot(dx, dy)
    rmax = np.max(np.stack([np.hypot(a, b) for a in (x0, x1) for b in (y0, y1)]), axis=0)
    interior = (rmin >= r_in) & (rmax <= r_out) & (z0 >= -half_h) & (z1 <= half_h)
    outside = (rmax <= r_in) | (rmin >= r_out) | (z1 <= -half_h) | (z0 >= half_h)
    cut = ~interior & ~outside
    # inside fraction of cut cells by a stratified (cell-centred) strata^3 sample
    s = (np.arange(strata) + 0.5) / strata
    g = np.stack(np.meshgrid(s, s, s, indexing="ij"), axis=-1).reshape(-1, 3)
    frac = np.where(interior, 1.0, 0.0)
    idx = np.nonzero(cut)[0]
    for start in range(0, idx.size, 512):
        sel = idx[start : start + 512]
        pts = lo[sel, None, :] + g[None, :, :] * (hi[sel] - lo[sel])[:, None, :]
        frac[sel] = _hollow_cylinder_inside(pts, r_in, r_out, half_h).mean(axis=1)
    # drop cuts below 1e-4 of the cell volume (fenris/src/embedding/embedding3d.rs:101-107)
    cut &= frac >= 1e-4
    return interior, cut, frac


def _rejection_points(lo, hi, count, rng, r_in, r_out, half_h):
    got = []
    need = count
    while need > 0:
        cand = lo + rng.random((max(64, 4 * need), 3)) * (hi - lo)
        ok = cand[_hollow_cylinder_inside(cand, r_in, r_out, half_h)]
        got.append(ok[:need])
        need -= min(need, ok.shape[0])
    return np.concatenate(got, axis=0)[:count]


def cylinder_scene(cells: int = 24, element_type: int = HEX20, points: int = 8, young: float = 5e5,
                   poisson: float = 0.45, density: float = 1000.0, dt: float = 1.0 / 200.0, steps: int = 20,
                   seed: int = 2024, r_in: float = 0.6, r_out: float = 0.8, height: float = 1.6,
                   mass_points: int = 64, schwarz_threshold: float = 0.5, quadrature: str = "sampled") -> Scene:
    """Embedded hollow cylinder in a cells^3 background grid over [-1, 1]^3.

    Interior cells share a Gauss rule with ``points`` points; every cut cell gets
    ``points`` rejection-sampled points (seeded per background cell) with equal weights
    summing to the cell's inside volume. The mass quadrature uses ``mass_points`` samples
    plus the 1e-8-scaled full-cell stabilisation of
    scene_runner/src/scenes/cylinder_shell.rs:749-762. Schwarz subdomains are the interface
    cells with volume fraction <= 0.5, all other nodes untreated
    (simulation_toolbox/src/fem/schwarz_precond.rs:65-127).

    quadrature="embedded" builds the cut cells' rules the reference's way instead
    (embedded_cylinder_scene).
    """
    if quadrature == "embedded":
        return embedded_cylinder_scene(cells, element_type, points, young=young, poisson=poisson, density=density,
                                       dt=dt, steps=steps, r_in=r_in, r_out=r_out, height=height,
                                       schwarz_threshold=schwarz_threshold)
    if quadrature != "sampled":
        raise ValueError(f"unknown quadrature {quadrature!r}")
    half_h = height / 2.0
    hv, hc = uniform_hex_grid(cells, 2.0, -1.0)
    h = 2.0 / cells
    interior, cut, frac = _classify_cells(hv, hc, r_in, r_out, half_h)
    interior_cells = np.nonzero(interior)[0]
    interface_cells = np.nonzero(cut)[0]
    verts, conn = refine(hv, hc, element_type)
    kept = np.concatenate([interior_cells, interface_cells])
    verts, conn = keep_cells(verts, conn, kept)
    det_ref = (h / 2.0) ** 3
    k_int = round(points ** (1.0 / 3.0))
    if k_int ** 3 != points:
        raise ValueError("points must be a cube number for the interior Gauss rule")
    iw, ixi = gauss_legendre_hex(k_int)
    mw, mxi = gauss_legendre_hex(3)
    sw, sxi = gauss_legendre_hex(4)
    nif = interface_cells.size
    s_off = np.arange(nif + 1, dtype=np.int64) * points
    s_w = np.empty(nif * points)
    s_xi = np.empty((nif * points, 3))
    m_off = np.arange(nif + 1, dtype=np.int64) * (mass_points + sw.size)
    m_w = np.empty(int(m_off[-1]))
    m_xi = np.empty((int(m_off[-1]), 3))
    vol_frac = np.empty(nif)
    for k, c in enumerate(interface_cells):
        lo, hi = hv[hc[c, 0]], hv[hc[c, 6]]
        rng = np.random.default_rng([seed, int(c)])
        v_in = frac[c] * h ** 3
        p = _rejection_points(lo, hi, points, rng, r_in, r_out, half_h)
        s_xi[k * points : (k + 1) * points] = (p - lo) / h * 2.0 - 1.0
        s_w[k * points : (k + 1) * points] = (v_in / points) / det_ref
        pm = _rejection_points(lo, hi, mass_points, rng, r_in, r_out, half_h)
        a = int(m_off[k])
        m_xi[a : a + mass_points] = (pm - lo) / h * 2.0 - 1.0
        m_w[a : a + mass_points] = (v_in / mass_points) / det_ref
        m_xi[a + mass_points : a + mass_points + sw.size] = sxi
        m_w[a + mass_points : a + mass_points + sw.size] = 1e-8 * sw
        vol_frac[k] = m_w[a : int(m_off[k + 1])].sum() * det_ref / h ** 3
    nint = interior_cells.size
    stiff = Quadrature(iw, ixi, s_off, s_w, s_xi)
    mass = Quadrature(mw, mxi, m_off, m_w, m_xi)
    # bottom cap: nodes of kept cells at or below the cap plane
    dir_nodes = np.nonzero(verts[:, 2] <= -half_h + 1e-12)[0].astype(np.int32)
    sel = np.nonzero(vol_frac <= schwarz_threshold)[0]
    sub_conn = conn[nint + sel]
    s_offsets = np.arange(sel.size + 1, dtype=np.int64) * conn.shape[1]
    treated = np.zeros(verts.shape[0], dtype=bool)
    treated[sub_conn.ravel()] = True
    untreated = np.nonzero(~treated)[0].astype(np.int32)
    return Scene(
        name=f"cylinder{cells}_{ELEMENT_NAMES[element_type].lower()}_q{points}",
        element_type=element_type,
        vertices=np.ascontiguousarray(verts, dtype=np.float64),
        connectivity=np.ascontiguousarray(conn, dtype=np.int32),
        num_interior=int(nint),
        mass_quadrature=mass,
        stiffness_quadrature=stiff,
        elliptic_quadrature=stiff,
        young=young,
        poisson=poisson,
        density=density,
        dirichlet_nodes=dir_nodes,
        subdomain_offsets=s_offsets,
        subdomain_nodes=np.ascontiguousarray(sub_conn.ravel(), dtype=np.int32),
        untreated_nodes=untreated,
        dt=dt,
        steps=steps,
        info={"interior_cells": int(nint), "interface_cells": int(nif), "schwarz_cells": int(sel.size)},
    )


def embedded_cylinder_scene(cells: int = 24, element_type: int = HEX20, points: int = 8, young: float = 5e5,
                            poisson: float = 0.45, density: float = 1000.0, dt: float = 1.0 / 200.0,
                            steps: int = 20, r_in: float = 0.6, r_out: float = 0.8, height: float = 1.6,
                            schwarz_threshold: float = 0.5, facet: float = 1.0) -> Scene:
    """The hollow cylinder embedded the reference's way (embedding.py): a Tet4 mesh of the cylinder
    with facets of about facet * h, background cells classified by the intersected volume, and the
    cut cells' rules built from tetrahedron rules on the intersection, mirroring
    cylinder_shell_embedded_hex20 / _hex27 (scene_runner/src/scenes/cylinder_shell.rs:687-800):
    stiffness from a strength 3 source reduced to strength 2 (Hex20, points = 8) or strength 5
    reduced to strength 5 (Hex27, points = 27) by optimize_quadrature; mass from a strength 5 rule on
    every piece (the reference uses strength 10 on its fine embedding mesh; here the pieces of the
    clipped tetrahedra are small and strength 5 keeps the rule at ~5k points per cut cell) plus the
    1e-8 stabilization with the strength 11 rule of the uncut cell. Interior cells use the k^3
    Gauss rule (k^3 = points) for stiffness and 3^3 for mass."""
    from . import embedding as E

    half_h = height / 2.0
    hv, hc = uniform_hex_grid(cells, 2.0, -1.0)
    h = 2.0 / cells
    nr = max(1, math.ceil((r_out - r_in) / (facet * h)))
    nth = max(8, math.ceil(2 * math.pi * r_out / (facet * h)))
    nz = max(1, math.ceil(height / (facet * h)))
    tv, tt = E.hollow_cylinder_tet_mesh(r_in, r_out, half_h, nr, nth, nz)
    lo_all, hi_all = hv[hc[:, 0]], hv[hc[:, 6]]
    emb = E.embed_tet_mesh(lo_all, hi_all, tv, tt)
    interior_cells, interface_cells = emb.interior_cells, emb.interface_cells
    verts, conn = refine(hv, hc, element_type)
    kept = np.concatenate([interior_cells, interface_cells])
    verts, conn = keep_cells(verts, conn, kept)
    k_int = round(points ** (1.0 / 3.0))
    if k_int ** 3 != points:
        raise ValueError("points must be a cube number for the interior Gauss rule")
    src, red, mass_strength = (3, 2, 5) if element_type != HEX27 else (5, 5, 5)
    iw, ixi = gauss_legendre_hex(k_int)
    mw, mxi = gauss_legendre_hex(3)
    sw, sxi = gauss_legendre_hex(6)  # strength 11
    stiff_rule, mass_rule = E.tet_rule(src), E.tet_rule(mass_strength)
    s_w, s_xi, s_cnt, m_w, m_xi, m_cnt = [], [], [], [], [], []
    for k, c in enumerate(interface_cells):
        lo, hi = lo_all[c], hi_all[c]
        w0, p0 = E.element_embedded_quadrature(emb.interface_pieces[k], lo, hi, stiff_rule)
        w, p = E.optimize_quadrature(w0, p0, red)
        s_w.append(w)
        s_xi.append(p)
        s_cnt.append(w.size)
        w, p = E.element_embedded_quadrature(emb.interface_pieces[k], lo, hi, mass_rule, stabilization=(1e-8, (sw, sxi)))
        m_w.append(w)
        m_xi.append(p)
        m_cnt.append(w.size)
    nif = interface_cells.size

    def pack(ws, xs, cnt):
        off = np.zeros(nif + 1, np.int64)
        off[1:] = np.cumsum(cnt)
        if nif == 0:
            return off, np.zeros(0), np.zeros((0, 3))
        return off, np.concatenate(ws), np.concatenate(xs, axis=0)

    s_off, s_w, s_xi = pack(s_w, s_xi, s_cnt)
    m_off, m_w, m_xi = pack(m_w, m_xi, m_cnt)
    stiff = Quadrature(iw, ixi, s_off, s_w, s_xi)
    mass = Quadrature(mw, mxi, m_off, m_w, m_xi)
    nint = interior_cells.size
    dir_nodes = np.nonzero(verts[:, 2] <= -half_h + 1e-12)[0].astype(np.int32)
    sel = np.nonzero(emb.volume_fractions[interface_cells] <= schwarz_threshold)[0]
    sub_conn = conn[nint + sel]
    s_offsets = np.arange(sel.size + 1, dtype=np.int64) * conn.shape[1]
    treated = np.zeros(verts.shape[0], dtype=bool)
    treated[sub_conn.ravel()] = True
    untreated = np.nonzero(~treated)[0].astype(np.int32)
    return Scene(
        name=f"cylinder{cells}_{ELEMENT_NAMES[element_type].lower()}_q{points}_emb",
        element_type=element_type,
        vertices=np.ascontiguousarray(verts, dtype=np.float64),
        connectivity=np.ascontiguousarray(conn, dtype=np.int32),
        num_interior=int(nint),
        mass_quadrature=mass,
        stiffness_quadrature=stiff,
        elliptic_quadrature=stiff,
        young=young,
        poisson=poisson,
        density=density,
        dirichlet_nodes=dir_nodes,
        subdomain_offsets=s_offsets,
        subdomain_nodes=np.ascontiguousarray(sub_conn.ravel(), dtype=np.int32),
        untreated_nodes=untreated,
        dt=dt,
        steps=steps,
        info={"interior_cells": int(nint), "interface_cells": int(nif), "schwarz_cells": int(sel.size),
              "embedded_tets": int(tt.shape[0]), "stiffness_points_interface": int(s_off[-1]),
              "mass_points_interface": int(m_off[-1])},
    )


def reorder(scene: Scene) -> Scene:
    """The scene after reorder_mesh_par (fenris/src/reorder.rs:52-88), as the reference's scenes apply
    it to their meshes (e.g. scene_runner/src/scenes/hollow_ball.rs:398): reverse Cuthill-McKee
    vertices and elements sorted by their smallest new vertex index, computed by the library
    (hoefem_mesh_reorder). Interior elements stay before interface elements, each group in the
    permuted order (as the reference's embedding of a reordered background mesh lists them); the
    per-interface-element quadrature rules, the Dirichlet nodes and the Schwarz node sets follow."""
    import ctypes as C

    from . import _lib

    conn = np.ascontiguousarray(scene.connectivity, dtype=np.int32)
    ne, n = conn.shape
    nv = scene.num_vertices
    vperm = np.zeros(nv, np.int64)
    eperm = np.zeros(ne, np.int64)
    st = _lib.load().hoefem_mesh_reorder(n, nv, ne, conn.ctypes.data, vperm.ctypes.data, eperm.ctypes.data)
    if st != 0:
        raise RuntimeError(f"hoefem_mesh_reorder failed ({st})")
    inv = np.empty(nv, np.int64)
    inv[vperm] = np.arange(nv)
    nint = scene.num_interior
    order = np.concatenate([eperm[eperm < nint], eperm[eperm >= nint]])
    iface = order[nint:] - nint

    def permute_rules(q: Quadrature) -> Quadrature:
        if q.interface_offsets is None or iface.size == 0:
            return q
        off = np.asarray(q.interface_offsets, np.int64)
        cnt = np.diff(off)[iface]
        noff = np.zeros(iface.size + 1, np.int64)
        noff[1:] = np.cumsum(cnt)
        idx = np.concatenate([np.arange(off[k], off[k + 1]) for k in iface])
        xi = np.asarray(q.interface_xi).reshape(-1, 3)
        return Quadrature(q.interior_w, q.interior_xi, noff, np.asarray(q.interface_w)[idx], xi[idx])

    so, sn = np.asarray(scene.subdomain_offsets, np.int64), np.asarray(scene.subdomain_nodes)
    subs = [inv[sn[so[k]:so[k + 1]]] for k in range(len(so) - 1)]
    sorder = sorted(range(len(subs)), key=lambda k: int(subs[k].min()) if subs[k].size else -1)
    new_so = np.zeros(len(subs) + 1, np.int64)
    new_so[1:] = np.cumsum([subs[k].size for k in sorder])
    new_sn = (np.concatenate([subs[k] for k in sorder]) if subs else np.zeros(0)).astype(np.int32)
    return dataclasses.replace(
        scene,
        vertices=np.ascontiguousarray(scene.vertices[vperm]),
        connectivity=np.ascontiguousarray(inv[conn[order]].astype(np.int32)),
        mass_quadrature=permute_rules(scene.mass_quadrature),
        stiffness_quadrature=permute_rules(scene.stiffness_quadrature),
        elliptic_quadrature=permute_rules(scene.elliptic_quadrature),
        dirichlet_nodes=np.sort(inv[scene.dirichlet_nodes]).astype(np.int32),
        subdomain_offsets=new_so,
        subdomain_nodes=new_sn,
        untreated_nodes=np.sort(inv[scene.untreated_nodes]).astype(np.int32),
        name=scene.name + "_rcm",
    )


def config_scene(index: int, **kw) -> Scene:
    """The workload configurations: 1 = Tet10 cube 8^3 with Schwarz subdomains of greedy
    64-element patches, 2 = Hex20 cylinder 24^3 (q=8), 3 = Hex27 cylinder 48^3 (q=27)."""
    if index == 1:
        kw.setdefault("schwarz", "patches")
        return cube_scene(8, TET10, **kw)
    if index == 2:
        return cylinder_scene(24, HEX20, 8, **kw)
    if index == 3:
        return cylinder_scene(48, HEX27, 27, young=5e5, poisson=0.45, **kw)
    raise ValueError(f"unknown configuration {index}")


def pull_scene(scene: Scene, z_top: float = 0.8, z_bottom: float = -0.8) -> Scene:
    """The scene with its top cap held by a Dirichlet condition too, for prescribed pulls: the nodes at
    or below z_bottom stay fixed and the nodes at or above z_top follow a prescribed displacement
    (the hollow cylinder's caps are at z = -0.8 and 0.8). Returns the new scene and the mask of the
    pulled nodes among its Dirichlet nodes."""
    z = scene.vertices[:, 2]
    bottom = np.nonzero(z <= z_bottom + 1e-12)[0]
    top = np.nonzero(z >= z_top - 1e-12)[0]
    nodes = np.union1d(bottom, top).astype(np.int32)
    if nodes.size != bottom.size + top.size:
        raise ValueError("the caps overlap")
    out = dataclasses.replace(scene, name=scene.name + "_pull", dirichlet_nodes=nodes)
    return out, np.isin(nodes, top)


def pull_schedule(steps: int, height: float, seed: int = 0, segment: int = 50, lo: float = 0.05, hi: float = 0.2):
    """Seeded random pull loads: every `segment` steps a new target displacement of the pulled cap (a
    uniformly random direction, magnitude uniform in [lo, hi] * height), reached by a linear ramp from
    the previous one over the segment. Returns the cap displacement at the end of each step, (steps, 3)."""
    rng = np.random.default_rng(seed)
    out = np.zeros((steps, 3))
    prev = np.zeros(3)
    for k0 in range(0, steps, segment):
        d = rng.standard_normal(3)
        target = d / np.linalg.norm(d) * rng.uniform(lo, hi) * height
        for j in range(min(segment, steps - k0)):
            out[k0 + j] = prev + (target - prev) * (j + 1) / segment
        prev = target
    return out


def lame_from_young_poisson(young: float, poisson: float):
    """From<YoungPoisson> for LameParameters (fenris/src/solid/materials.rs:27-38), with
    the reference's lambda = 0.5*mu*nu/(1-2nu) (errata/tex/errata.tex:24-38)."""
    mu = 0.5 * young / (1.0 + poisson)
    lam = 0.5 * mu * poisson / (1.0 - 2.0 * poisson)
    return mu, lam
