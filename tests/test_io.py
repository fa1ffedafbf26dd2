"""Scene IO (CPU): MSH round trips and the VTK node orders, pinned by geometry (each VTK edge or
face node must sit at the midpoint of the corners VTK's cell definition assigns to it)."""
import numpy as np
import pytest

from higher_order_embedded_fem_amd import io, scenes

# VTK's definitions: edge nodes of the quadratic tetra / hexahedron, face nodes of the
# triquadratic hexahedron (x-, x+, y-, y+, z-, z+)
VTK_TET_EDGES = [(0, 1), (1, 2), (2, 0), (0, 3), (1, 3), (2, 3)]
VTK_HEX_EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (4, 5), (5, 6), (6, 7), (7, 4), (0, 4), (1, 5), (2, 6), (3, 7)]
VTK_HEX_FACES = [(0, 3, 7, 4), (1, 2, 6, 5), (0, 1, 5, 4), (3, 2, 6, 7), (0, 1, 2, 3), (4, 5, 6, 7)]


def test_msh_round_trip(tmp_path):
    hv, hc = scenes.uniform_hex_grid(2, 1.0)
    tets = scenes.hex_to_tets(hv, hc)
    p = tmp_path / "cube.msh"
    io.write_msh(str(p), hv, tets)
    v, t = io.load_msh(str(p))
    assert np.array_equal(v, hv) and np.array_equal(t, tets)


def test_msh22_with_other_elements(tmp_path):
    text = """$MeshFormat
2.2 0 8
$EndMeshFormat
$Nodes
5
10 0 0 0
11 1 0 0
12 0 1 0
13 0 0 1
14 1 1 1
$EndNodes
$Elements
3
1 2 2 0 1 10 11 12
2 4 2 0 1 10 11 12 13
3 4 2 0 1 11 12 13 14
$EndElements
"""
    p = tmp_path / "m22.msh"
    p.write_text(text)
    with pytest.warns(UserWarning):
        v, t = io.load_msh(str(p))
    assert v.shape == (5, 3) and t.tolist() == [[0, 1, 2, 3], [1, 2, 3, 4]]


def _check_vtk_order(verts, conn, etype):
    ctype, perm = io._VTK[etype]
    c = verts[np.asarray(conn)[:, perm]]
    if etype == scenes.TET10:
        for k, (a, b) in enumerate(VTK_TET_EDGES):
            assert np.allclose(c[:, 4 + k], 0.5 * (c[:, a] + c[:, b]))
    if etype in (scenes.HEX20, scenes.HEX27):
        for k, (a, b) in enumerate(VTK_HEX_EDGES):
            assert np.allclose(c[:, 8 + k], 0.5 * (c[:, a] + c[:, b]))
    if etype == scenes.HEX27:
        for k, f in enumerate(VTK_HEX_FACES):
            assert np.allclose(c[:, 20 + k], c[:, list(f)].mean(axis=1))
        assert np.allclose(c[:, 26], c[:, :8].mean(axis=1))


@pytest.mark.parametrize("etype", [scenes.TET10, scenes.HEX20, scenes.HEX27])
def test_vtk_node_orders(etype, tmp_path):
    if etype == scenes.TET10:
        sc = scenes.cube_scene(2, etype)
        verts, conn = sc.vertices, sc.connectivity
    else:
        hv, hc = scenes.uniform_hex_grid(2, 1.0)
        verts, conn = scenes.refine(hv, hc, etype)
    _check_vtk_order(verts, conn, etype)
    p = tmp_path / "out.vtk"
    u = np.random.default_rng(0).standard_normal((verts.shape[0], 3))
    io.write_vtk(str(p), verts, conn, etype, {"displacement": u, "magnitude": np.linalg.norm(u, axis=1)})
    lines = p.read_text().splitlines()
    assert lines[4] == f"POINTS {verts.shape[0]} double"
    assert any(line.startswith(f"CELLS {conn.shape[0]} ") for line in lines)
    assert "VECTORS displacement double" in lines and "SCALARS magnitude double 1" in lines
