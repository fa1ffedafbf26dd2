"""Scene IO: Gmsh meshes in, VTK out (setup and output, outside the time step).

* ``load_msh`` reads the Tet4 volume meshes the reference's scenes load
  (simulation_toolbox/src/io/msh.rs:9-260, through the mshio crate): MSH 4.1 or 2.2 ASCII, 3D
  nodes, tetrahedra (Gmsh element type 4); other element types are skipped with the reference's
  warning semantics. Node tags are renumbered consecutively in file order.
* ``write_msh`` writes a Tet4 mesh as MSH 4.1 ASCII (for round trips and tests).
* ``write_vtk`` writes a legacy ASCII VTK unstructured grid with point data, the format of the
  reference's VtkFemOutput (simulation_toolbox/src/io/vtk.rs): linear and quadratic cells in VTK's
  node order (Tet4 10, Tet10 24, Hex8 12, Hex20 25, Hex27 29); Tet20 is written by its corners.
"""

from __future__ import annotations

import warnings
from typing import Dict, Optional, Tuple

import numpy as np

from .scenes import HEX8, HEX20, HEX27, TET4, TET10, TET20

# element node order -> VTK node order (VTK_QUADRATIC_TETRA: edges (0,1) (1,2) (2,0) (0,3) (1,3)
# (2,3); VTK_QUADRATIC_HEXAHEDRON: bottom edges, top edges, vertical edges; VTK_TRIQUADRATIC_
# HEXAHEDRON: then the faces x-, x+, y-, y+, z-, z+ and the centre)
_VTK = {
    TET4: (10, list(range(4))),
    TET10: (24, [0, 1, 2, 3, 4, 5, 6, 7, 9, 8]),
    TET20: (10, list(range(4))),
    HEX8: (12, list(range(8))),
    HEX20: (25, list(range(8)) + [8, 11, 13, 9, 16, 18, 19, 17, 10, 12, 14, 15]),
    HEX27: (29, list(range(8)) + [8, 11, 13, 9, 16, 18, 19, 17, 10, 12, 14, 15] + [22, 23, 21, 24, 20, 25, 26]),
}


def _sections(text: str) -> Dict[str, list]:
    out: Dict[str, list] = {}
    lines = text.splitlines()
    i = 0
    while i < len(lines):
        line = lines[i].strip()
        if line.startswith("$") and not line.startswith("$End"):
            name = line[1:]
            j = i + 1
            body = []
            while not lines[j].strip().startswith("$End" + name):
                body.append(lines[j])
                j += 1
            out[name] = body
            i = j + 1
        else:
            i += 1
    return out


def load_msh(path: str) -> Tuple[np.ndarray, np.ndarray]:
    """(vertices[n, 3], tetrahedra[m, 4]) of a Gmsh ASCII file (format 4.1 or 2.2)."""
    with open(path, "r") as f:
        sec = _sections(f.read())
    fmt = sec["MeshFormat"][0].split()
    version, filetype = float(fmt[0]), int(fmt[1])
    if filetype != 0:
        raise ValueError("only ASCII MSH files are supported")
    tags, coords, tets = [], [], []
    skipped = False
    if version >= 4.0:
        body = sec["Nodes"]
        nblocks = int(body[0].split()[0])
        k = 1
        for _ in range(nblocks):
            _, _, _, n = (int(x) for x in body[k].split())
            k += 1
            tags += [int(body[k + i]) for i in range(n)]
            coords += [[float(x) for x in body[k + n + i].split()[:3]] for i in range(n)]
            k += 2 * n
        body = sec["Elements"]
        nblocks = int(body[0].split()[0])
        k = 1
        for _ in range(nblocks):
            _, _, etype, n = (int(x) for x in body[k].split())
            k += 1
            if etype == 4:
                tets += [[int(x) for x in body[k + i].split()[1:5]] for i in range(n)]
            else:
                skipped = True
            k += n
    else:
        body = sec["Nodes"]
        n = int(body[0])
        for i in range(n):
            parts = body[1 + i].split()
            tags.append(int(parts[0]))
            coords.append([float(x) for x in parts[1:4]])
        body = sec["Elements"]
        n = int(body[0])
        for i in range(n):
            parts = [int(x) for x in body[1 + i].split()]
            if parts[1] == 4:
                ntags = parts[2]
                tets.append(parts[3 + ntags:7 + ntags])
            else:
                skipped = True
    if skipped:
        warnings.warn("MSH file holds elements other than tetrahedra; they are ignored")
    index = {t: i for i, t in enumerate(tags)}
    verts = np.array(coords, dtype=np.float64).reshape(-1, 3)
    conn = np.array([[index[t] for t in e] for e in tets], dtype=np.int64).reshape(-1, 4)
    return verts, conn


def write_msh(path: str, vertices: np.ndarray, tets: np.ndarray) -> None:
    """A Tet4 mesh as MSH 4.1 ASCII: one node block and one element block on volume entity 1."""
    n, m = vertices.shape[0], tets.shape[0]
    with open(path, "w") as f:
        f.write("$MeshFormat\n4.1 0 8\n$EndMeshFormat\n")
        f.write(f"$Nodes\n1 {n} 1 {n}\n3 1 0 {n}\n")
        f.writelines(f"{i + 1}\n" for i in range(n))
        f.writelines(f"{x:.17g} {y:.17g} {z:.17g}\n" for x, y, z in vertices)
        f.write("$EndNodes\n")
        f.write(f"$Elements\n1 {m} 1 {m}\n3 1 4 {m}\n")
        f.writelines(f"{e + 1} {a + 1} {b + 1} {c + 1} {d + 1}\n" for e, (a, b, c, d) in enumerate(tets))
        f.write("$EndElements\n")


def write_vtk(path: str, vertices: np.ndarray, connectivity: np.ndarray, element_type: int,
              point_data: Optional[Dict[str, np.ndarray]] = None, title: str = "hoefem") -> None:
    """Legacy ASCII VTK unstructured grid; point_data arrays of shape (n,) or (n, 3)."""
    ctype, perm = _VTK[element_type]
    cells = np.asarray(connectivity)[:, perm]
    n, m = vertices.shape[0], cells.shape[0]
    k = cells.shape[1]
    with open(path, "w") as f:
        f.write(f"# vtk DataFile Version 3.0\n{title}\nASCII\nDATASET UNSTRUCTURED_GRID\n")
        f.write(f"POINTS {n} double\n")
        f.writelines(f"{x:.17g} {y:.17g} {z:.17g}\n" for x, y, z in vertices)
        f.write(f"CELLS {m} {m * (k + 1)}\n")
        f.writelines(f"{k} " + " ".join(str(int(v)) for v in c) + "\n" for c in cells)
        f.write(f"CELL_TYPES {m}\n")
        f.writelines(f"{ctype}\n" for _ in range(m))
        if point_data:
            f.write(f"POINT_DATA {n}\n")
            for name, arr in point_data.items():
                a = np.asarray(arr, dtype=np.float64).reshape(n, -1)
                if a.shape[1] == 3:
                    f.write(f"VECTORS {name} double\n")
                    f.writelines(f"{x:.17g} {y:.17g} {z:.17g}\n" for x, y, z in a)
                else:
                    f.write(f"SCALARS {name} double 1\nLOOKUP_TABLE default\n")
                    f.writelines(f"{x:.17g}\n" for x in a[:, 0])
