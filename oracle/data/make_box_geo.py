#!/usr/bin/env python3
"""Write ParMooN GEO files (the text format TDomain::ReadGeo reads,
ReadGeo.C:46-200) for a box split into m_x x m_y x m_z macro bricks.

Vertices carry explicit coordinates (KNPR 0), so the reader does not scale
them (the parametrised branch at ReadGeo.C:1733-1755 is never taken).  Cells
are listed with x fastest, then y, then z; each cell's vertices in the
hexahedron order of the reference (bottom face counter-clockwise, then the
top face), so every macro cell has the identity local frame.  Boundary faces
use the hexahedron face numbering of the reference (face 1 = 1 2 3 4, 2 =
1 2 6 5, 3 = 2 3 7 6, 4 = 3 4 8 7, 5 = 4 1 5 8, 6 = 5 6 7 8) and the six
plane components of unitcube.PRM (1 z=0, 2 x=0, 3 y=0, 4 z=1, 5 x=1, 6 y=1).

A deformation can be given to move the interior vertices (general
hexahedra, trilinear maps); boundary vertices stay on their planes.

usage: python oracle/data/make_box_geo.py m|kind [out.GEO]   (kind: m3, m3w, ...)
"""
import math
import sys

import numpy as np

# local vertex offsets of a hexahedron in the reference's order
HEXV = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]
# local faces (1-based vertex numbers) and the PRM component of the box face
FACES = [((1, 2, 3, 4), ("z", 0)), ((1, 2, 6, 5), ("y", 0)), ((2, 3, 7, 6), ("x", 1)),
         ((3, 4, 8, 7), ("y", 1)), ((4, 1, 5, 8), ("x", 0)), ((5, 6, 7, 8), ("z", 1))]
COMP = {("z", 0): 1, ("x", 0): 2, ("y", 0): 3, ("z", 1): 4, ("x", 1): 5, ("y", 1): 6}
AX = {"x": 0, "y": 1, "z": 2}


def box_geo(m, box=(1.0, 1.0, 1.0), deform=None):
    """GEO text of the box [0,box] split into m^3 (or m = (mx, my, mz)) bricks.

    deform(i, j, k) -> (dx, dy, dz) in units of the macro cell size moves the
    interior lattice vertex (i, j, k)."""
    mx, my, mz = (m, m, m) if isinstance(m, int) else m
    nv = (mx + 1) * (my + 1) * (mz + 1)
    vid = lambda i, j, k: 1 + i + (mx + 1) * (j + (my + 1) * k)
    lines = ["Grobgitter 3D", f"Box of {mx} x {my} x {mz} macro bricks, explicit vertex coordinates"]
    cells = [(i, j, k) for k in range(mz) for j in range(my) for i in range(mx)]
    bfaces = []
    for c, (i, j, k) in enumerate(cells):
        for f, (verts, (ax, side)) in enumerate(FACES):
            a = AX[ax]
            pos = (i, j, k)[a] + side
            lim = (mx, my, mz)[a]
            if (side == 0 and pos == 0) or (side == 1 and pos == lim):
                bfaces.append((c, f, verts))
    lines.append(f"  {len(cells):6d} {nv:6d}    4    8 {len(bfaces):6d}         NEL NVT NVpF NVpEL NBF")
    lines.append("DCORVG (comp, t, s  or  x, y, z)")
    for k in range(mz + 1):
        for j in range(my + 1):
            for i in range(mx + 1):
                d = (0.0, 0.0, 0.0)
                interior = 0 < i < mx and 0 < j < my and 0 < k < mz
                if deform is not None and interior:
                    d = deform(i, j, k)
                x = (i + d[0]) * box[0] / mx
                y = (j + d[1]) * box[1] / my
                z = (k + d[2]) * box[2] / mz
                lines.append(f" {x!r} {y!r} {z!r}")
    lines.append("KVERT")
    for (i, j, k) in cells:
        lines.append(" " + " ".join(f"{vid(i + a, j + b, k + c):5d}" for a, b, c in HEXV))
    lines.append("KNPR")
    knpr = ["0"] * nv
    for s in range(0, nv, 16):
        lines.append(" " + " ".join(f"{x:>3}" for x in knpr[s:s + 16]))
    lines.append("BoundFaces")
    for c, f, verts in bfaces:
        i, j, k = cells[c]
        lines.append(" " + " ".join(f"{vid(i + HEXV[v - 1][0], j + HEXV[v - 1][1], k + HEXV[v - 1][2]):5d}"
                                    for v in verts))
    lines.append("FaceParam (cell, locface, boundary part/component)")
    for c, f, verts in bfaces:
        lines.append(f" {c + 1:5d} {f + 1:3d}   1 {COMP[FACES[f][1]]:3d}")
    return "\n".join(lines) + "\n"


def wavy(i, j, k):
    """Displacement of interior macro vertex (i, j, k), in macro cell sizes:
    up to 0.15 per axis, so that neighbouring vertices keep their order and
    every cell stays a valid (non-parallelepiped) hexahedron."""
    return (0.15 * math.sin(2.1 * i + 1.3 * j + 0.7 * k),
            0.15 * math.sin(0.9 * i + 2.3 * j + 1.1 * k + 1.0),
            0.15 * math.sin(1.7 * i + 0.5 * j + 2.9 * k + 2.0))


DEFORMS = {"w": wavy}


def geo_for_kind(kind):
    """GEO text for a mesh kind "m<m>" (bricks) or "m<m><d>" (interior
    vertices moved by DEFORMS[d])."""
    body = kind[1:]
    d = body[-1] if body[-1] in DEFORMS else None
    m = int(body[:-1] if d else body)
    return m, box_geo(m, deform=DEFORMS[d] if d else None)


def macro_vertices(text):
    """(m+1)^3 x 3 vertex coordinates of a GEO text as box_geo writes it."""
    lines = text.split("\n")
    nv = int(lines[2].split()[1])
    return np.array([[float(x) for x in lines[4 + k].split()] for k in range(nv)])


def main():
    kind = sys.argv[1] if sys.argv[1].startswith("m") else "m" + sys.argv[1]
    m, text = geo_for_kind(kind)
    out = sys.argv[2] if len(sys.argv) > 2 else f"box_{kind}.GEO"
    with open(out, "w") as f:
        f.write(text)


if __name__ == "__main__":
    main()
