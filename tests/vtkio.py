"""Minimal legacy-VTK (ASCII, tetrahedra) reader used by the tests."""
from __future__ import annotations

import numpy as np


def read_vtk(path):
    """Returns dict with xyz (n,3), cells (m,4) in file order, cell_types,
    elem_tag, and point_data {name: array}."""
    with open(path) as f:
        lines = f.read().split("\n")
    out = {"point_data": {}}
    i = 0
    n = len(lines)
    while i < n:
        ln = lines[i].strip()
        if ln.startswith("POINTS"):
            npts = int(ln.split()[1])
            vals = " ".join(lines[i + 1:i + 1 + npts]).split()
            out["xyz"] = np.array(vals, dtype=np.float64).reshape(npts, 3)
            i += npts + 1
            continue
        if ln.startswith("CELLS"):
            nc = int(ln.split()[1])
            rows = [list(map(int, lines[i + 1 + k].split())) for k in range(nc)]
            out["cells"] = np.array([r[1:5] for r in rows], dtype=np.int64)
            i += nc + 1
            continue
        if ln.startswith("CELL_TYPES"):
            nc = int(ln.split()[1])
            out["cell_types"] = np.array([int(lines[i + 1 + k]) for k in range(nc)])
            i += nc + 1
            continue
        if ln.startswith("CELL_DATA"):
            nc = int(ln.split()[1])
            # SCALARS elemTag int 1 / LOOKUP_TABLE default
            out["elem_tag"] = np.array([int(lines[i + 3 + k].split()[0]) for k in range(nc)])
            i += nc + 3
            continue
        if ln.startswith("POINT_DATA"):
            npd = int(ln.split()[1])
            i += 1
            while i < n and lines[i].strip().startswith("SCALARS"):
                name = lines[i].split()[1]
                vals = [float(lines[i + 2 + k].split()[0]) for k in range(npd)]
                out["point_data"][name] = np.array(vals)
                i += npd + 2
            continue
        i += 1
    return out
