"""Synthetic workloads (sizes in micrometres, speeds in micrometres per
microsecond = m/s, times in microseconds), seeded and deterministic."""
from __future__ import annotations

import numpy as np

from . import Mesh

# unit cube 1 cm, 20^3 hexes x 6 tets (48,000 tets, 9,261 vertices), interior
# vertices jittered +-10 % of the spacing, isotropic 0.6 m/s, pacing at the
# origin
CUBE = dict(lx=10000.0, ly=10000.0, lz=10000.0, nx=20, ny=20, nz=20, jitter=0.1,
            angle0=0.0, angle1=0.0, vf=0.6, vc=0.6)
# slab 10 x 10 x 2 cm, 160 x 160 x 13 hexes x 6 tets (1,996,800 tets,
# 362,894 vertices), fibres rotating -60..+60 degrees through the thickness,
# v_fibre 0.6, v_cross 0.24 m/s, five pacing vertices on z = 0
SLAB = dict(lx=100000.0, ly=100000.0, lz=20000.0, nx=160, ny=160, nz=13, jitter=0.1,
            angle0=-60.0, angle1=60.0, vf=0.6, vc=0.24)


def splitmix64(seed):
    """Deterministic stream of 64-bit integers (SplitMix64)."""
    x = seed & 0xFFFFFFFFFFFFFFFF
    while True:
        x = (x + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = x
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        yield z ^ (z >> 31)


def pick(n_choices, k, seed):
    """k distinct indices in [0, n_choices) drawn from SplitMix64(seed)."""
    g = splitmix64(seed)
    out = []
    while len(out) < k:
        v = next(g) % n_choices
        if v not in out:
            out.append(v)
    return out


def scaled(cfg, f):
    """Same physical box with the hex counts divided by f."""
    c = dict(cfg)
    for k in ("nx", "ny", "nz"):
        c[k] = max(2, int(round(cfg[k] / f)))
    return c


def box(cfg, seed) -> Mesh:
    return Mesh.gen_box(seed=seed, **cfg)


def face_vertices(cfg, z_index=0):
    NX, NY = cfg["nx"] + 1, cfg["ny"] + 1
    return np.array([i + NX * (j + NY * z_index) for j in range(NY) for i in range(NX)],
                    dtype=np.int64)


def pacing_sites(cfg, k, seed):
    """k distinct vertices of the z = 0 face, from the seed."""
    face = face_vertices(cfg, 0)
    return face[pick(face.shape[0], k, seed)]


# ---------------------------------------------------------------------------
# batched workload over four meshes of uneven size, built as the cube and the
# slab above: 48k, ~400k, ~1M and ~2M tets (the slabs keep the 625 um
# in-plane spacing and shrink the square)
# ---------------------------------------------------------------------------
MULTI = [
    ("cube_48k", CUBE),
    ("slab_404k", dict(SLAB, nx=72, ny=72, lx=45000.0, ly=45000.0)),
    ("slab_996k", dict(SLAB, nx=113, ny=113, lx=70625.0, ly=70625.0)),
    ("slab_2M", SLAB),
]


def quadrant_regions(mesh: Mesh) -> np.ndarray:
    """Region id 0..3 of each element: quadrant of its centroid in the xy
    plane around the mesh's bounding-box centre."""
    xyz, e2n, _ = mesh.arrays()
    c = xyz[e2n].mean(axis=1)
    mid = 0.5 * (xyz.min(axis=0) + xyz.max(axis=0))
    return ((c[:, 0] > mid[0]).astype(np.uint8) + 2 * (c[:, 1] > mid[1]).astype(np.uint8))


def multi_mesh(parts, seed):
    """Union of box meshes with quadrant regions.  Returns the union mesh,
    the vertex offset of each part and the parts' configs."""
    meshes, offs, off = [], [], 0
    for k, (_, cfg) in enumerate(parts):
        m = box(cfg, seed + k)
        m.set_regions(quadrant_regions(m))
        meshes.append(m)
        offs.append(off)
        off += m.nn
    return Mesh.concat(meshes), offs, [cfg for _, cfg in parts]


def multi_solves(parts, offs, n, seed):
    """n solves spread over the parts in turn: 1-3 pacing vertices on the
    part's z = 0 face and four regional speed factors in [0.5, 1.5]."""
    g = splitmix64(seed)
    out = []
    for i in range(n):
        k = i % len(parts)
        cfg = parts[k]
        npace = 1 + next(g) % 3
        sites = pacing_sites(cfg, npace, next(g)) + offs[k]
        scale = np.array([0.5 + (next(g) >> 11) * 2.0 ** -53 for _ in range(4)], dtype=np.float32)
        out.append(dict(part=k, sites=sites, scale=scale))
    return out


# ---------------------------------------------------------------------------
# prolate-ellipsoid shell, semi-axes 5/3/3 cm outer and 4.2/2.2/2.2 cm inner,
# hexes of 395 um -> ~10M tets; twelve scars of radius 0.3-1.0 cm with
# 0.05 m/s; fibres +60 -> -60 deg transmurally; v_f 0.6 / v_c 0.24 m/s
# ---------------------------------------------------------------------------
SHELL = dict(ao=50000.0, bo=30000.0, ai=42000.0, bi=22000.0, h=395.0, n_scars=12, r0=3000.0,
             r1=10000.0, vf=0.6, vc=0.24, vs=0.05)


def shell(seed, h=None) -> Mesh:
    cfg = dict(SHELL)
    if h is not None:
        cfg["h"] = h
    return Mesh.gen_shell(seed=seed, **cfg)


def inner_surface_sites(mesh: Mesh, k, seed, cfg=None):
    """k distinct vertices within 1.5 hex sides of the inner ellipsoid."""
    cfg = cfg or SHELL
    xyz, e2n, _ = mesh.arrays()
    h = float(np.abs(xyz[e2n[0, 1]] - xyz[e2n[0, 0]]).max())
    r = np.sqrt((xyz[:, 0] / cfg["ai"]) ** 2 + (xyz[:, 1] ** 2 + xyz[:, 2] ** 2) / cfg["bi"] ** 2)
    cand = np.flatnonzero(r <= 1.0 + 1.5 * h / cfg["bi"])
    return cand[pick(cand.shape[0], k, seed)]
