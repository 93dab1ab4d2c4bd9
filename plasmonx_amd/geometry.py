"""Deterministic synthetic structures (the reference's geometry generator is
an empty submodule, .gitmodules:1-6): fcc spheres and core-shell spheres,
honeycomb graphene flakes, icosphere and hemispherically capped rod meshes.
Coordinates in Angstrom unless stated.
"""
import math

import numpy as np


def fcc_sphere(a, radius, center_site=True):
    """fcc lattice (constant a, a site at the origin) inside |r| <= radius."""
    m = int(math.ceil(radius / a)) + 1
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    g = np.arange(-m, m + 1)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), axis=-1).reshape(-1, 3)
    pts = (cells[:, None, :] + base[None, :, :]).reshape(-1, 3) * a
    r = np.linalg.norm(pts, axis=1)
    pts = pts[r <= radius + 1e-9]
    order = np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0]))
    return pts[order]


def core_shell(a, r_core, r_shell, core="Au", shell="Ag"):
    pts = fcc_sphere(a, r_shell)
    r = np.linalg.norm(pts, axis=1)
    sym = [core if x <= r_core else shell for x in r]
    return sym, pts


def honeycomb(shape, size, a_cc=1.42, rng=None):
    """Graphene flake: shape in {'disk','ribbon','triangle','ring'}, size ~ the
    number of atoms wanted.  Returns (n, 3) with z = 0."""
    a1 = np.array([1.5 * a_cc, math.sqrt(3) / 2 * a_cc])
    a2 = np.array([1.5 * a_cc, -math.sqrt(3) / 2 * a_cc])
    basis = [np.zeros(2), np.array([a_cc, 0.0])]
    area_per_atom = 3 * math.sqrt(3) / 4 * a_cc ** 2
    R = math.sqrt(size * area_per_atom / math.pi) * (1.15 if shape == "ring" else 1.0)
    m = int(R / a_cc) + 4
    ij = np.stack(np.meshgrid(np.arange(-m, m + 1), np.arange(-m, m + 1), indexing="ij"), -1).reshape(-1, 2)
    pts = np.concatenate([ij[:, :1] * a1 + ij[:, 1:] * a2 + b for b in basis])
    x, y = pts[:, 0], pts[:, 1]
    if shape == "disk":
        keep = x ** 2 + y ** 2 <= R ** 2
    elif shape == "ring":
        rr = x ** 2 + y ** 2
        keep = (rr <= R ** 2) & (rr >= (0.45 * R) ** 2)
    elif shape == "ribbon":
        w = R * 0.6
        L = size * area_per_atom / (2 * w)
        keep = (np.abs(y) <= w) & (np.abs(x) <= L / 2)
    elif shape == "triangle":
        s = math.sqrt(4 * size * area_per_atom / math.sqrt(3))
        h = s * math.sqrt(3) / 2
        keep = (y >= -h / 3) & (math.sqrt(3) * x + y <= 2 * h / 3) & (-math.sqrt(3) * x + y <= 2 * h / 3)
    else:
        raise ValueError(shape)
    p = pts[keep]
    p = p[np.lexsort((p[:, 1], p[:, 0]))]
    return np.column_stack([p, np.zeros(len(p))])


def graphene_batch(n_flakes=96, seed=20261203, lo=500, hi=12000):
    """The batch of seeded flakes: list of dicts (shape, atoms xyz, E_F eV)."""
    rng = np.random.default_rng(seed)
    shapes = ["disk", "ribbon", "triangle", "ring"]
    out = []
    for k in range(n_flakes):
        shape = shapes[int(rng.integers(0, 4))]
        size = int(rng.integers(lo, hi + 1))
        ef = float(rng.uniform(0.2, 0.6))
        out.append({"shape": shape, "xyz": honeycomb(shape, size), "fermi_energy": ef})
    return out


def icosphere(radius, level, seed=None):
    """Icosahedron refined `level` times, projected on the sphere (20*4^level
    triangles, outward orientation), optionally rotated by a seeded random
    rotation.  Returns (n, 3, 3)."""
    t = (1 + 5 ** 0.5) / 2
    v = [(-1, t, 0), (1, t, 0), (-1, -t, 0), (1, -t, 0), (0, -1, t), (0, 1, t), (0, -1, -t), (0, 1, -t),
         (t, 0, -1), (t, 0, 1), (-t, 0, -1), (-t, 0, 1)]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    verts = [np.array(p, dtype=np.float64) / np.linalg.norm(p) for p in v]
    faces = list(f)
    for _ in range(level):
        cache = {}

        def mid(i, j):
            key = (min(i, j), max(i, j))
            if key not in cache:
                p = verts[i] + verts[j]
                verts.append(p / np.linalg.norm(p))
                cache[key] = len(verts) - 1
            return cache[key]

        nf = []
        for a, b, c in faces:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        faces = nf
    V = np.array(verts) * radius
    if seed is not None:
        rng = np.random.default_rng(seed)
        q, r = np.linalg.qr(rng.normal(size=(3, 3)))
        q = q @ np.diag(np.sign(np.diag(r)))
        if np.linalg.det(q) < 0:
            q[:, 0] = -q[:, 0]
        V = V @ q.T
    tris = V[np.array(faces)]
    # outward orientation: (v2-v1) x (v3-v1) along the centroid
    n = np.cross(tris[:, 1] - tris[:, 0], tris[:, 2] - tris[:, 0])
    flip = np.einsum("ij,ij->i", n, tris.mean(axis=1)) < 0
    tris[flip] = tris[flip][:, [0, 2, 1]]
    return tris


def capped_rod(length, radius, n_ring=40, n_len=None, n_cap=10):
    """Cylinder of total length `length` with hemispherical caps along z:
    quadrilateral strips split in triangles.  Returns (n, 3, 3) outward."""
    body = length - 2 * radius
    if n_len is None:
        n_len = max(1, int(round(body / (2 * math.pi * radius / n_ring))))
    rings = []
    # south cap (z < -body/2), body rings, north cap
    for k in range(n_cap, 0, -1):
        th = (k / n_cap) * (math.pi / 2)
        rings.append((radius * math.cos(th), -body / 2 - radius * math.sin(th)))
    for k in range(n_len + 1):
        rings.append((radius, -body / 2 + body * k / n_len))
    for k in range(1, n_cap + 1):
        th = (k / n_cap) * (math.pi / 2)
        rings.append((radius * math.cos(th), body / 2 + radius * math.sin(th)))
    tris = []
    phis = np.linspace(0, 2 * math.pi, n_ring, endpoint=False)

    def P(r, z, p):
        return np.array([r * math.cos(p), r * math.sin(p), z])

    for (r0, z0), (r1, z1) in zip(rings[:-1], rings[1:]):
        for m in range(n_ring):
            p0, p1 = phis[m], phis[(m + 1) % n_ring]
            a, b, c, d = P(r0, z0, p0), P(r0, z0, p1), P(r1, z1, p1), P(r1, z1, p0)
            if r0 < 1e-12:
                tris.append((a, c, d))
            elif r1 < 1e-12:
                tris.append((a, b, d))
            else:
                tris.append((a, b, c))
                tris.append((a, c, d))
    T = np.array(tris)
    n = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    cen = T.mean(axis=1)
    axis_pt = np.column_stack([np.zeros(len(T)), np.zeros(len(T)), np.clip(cen[:, 2], -body / 2, body / 2)])
    flip = np.einsum("ij,ij->i", n, cen - axis_pt) < 0
    T[flip] = T[flip][:, [0, 2, 1]]
    return T


def fcc_rod(a, length, radius):
    """fcc lattice (constant a, a site at the origin) inside a cylinder of
    total length `length` along z with hemispherical caps of radius
    `radius` (all in the units of a)."""
    m_xy = int(math.ceil(radius / a)) + 1
    m_z = int(math.ceil(length / 2 / a)) + 1
    base = np.array([[0, 0, 0], [0.5, 0.5, 0], [0.5, 0, 0.5], [0, 0.5, 0.5]])
    gx = np.arange(-m_xy, m_xy + 1)
    gz = np.arange(-m_z, m_z + 1)
    cells = np.stack(np.meshgrid(gx, gx, gz, indexing="ij"), axis=-1).reshape(-1, 3)
    pts = (cells[:, None, :] + base[None, :, :]).reshape(-1, 3) * a
    half = max(0.0, length / 2 - radius)
    zc = np.clip(pts[:, 2], -half, half)
    d = np.sqrt(pts[:, 0] ** 2 + pts[:, 1] ** 2 + (pts[:, 2] - zc) ** 2)
    pts = pts[d <= radius + 1e-9]
    order = np.lexsort((pts[:, 2], pts[:, 1], pts[:, 0]))
    return pts[order]
