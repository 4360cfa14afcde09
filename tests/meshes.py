"""Triangle meshes for the voxelizer tests (numpy, grid units).

`parse_obj` reads Wavefront OBJ text the way the reference's
ObjLoader::load_from_string does (obj_loader.cpp:21-83: `v x y z`,
`f i j k ...` with 1-based or negative indices, fan triangulation).
"""
import numpy as np

# the unit cube of tests/unit/test_geometry.cpp:32-47 (6 quads)
CUBE_OBJ = """v -1.0 -1.0 -1.0
v  1.0 -1.0 -1.0
v  1.0  1.0 -1.0
v -1.0  1.0 -1.0
v -1.0 -1.0  1.0
v  1.0 -1.0  1.0
v  1.0  1.0  1.0
v -1.0  1.0  1.0
f 1 2 3 4
f 5 8 7 6
f 1 5 6 2
f 2 6 7 3
f 3 7 8 4
f 5 1 4 8
"""


def parse_obj(text):
    verts, tris = [], []
    for line in text.splitlines():
        parts = line.split()
        if not parts or parts[0].startswith("#"):
            continue
        if parts[0] == "v" and len(parts) >= 4:
            verts.append([float(p) for p in parts[1:4]])
        elif parts[0] == "f":
            idx = []
            for tok in parts[1:]:
                i = int(tok.split("/")[0])
                idx.append(i - 1 if i > 0 else len(verts) + i)
            for k in range(1, len(idx) - 1):
                tri = (idx[0], idx[k], idx[k + 1])
                if all(0 <= i < len(verts) for i in tri):
                    tris.append([verts[i] for i in tri])
    return np.array(verts, np.float32), np.array(tris, np.float32).reshape(-1, 3, 3)


def fit_into_box(tris, lo, hi):
    """obj_loader.cpp:185-206 in FP32: uniform scale into [lo, hi], centred."""
    t = np.asarray(tris, np.float32).copy()
    v = t.reshape(-1, 3)
    cmin, cmax = v.min(0), v.max(0)
    size = cmax - cmin
    target = np.float32(hi) - np.float32(lo)
    s = np.float32(np.min(target / size))
    v -= cmin
    v *= s
    v += np.float32(lo) + (target - size * s) * np.float32(0.5)
    return t


def box_mesh(x0, y0, z0, x1, y1, z1, open_face=None):
    """Closed axis-aligned box surface, 12 triangles (10 without open_face:
    0..5 = -x, +x, -y, +y, -z, +z)."""
    c = np.array([[x0, y0, z0], [x1, y0, z0], [x1, y1, z0], [x0, y1, z0],
                  [x0, y0, z1], [x1, y0, z1], [x1, y1, z1], [x0, y1, z1]], np.float32)
    faces = [(0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (3, 7, 6, 2), (0, 3, 2, 1), (4, 5, 6, 7)]
    tris = []
    for f, (a, b, cc, d) in enumerate(faces):
        if f == open_face:
            continue
        tris += [[c[a], c[b], c[cc]], [c[a], c[cc], c[d]]]
    return np.array(tris, np.float32)


def icosphere(subdiv, radius=1.0, center=(0.0, 0.0, 0.0)):
    """Subdivided icosahedron with every vertex on the sphere."""
    p = (1.0 + 5 ** 0.5) / 2
    v = [[-1, p, 0], [1, p, 0], [-1, -p, 0], [1, -p, 0], [0, -1, p], [0, 1, p], [0, -1, -p], [0, 1, -p],
         [p, 0, -1], [p, 0, 1], [-p, 0, -1], [-p, 0, 1]]
    v = [np.array(x, float) / np.linalg.norm(x) for x in v]
    f = [(0, 11, 5), (0, 5, 1), (0, 1, 7), (0, 7, 10), (0, 10, 11), (1, 5, 9), (5, 11, 4), (11, 10, 2), (10, 7, 6),
         (7, 1, 8), (3, 9, 4), (3, 4, 2), (3, 2, 6), (3, 6, 8), (3, 8, 9), (4, 9, 5), (2, 4, 11), (6, 2, 10),
         (8, 6, 7), (9, 8, 1)]
    for _ in range(subdiv):
        cache, nf = {}, []

        def mid(a, b):
            k = (min(a, b), max(a, b))
            if k not in cache:
                m = v[a] + v[b]
                v.append(m / np.linalg.norm(m))
                cache[k] = len(v) - 1
            return cache[k]

        for a, b, c in f:
            ab, bc, ca = mid(a, b), mid(b, c), mid(c, a)
            nf += [(a, ab, ca), (b, bc, ab), (c, ca, bc), (ab, bc, ca)]
        f = nf
    V = np.array(v) * radius + np.asarray(center, float)
    return V[np.array(f)].astype(np.float32)


def torus(R, r, nu, nv, center, axes=None):
    """Closed torus surface (major radius R, minor r) around `center`."""
    u = np.linspace(0, 2 * np.pi, nu, endpoint=False)
    w = np.linspace(0, 2 * np.pi, nv, endpoint=False)
    uu, ww = np.meshgrid(u, w, indexing="ij")
    P = np.stack([(R + r * np.cos(ww)) * np.cos(uu), (R + r * np.cos(ww)) * np.sin(uu), r * np.sin(ww)], -1)
    if axes is not None:
        P = P @ np.asarray(axes, float)
    P = P + np.asarray(center, float)
    tris = []
    for i in range(nu):
        for j in range(nv):
            a, b = P[i, j], P[(i + 1) % nu, j]
            c, d = P[(i + 1) % nu, (j + 1) % nv], P[i, (j + 1) % nv]
            tris += [[a, b, c], [a, c, d]]
    return np.array(tris, np.float32)


def clip_overlap(tri, half):
    """Exact (float64) triangle / centred box [-half, half]^3 intersection by
    Sutherland-Hodgman clipping of the triangle against the six box planes:
    True when a non-empty piece (a point counts) survives."""
    poly = [np.asarray(p, float) for p in tri]
    for axis in range(3):
        for sgn in (1.0, -1.0):
            out = []
            n = len(poly)
            for k in range(n):
                a, b = poly[k], poly[(k + 1) % n]
                da, db = half - sgn * a[axis], half - sgn * b[axis]  # >= 0 inside
                if da >= 0:
                    out.append(a)
                if (da >= 0) != (db >= 0):
                    t = da / (da - db)
                    out.append(a + t * (b - a))
            poly = out
            if not poly:
                return False
    return True
