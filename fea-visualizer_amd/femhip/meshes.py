"""Surface meshes for the FEM backend: synthetic generators and the OBJ reader.

A mesh is a dict with ``xyz`` (float32 [nv, 3]), ``tri`` (uint32 [T, 3]) and
``on_boundary`` (uint8 [nv]) -- the fields the reference's FEMContext reads
from ``Surface`` (include/Utils/Surface.hpp:36-40,50).

Boundary detection follows Surface::init_from_obj (src/Utils/Surface.cpp:104-127):
a vertex is on the boundary when it belongs to an edge used by exactly one
triangle.
"""
from __future__ import annotations

import numpy as np

__all__ = [
    "boundary_from_edges",
    "structured_square",
    "jittered_square",
    "icosphere",
    "read_obj",
    "mesh_stats",
]


def boundary_from_edges(tri: np.ndarray, nv: int) -> np.ndarray:
    """Edge-count boundary rule of Surface.cpp:104-127.

    Every directed triangle edge (a, b) increments the counts of both (a, b)
    and (b, a); endpoints of an edge whose count is 1 are boundary vertices.
    """
    tri = np.asarray(tri, dtype=np.int64)
    onb = np.zeros(nv, dtype=np.uint8)
    if tri.size == 0:
        return onb
    a = np.concatenate([tri[:, 0], tri[:, 1], tri[:, 2]])
    b = np.concatenate([tri[:, 1], tri[:, 2], tri[:, 0]])
    lo = np.minimum(a, b)
    hi = np.maximum(a, b)
    key = lo * np.int64(nv) + hi
    uniq, counts = np.unique(key, return_counts=True)
    once = uniq[counts == 1]
    onb[once // nv] = 1
    onb[once % nv] = 1
    return onb


def _pack(xyz, tri, onb=None):
    xyz = np.ascontiguousarray(xyz, dtype=np.float32)
    tri = np.ascontiguousarray(tri, dtype=np.uint32)
    if onb is None:
        onb = boundary_from_edges(tri, xyz.shape[0])
    return {"xyz": xyz, "tri": tri, "on_boundary": np.ascontiguousarray(onb, dtype=np.uint8)}


def structured_square(n: int):
    """Unit square, (n+1)^2 vertices (i/n, j/n, 0), vertex id j*(n+1)+i, each
    cell split along (i,j)-(i+1,j+1) with counter-clockwise triangles."""
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    xyz = np.stack([i.ravel() / n, j.ravel() / n, np.zeros((n + 1) ** 2)], axis=1)
    ci, cj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    v00 = (cj * (n + 1) + ci).ravel()
    v10 = v00 + 1
    v01 = v00 + (n + 1)
    v11 = v01 + 1
    t1 = np.stack([v00, v10, v11], axis=1)
    t2 = np.stack([v00, v11, v01], axis=1)
    tri = np.stack([t1, t2], axis=1).reshape(-1, 3)
    return _pack(xyz, tri)


def jittered_square(n: int, seed: int, jitter: float = 0.25):
    """(n+1)^2 lattice on [0,1]^2, interior vertices moved by U(-jitter*h,
    +jitter*h) in x and y (numpy PCG64, `seed`), boundary fixed.  Each quad is
    split along its Delaunay diagonal: (i,j)-(i+1,j+1) unless the two angles
    opposite it sum to more than pi (ties keep that diagonal)."""
    h = 1.0 / n
    rng = np.random.Generator(np.random.PCG64(seed))
    i, j = np.meshgrid(np.arange(n + 1), np.arange(n + 1), indexing="xy")
    x = i.ravel() * h
    y = j.ravel() * h
    d = rng.uniform(-jitter * h, jitter * h, size=(2, (n + 1) ** 2))
    interior = ((i > 0) & (i < n) & (j > 0) & (j < n)).ravel()
    x = np.where(interior, x + d[0], x)
    y = np.where(interior, y + d[1], y)
    xyz = np.stack([x, y, np.zeros_like(x)], axis=1).astype(np.float32)
    P = xyz[:, :2].astype(np.float64)
    ci, cj = np.meshgrid(np.arange(n), np.arange(n), indexing="xy")
    v00 = (cj * (n + 1) + ci).ravel()
    v10 = v00 + 1
    v01 = v00 + (n + 1)
    v11 = v01 + 1

    def angle(at, p, q):
        u = P[p] - P[at]
        w = P[q] - P[at]
        c = (u * w).sum(1) / (np.linalg.norm(u, axis=1) * np.linalg.norm(w, axis=1))
        return np.arccos(np.clip(c, -1.0, 1.0))

    opp = angle(v10, v00, v11) + angle(v01, v00, v11)
    keep = opp <= np.pi
    ta = np.where(keep[:, None], np.stack([v00, v10, v11], 1), np.stack([v00, v10, v01], 1))
    tb = np.where(keep[:, None], np.stack([v00, v11, v01], 1), np.stack([v10, v11, v01], 1))
    tri = np.stack([ta, tb], axis=1).reshape(-1, 3)
    return _pack(xyz, tri)


def icosphere(level: int, radius: float = 1.0):
    """Icosahedron refined `level` times by edge-midpoint subdivision and
    projected to the sphere.  Vertex order: the 12 icosahedron vertices, then
    per level the new midpoints in ascending (min, max) parent-id order."""
    t = (1.0 + 5.0 ** 0.5) / 2.0
    V = np.array(
        [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0],
         [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
         [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    F = np.array(
        [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11],
         [1, 5, 9], [5, 11, 4], [11, 10, 2], [10, 7, 6], [7, 1, 8],
         [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9],
         [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        nv = V.shape[0]
        e = np.concatenate([F[:, [0, 1]], F[:, [1, 2]], F[:, [2, 0]]])
        lo = e.min(1)
        hi = e.max(1)
        key = lo * nv + hi
        uniq, inv = np.unique(key, return_inverse=True)
        mid = (V[uniq // nv] + V[uniq % nv]) * 0.5
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        m = inv.reshape(3, -1) + nv  # midpoint ids of edges (01, 12, 20)
        m01, m12, m20 = m[0], m[1], m[2]
        a, b, c = F[:, 0], F[:, 1], F[:, 2]
        F = np.stack([
            np.stack([a, m01, m20], 1),
            np.stack([m01, b, m12], 1),
            np.stack([m20, m12, c], 1),
            np.stack([m01, m12, m20], 1)], axis=1).reshape(-1, 3)
        V = np.concatenate([V, mid])
    return _pack(V * radius, F)


def read_obj(path: str):
    """Surface::init_from_obj (src/Utils/Surface.cpp:57-133) through the
    library's reader (fem_surface_from_obj): the mesh dict plus the vertex
    ``normals`` (float32 [nv, 3]).  Faces are triangulated (quads along the
    shorter diagonal, larger polygons fanned), every object repeats the
    vertex list as the reference does, boundary by the edge-count rule."""
    import ctypes as C

    from . import _lib as L
    lib = L.load()
    h = C.c_void_p()
    L.check(lib.fem_surface_from_obj(str(path).encode(), C.byref(h)))
    try:
        m = L.fem_mesh()
        nrm = C.POINTER(C.c_float)()
        nb = C.c_int32()
        L.check(lib.fem_surface_get(h, C.byref(m), C.byref(nrm), C.byref(nb)))
        nv, nt = m.num_vertices, m.num_triangles
        xyz = np.ctypeslib.as_array(m.xyz, shape=(nv, 3)).copy() if nv else np.zeros((0, 3), np.float32)
        tri = np.ctypeslib.as_array(m.triangles, shape=(nt, 3)).copy() if nt else np.zeros((0, 3), np.uint32)
        onb = np.ctypeslib.as_array(m.on_boundary, shape=(nv,)).copy() if nv else np.zeros(0, np.uint8)
        normals = np.ctypeslib.as_array(nrm, shape=(nv, 3)).copy() if nv else np.zeros((0, 3), np.float32)
    finally:
        lib.fem_surface_destroy(h)
    out = _pack(xyz, tri, onb)
    out["normals"] = normals
    return out


def mesh_stats(mesh, bc: int = 0):
    """(nv, T, boundary count, N) for a mesh under boundary condition bc."""
    nv = mesh["xyz"].shape[0]
    nb = int(mesh["on_boundary"].sum())
    n = nv - (nb if bc == 0 else 0)
    return nv, mesh["tri"].shape[0], nb, n
