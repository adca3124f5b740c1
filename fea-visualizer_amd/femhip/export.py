"""PLY export of a solved surface with threshold clipping.

Surface::export_to_ply (src/Utils/Surface.cpp:138-285) runs in the library
(fem_export_ply, or fem_solver_export_ply for the solver's current field):
vertices are extruded along their normals by value * max(0, extrusion),
triangles clipped at `threshold`, closed by a projected copy for the CLOSED
and MIRRORED mesh types, positions de-duplicated in first-use order, each
vertex coloured by the Viridis polynomial.  This module passes arrays across
the C ABI and reads files back.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib as L

OPEN, CLOSED, MIRRORED = L.FEM_MESH_OPEN, L.FEM_MESH_CLOSED, L.FEM_MESH_MIRRORED  # MeshType (Surface.hpp:25-29)


def export_ply(path: str, xyz, tri, normals, values, vertex_extrusion: float, threshold: float,
               mesh_type: int = OPEN):
    """Write the clipped, extruded surface as ASCII PLY (fem_export_ply)."""
    lib = L.load()
    x = np.ascontiguousarray(xyz, dtype=np.float32).reshape(-1, 3)
    t = np.ascontiguousarray(tri, dtype=np.uint32).reshape(-1, 3)
    n = np.ascontiguousarray(normals, dtype=np.float32).reshape(-1, 3)
    v = np.ascontiguousarray(values, dtype=np.float32).reshape(-1)
    if n.shape != x.shape or v.shape[0] != x.shape[0]:
        raise ValueError("normals and values need one entry per vertex")
    onb = np.zeros(x.shape[0], dtype=np.uint8)
    m = L.fem_mesh(x.shape[0], t.shape[0], L.ptr(x, C.c_float), L.ptr(t, C.c_uint32), L.ptr(onb, C.c_uint8))
    L.check(lib.fem_export_ply(C.byref(m), L.ptr(n, C.c_float), L.ptr(v, C.c_float), float(vertex_extrusion),
                               float(threshold), int(mesh_type), str(path).encode()))


def read_ply(path: str):
    """Positions, colours and faces of an ASCII PLY written by export_ply."""
    with open(path) as f:
        lines = f.read().splitlines()
    nv = int(next(l for l in lines if l.startswith("element vertex")).split()[-1])
    nf = int(next(l for l in lines if l.startswith("element face")).split()[-1])
    h = lines.index("end_header") + 1
    vert = np.array([[float(z) for z in l.split()] for l in lines[h:h + nv]], dtype=np.float64).reshape(-1, 6)
    faces = np.array([[int(z) for z in l.split()[1:]] for l in lines[h + nv:h + nv + nf]], dtype=np.int64)
    return vert[:, :3], vert[:, 3:], faces.reshape(-1, 3)
