"""PLY export of a solved surface with threshold clipping.

Mirrors Surface::export_to_ply (src/Utils/Surface.cpp:138-285): vertices are
extruded along their normals by value * max(0, extrusion); triangles are
clipped at `threshold` (vertices whose value exceeds it by at least 1e-9 are
kept, crossing edges are cut by linear interpolation); with a CLOSED or
MIRRORED mesh type the kept part is closed by a projected copy.  Positions
are de-duplicated exactly (float32 bit equality) in first-use order, and
every vertex carries the colour of its value under the colour map.

This is file output: it runs on the host, on the values the GPU solver
returns (HIPSolver.get_state).  Arithmetic is float32, as the reference's
glm::vec3.
"""
from __future__ import annotations

from typing import Callable, Optional, Sequence

import numpy as np

OPEN, CLOSED, MIRRORED = 0, 1, 2  # enum class MeshType (include/Utils/Surface.hpp:25-29)

# Viridis as an order-6 polynomial per channel: the reference's default colour
# map (src/Utils/AssetStorage.cpp:141-152, from shadertoy.com/view/Nd3fR2).
VIRIDIS = np.array([
    [0.274344, 0.004462, 0.331359],
    [0.108915, 1.397291, 1.388110],
    [-0.319631, 0.243490, 0.156419],
    [-4.629188, -5.882803, -19.646115],
    [6.181719, 14.388598, 57.442181],
    [4.876952, -13.955112, -66.125783],
    [-5.513165, 4.709245, 26.582180],
], dtype=np.float32)


def polynomial_color_map(coeffs: np.ndarray) -> Callable[[float], np.ndarray]:
    """ColorMap::get_color (src/Utils/ColorMap.cpp:12-14): Horner evaluation."""
    c = np.asarray(coeffs, dtype=np.float32)

    def color(t):
        t = np.float32(t)
        acc = c[6]
        for k in range(5, -1, -1):
            acc = c[k] + t * acc
        return acc
    return color


def _mix(x, y, a):
    a = np.float32(a)
    return x * (np.float32(1.0) - a) + y * a


def clip_surface(xyz, tri, normals, values, vertex_extrusion: float, threshold: float, mesh_type: int = OPEN):
    """Clipped positions, their values and triangles (Surface.cpp:138-255)."""
    f = np.float32
    V = np.asarray(xyz, dtype=f).reshape(-1, 3)
    N = np.asarray(normals, dtype=f).reshape(-1, 3)
    val = np.asarray(values, dtype=f).reshape(-1)
    T = np.asarray(tri, dtype=np.int64).reshape(-1, 3)
    ext, thr = f(vertex_extrusion), f(threshold)
    positions, pvalues, faces = [], [], []
    index = {}

    def add(pos, value):
        key = pos.astype(f).tobytes()
        if key not in index:
            index[key] = len(positions)
            positions.append(pos.astype(f))
            pvalues.append(f(value))
        return index[key]

    def extruded(i):
        return val[i] * max(f(0.0), ext) * N[i] + V[i]

    def projected(i):
        scale = (thr - val[i] * ext) if mesh_type == MIRRORED else (thr * ext)
        return f(scale) * N[i] + V[i]

    for t in T:
        above = [bool(float(val[i]) - float(thr) >= 1e-9) for i in t]
        count = sum(above)
        if count == 1:
            ai = above.index(True)
            clipped = [0, 0, 0]
            for i in range(3):
                if i != ai:
                    w = (thr - val[t[i]]) / (val[t[ai]] - val[t[i]])
                    clipped[i] = add(_mix(extruded(t[i]), extruded(t[ai]), w), thr)
                else:
                    clipped[i] = add(extruded(t[i]), val[t[ai]])
            faces.append(tuple(clipped))
            if mesh_type != OPEN:
                proj = [clipped[i] if i != ai else add(projected(t[i]), thr) for i in range(3)]
                faces.append(tuple(proj))
        elif count == 2:
            a1 = 0 if above[0] else 1
            a2 = 1 if (above[0] and above[1]) else 2
            b = 3 - (a1 + a2)
            p1, p2 = extruded(t[a1]), extruded(t[a2])
            w1 = (thr - val[t[b]]) / (val[t[a1]] - val[t[b]])
            w2 = (thr - val[t[b]]) / (val[t[a2]] - val[t[b]])
            q1 = _mix(extruded(t[b]), p1, w1)
            q2 = _mix(extruded(t[b]), p2, w2)
            i1, i2 = add(p1, val[t[a1]]), add(p2, val[t[a2]])
            j1, j2 = add(q1, thr), add(q2, thr)
            faces.append((j1, i1, j2))
            faces.append((j2, i1, i2))
            if mesh_type != OPEN:
                k1 = add(projected(t[a1]), val[t[a1]])
                k2 = add(projected(t[a2]), val[t[a2]])
                faces.append((j1, k1, j2))
                faces.append((j2, k1, k2))
        elif count == 3:
            clipped = tuple(add(extruded(i), val[i]) for i in t)
            faces.append(clipped)
            if mesh_type != OPEN:
                faces.append(tuple(add(projected(i), thr) for i in t))
    return (np.array(positions, dtype=f).reshape(-1, 3), np.array(pvalues, dtype=f),
            np.array(faces, dtype=np.int64).reshape(-1, 3))


def _num(x) -> str:
    """Shortest round-trip text of a float32 (std::format's "{}" prints
    integral values without a fractional part)."""
    s = str(np.float32(x))
    return s[:-2] if s.endswith(".0") else s


def export_ply(path: str, xyz, tri, normals, values, vertex_extrusion: float, threshold: float,
               mesh_type: int = OPEN, color_map: Optional[Callable] = None):
    """Write the clipped, extruded surface as ASCII PLY (Surface.cpp:257-285).
    Returns (positions, values, faces)."""
    cmap = color_map or polynomial_color_map(VIRIDIS)
    pos, pval, faces = clip_surface(xyz, tri, normals, values, vertex_extrusion, threshold, mesh_type)
    with open(path, "w") as f:
        f.write("ply\nformat ascii 1.0\n")
        f.write(f"element vertex {len(pos)}\n")
        for q in ("x", "y", "z", "red", "green", "blue"):
            f.write(f"property float {q}\n")
        f.write(f"element face {len(faces)}\n")
        f.write("property list uchar uint vertex_indices\nend_header\n")
        for p, v in zip(pos, pval):
            c = cmap(v)
            f.write(" ".join(_num(z) for z in (p[0], p[1], p[2], c[0], c[1], c[2])) + "\n")
        for a, b, c in faces:
            f.write(f"3 {a} {b} {c}\n")
    return pos, pval, faces


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
