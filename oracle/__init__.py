"""Python view of the oracle (CPU restatement of the reference FEM hot path).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke entry and the CPU-baseline
leg of bench.py as the checker.  The product (fea-visualizer_amd/femhip) never
imports it.  See fem_oracle.h for the reference file:line each routine
follows.

Parity status of each function (how it is pinned, see tests/test_oracle.py):
  idx_map, pattern        brute-force set construction on small meshes and
                          closed-form counts on the shipped icosphere/tube
  assemble (K, M, A)      closed forms: 5-point stencil on the structured
                          square, cotangent formula, row sums K.1 = 0,
                          sum(M) = area, A.1 = 0 on interior rows, A = area/3 v.grad
  cg / bicgstab           scipy direct solve (FP64) on the same system
  step (heat, advdiff,    scipy direct solve of the same linear system per step;
  wave, grayscott)        heat mass conservation under Neumann; wave energy decay
  newmark                 discrete energy conservation (beta=1/4, gamma=1/2)
  adjoint                 central finite differences and a dot-product test
  gl_step                 its CG reaches the direct solution given enough
                          iterations (heat), and it follows the GL stopping rule
  smooth_normals          closed form on flat meshes (2 x incident area along z),
                          outward normals on the sphere
  port_run_f              same steps as orc_step_f (FP32) up to CG tolerance
  read_obj                hand-built OBJ files with known arrays (triangles,
                          quads split either way, a pentagon, negative and
                          v//n indices, two objects), the shipped icosphere's
                          golden arrays and the shipped tube's counts; the
                          quad-split tie rule of tinyobjloader itself is
                          parity unpinned (the library is absent)
  ply_clip, viridis       hand-computed clipping cases (one, two, three
                          vertices above the threshold, closed and mirrored
                          caps), Viridis endpoints
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libfem_oracle.so")

HEAT, WAVE, ADVDIFF, GRAYSCOTT, NEWMARK = 0, 1, 2, 3, 4
DIRICHLET, NEUMANN = 0, 1
CG, BICGSTAB = 0, 1

_lib = None


def build():
    """Compile the restatement with the reference's release flags."""
    res = subprocess.run(["make", "-s", "-C", HERE], capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError("oracle build failed:\n" + res.stdout + res.stderr)
    return LIB


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        build()
    lib = C.CDLL(LIB)
    d, f, i32, i64 = C.c_double, C.c_float, C.c_int32, C.c_int64
    P = C.c_void_p
    sig = {
        "orc_idx_map": (i32, [i32, P, i32, P]),
        "orc_pattern": (i64, [i32, i32, P, P, i32, P, i32, P, P, i64]),
        "orc_assemble_d": (None, [i32, i32, P, P, P, i32, P, P, P, P, P, i32, P, P, P, P]),
        "orc_assemble_f": (None, [i32, i32, P, P, P, i32, P, P, P, P, P, i32, P, P, P, P]),
        "orc_spmv_d": (None, [i32, P, P, P, P, P, d, d, d, P, P]),
        "orc_cg_d": (i32, [i32, P, P, P, P, P, d, i32, P]),
        "orc_bicgstab_d": (i32, [i32, P, P, P, P, P, d, i32, P]),
        "orc_step_d": (i32, [i32, P, P, P, P, P, i32, P, P, P, P, i32, d, i32, i32]),
        "orc_step_f": (i32, [i32, P, P, P, P, P, i32, P, P, P, P, i32, d, i32, i32]),
        "orc_newmark_init_d": (i32, [i32, P, P, P, P, d, P, P, d]),
        "orc_adjoint_d": (d, [i32, i32, P, P, P, i32, P, P, P, P, P, d, i32, P, P, d, P, P, P]),
        "orc_port_run_f": (i64, [i32, P, P, P, P, P, P, P, P, P, i32, P, P, P, i32]),
        "orc_unstable_f": (i32, [i32, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class System:
    """idx_map, CSR pattern and assembled K, M, A of a mesh (double or float)."""

    def __init__(self, mesh, bc=DIRICHLET, velocity=(1.0, 0.0, 0.0), velocity_e=None,
                 normalize=True, kappa_e=None, dtype=np.float64):
        lib = load()
        self.mesh = mesh
        self.xyz = np.ascontiguousarray(mesh["xyz"], dtype=np.float32)
        self.tri = np.ascontiguousarray(mesh["tri"], dtype=np.uint32)
        self.onb = np.ascontiguousarray(mesh["on_boundary"], dtype=np.uint8)
        self.nv = self.xyz.shape[0]
        self.ntri = self.tri.shape[0]
        self.bc = bc
        self.idx_map = np.empty(self.nv, dtype=np.int32)
        self.n = int(lib.orc_idx_map(self.nv, _p(self.onb), bc, _p(self.idx_map)))
        cap = 9 * self.ntri + self.n + 1
        self.row_ptr = np.zeros(self.n + 1, dtype=np.int32)
        ci = np.empty(cap, dtype=np.int32)
        nnz = int(lib.orc_pattern(self.nv, self.ntri, _p(self.tri), _p(self.onb), bc, _p(self.idx_map),
                                  self.n, _p(self.row_ptr), _p(ci), cap))
        assert nnz >= 0
        self.col_idx = ci[:nnz].copy()
        self.nnz = nnz
        self.dtype = np.dtype(dtype)
        self.velocity = np.ascontiguousarray(velocity, dtype=np.float64)
        self.velocity_e = None if velocity_e is None else np.ascontiguousarray(velocity_e, dtype=np.float64)
        self.kappa_e = None if kappa_e is None else np.ascontiguousarray(kappa_e, dtype=np.float64)
        self.normalize = normalize
        self.K = np.zeros(nnz, dtype=dtype)
        self.M = np.zeros(nnz, dtype=dtype)
        self.A = np.zeros(nnz, dtype=dtype)
        fn = lib.orc_assemble_d if self.dtype == np.float64 else lib.orc_assemble_f
        fn(self.nv, self.ntri, _p(self.xyz), _p(self.tri), _p(self.onb), bc, _p(self.idx_map),
           _p(self.row_ptr), _p(self.col_idx), _p(self.velocity), _p(self.velocity_e),
           1 if normalize else 0, _p(self.kappa_e), _p(self.K), _p(self.M), _p(self.A))

    def scipy(self, which):
        import scipy.sparse as sp
        vals = {"K": self.K, "M": self.M, "A": self.A}[which]
        return sp.csr_matrix((vals.astype(np.float64), self.col_idx, self.row_ptr), shape=(self.n, self.n))

    def gather(self, values_nv):
        """CPUSolver::get_surface_value_vector (CPUSolver.cpp:13-21)."""
        out = np.zeros(self.n, dtype=np.float64)
        m = self.idx_map >= 0
        out[self.idx_map[m]] = np.asarray(values_nv, dtype=np.float64)[m]
        return out

    def scatter(self, x):
        """CPUSolver::map_vector_to_surface (CPUSolver.cpp:28-37): boundary -> 0."""
        out = np.zeros(self.nv, dtype=np.float64)
        m = self.idx_map >= 0
        out[m] = x[self.idx_map[m]]
        return out

    def step(self, equation, prm, u, v=None, a=None, krylov=CG, tol=1e-13, maxit=0, warm=False):
        """One oracle time step in place; returns Krylov iterations."""
        lib = load()
        dt = self.dtype
        u = np.ascontiguousarray(u, dtype=dt)
        v = np.zeros_like(u) if v is None else np.ascontiguousarray(v, dtype=dt)
        a = np.zeros_like(u) if a is None else np.ascontiguousarray(a, dtype=dt)
        pr = np.ascontiguousarray(prm, dtype=np.float64)
        fn = lib.orc_step_d if dt == np.float64 else lib.orc_step_f
        its = fn(self.n, _p(self.row_ptr), _p(self.col_idx), _p(self.K), _p(self.M), _p(self.A), equation,
                 _p(pr), _p(u), _p(v), _p(a), krylov, tol, maxit, 1 if warm else 0)
        if its < 0:
            raise RuntimeError("oracle step failed")
        return its, u, v, a

    def run(self, equation, prm, nsteps, u, v=None, a=None, **kw):
        its = 0
        for _ in range(nsteps):
            k, u, v, a = self.step(equation, prm, u, v, a, **kw)
            its += k
        return its, u, v, a

    def cg(self, S, b, x0=None, tol=1e-13, maxit=0):
        lib = load()
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        err = C.c_double()
        its = lib.orc_cg_d(self.n, _p(self.row_ptr), _p(self.col_idx), _p(np.ascontiguousarray(S)),
                           _p(np.ascontiguousarray(b, dtype=np.float64)), _p(x), tol, maxit or 2 * self.n,
                           C.byref(err))
        return x, its, err.value

    def bicgstab(self, S, b, x0=None, tol=1e-13, maxit=0):
        lib = load()
        x = np.zeros(self.n) if x0 is None else np.array(x0, dtype=np.float64)
        err = C.c_double()
        its = lib.orc_bicgstab_d(self.n, _p(self.row_ptr), _p(self.col_idx), _p(np.ascontiguousarray(S)),
                                 _p(np.ascontiguousarray(b, dtype=np.float64)), _p(x), tol,
                                 maxit or 20 * self.n, C.byref(err))
        return x, its, err.value

    def newmark_init(self, c, u0, tol=1e-14):
        lib = load()
        a0 = np.zeros(self.n)
        lib.orc_newmark_init_d(self.n, _p(self.row_ptr), _p(self.col_idx), _p(self.K), _p(self.M), c,
                               _p(np.ascontiguousarray(u0, dtype=np.float64)), _p(a0), tol)
        return a0

    def adjoint(self, dt, nsteps, u0, target, tol=1e-14):
        """Loss, u_T, dL/du0, dL/dkappa_e of the backward-Euler advection-diffusion
        forward with S = M/dt + K_kappa - A (K assembled with kappa_e)."""
        lib = load()
        uT = np.zeros(self.n)
        gu = np.zeros(self.n)
        gk = np.zeros(self.ntri)
        loss = lib.orc_adjoint_d(self.nv, self.ntri, _p(self.xyz), _p(self.tri), _p(self.idx_map), self.n,
                                 _p(self.row_ptr), _p(self.col_idx), _p(self.K), _p(self.M), _p(self.A),
                                 dt, nsteps, _p(np.ascontiguousarray(u0, dtype=np.float64)),
                                 _p(np.ascontiguousarray(target, dtype=np.float64)), tol, _p(uT), _p(gu), _p(gk))
        return loss, uT, gu, gk


def gl_step(sysf, equation, prm, u, v, max_it=10):
    """One frame of the reference's OpenGL solver (GPUSolver.cpp:74-172 with
    cgm_helper.glsl:74-117 and cgm.glsl:75-148) in float32: warm start,
    unpreconditioned CG for at most max_it iterations while r.r > 1e-10,
    stopping once r.r <= 1e-10 r0.r0; Gray-Scott reaction summed over the
    row's slots with the neighbours' values, u solved first and the v solve
    using the new u.  Returns (u, v, last r.r).  (The shader's reductions and
    direction update race; this restatement uses their intended values.)"""
    f = np.float32
    n, rp, ci = sysf.n, sysf.row_ptr, sysf.col_idx
    rows = np.repeat(np.arange(n), np.diff(rp))
    Mf, Kf, Af = (np.asarray(x, dtype=f) for x in (sysf.M, sysf.K, sysf.A))
    dt = f(prm[0])

    def rowsum(vals):
        return np.bincount(rows, weights=np.asarray(vals, dtype=np.float64), minlength=n).astype(f)

    def coef(eq):
        if eq == HEAT:
            return Mf / dt + f(prm[1]) * Kf
        if eq == ADVDIFF:
            return Mf / dt + f(prm[1]) * Kf - Af
        if eq == WAVE:
            return Mf / dt + f(prm[1]) * f(prm[1]) * Kf * dt
        if eq == "gs_u":
            return Mf / dt + f(prm[1]) * Kf
        return Mf / dt + f(prm[2]) * Kf

    def cgm(S, b, x):
        r = (b - rowsum(S * x[ci])).astype(f)
        d = r.copy()
        ri = f(np.dot(r.astype(np.float64), r))
        r0 = ri
        it = 0
        while ri > f(1e-10) and it < max_it:
            ad = rowsum(S * d[ci])
            dad = f(np.dot(d.astype(np.float64), ad))
            alpha = f(ri / dad)
            r = (r - alpha * ad).astype(f)
            x = (x + alpha * d).astype(f)
            ri1 = f(np.dot(r.astype(np.float64), r))
            d = (r + f(ri1 / ri) * d).astype(f)
            ri = ri1
            it += 1
            if ri <= f(1e-10) * r0:
                break
        return x, ri

    u = np.asarray(u, dtype=f).copy()
    v = np.asarray(v, dtype=f).copy()
    if equation == GRAYSCOTT:
        F, kr = f(prm[3]), f(prm[4])
        uc, vc = u[ci], v[ci]
        b = rowsum(uc * (Mf / dt) - uc * vc * vc + F * (f(1) - uc))
        u, rr = cgm(coef("gs_u"), b, u)
        uc = u[ci]
        b = rowsum(vc * (Mf / dt) + uc * vc * vc - vc * (F + kr))
        v, rr = cgm(coef("gs_v"), b, v)
        return u, v, rr
    if equation == WAVE:
        b = rowsum(v[ci] * (Mf / dt) - f(prm[1]) * f(prm[1]) * u[ci] * Kf)
        v, rr = cgm(coef(WAVE), b, v)
        return (u + v * dt).astype(f), v, rr
    b = rowsum(u[ci] * (Mf / dt))
    u, rr = cgm(coef(equation), b, u)
    return u, v, rr


def smooth_normals(xyz, tri, normals, values, extrusion):
    """shaders/FEM/smooth_normals.glsl:22-57 in float32: positions moved along
    their normals by min(1, extrusion max(0, value)), each triangle's
    unnormalised normal cross(b - a, c - a) summed into its vertices (here in
    triangle order; the shader's float compare-and-swap order is arbitrary)."""
    f = np.float32
    x = np.asarray(xyz, dtype=f)
    nrm = np.asarray(normals, dtype=f).reshape(-1, 3)
    v = np.asarray(values, dtype=f)
    h = np.minimum(f(1.0), f(extrusion) * np.maximum(f(0.0), v))
    p = x + h[:, None] * nrm
    t = np.asarray(tri, dtype=np.int64)
    a, b, c = p[t[:, 0]], p[t[:, 1]], p[t[:, 2]]
    n = np.cross(b - a, c - a).astype(f)
    out = np.zeros_like(x)
    for k in range(3):
        np.add.at(out, t[:, k], n)
    return out


def csr_to_csc(n, rp, ci, vals):
    """CSC arrays (colptr, rowidx, vals) of the CSR matrix (rp, ci, vals)."""
    import scipy.sparse as sp
    m = sp.csr_matrix((vals, ci, rp), shape=(n, n)).tocsc()
    m.sort_indices()
    return (m.indptr.astype(np.int32), m.indices.astype(np.int32), np.ascontiguousarray(m.data))


def port_prepare(sysf: System):
    """CSC float32 copies of K, M, A (FEMContext's Eigen storage order)."""
    n = sysf.n
    return {
        "n": n,
        "K": csr_to_csc(n, sysf.row_ptr, sysf.col_idx, sysf.K.astype(np.float32)),
        "M": csr_to_csc(n, sysf.row_ptr, sysf.col_idx, sysf.M.astype(np.float32)),
        "A": csr_to_csc(n, sysf.row_ptr, sysf.col_idx, sysf.A.astype(np.float32)),
    }


def port_run(sysf, equation, prm, nsteps, u, v=None):
    """Reference-cost FP32 port (CPUSolver semantics, Eigen CG defaults).

    sysf: a System or the result of port_prepare.  Advances nsteps on float32
    copies; returns (iterations, u, v)."""
    lib = load()
    prep = sysf if isinstance(sysf, dict) else port_prepare(sysf)
    n, K, M, A = prep["n"], prep["K"], prep["M"], prep["A"]
    u = np.ascontiguousarray(u, dtype=np.float32).copy()
    v = np.zeros_like(u) if v is None else np.ascontiguousarray(v, dtype=np.float32).copy()
    pr = np.ascontiguousarray(prm, dtype=np.float64)
    its = lib.orc_port_run_f(n, _p(K[0]), _p(K[1]), _p(K[2]), _p(M[0]), _p(M[1]), _p(M[2]),
                             _p(A[0]), _p(A[1]), _p(A[2]), equation, _p(pr), _p(u), _p(v), nsteps)
    return its, u, v


# ---- surface input and output (Surface.cpp:57-133, 138-285) ---------------

def read_obj(path):
    """Surface::init_from_obj (src/Utils/Surface.cpp:57-133) restated:
    xyz, tri, on_boundary and vertex normals.  Faces are triangulated as
    tinyobjloader's triangulate option does: triangles kept, quads split along
    the shorter diagonal (squared lengths in float32; [0,1,2],[0,2,3] when
    |v2-v0|^2 < |v3-v1|^2, else [0,1,3],[1,2,3]), larger polygons fanned.  Each
    object/group re-appends the whole vertex list while its indices stay
    unshifted (Surface.cpp:77-80); vertex normals are the float32 sums of the
    triangulated corners' normals, normalised as glm::normalize (NaN where no
    corner contributes); boundary by the edge-count rule (:104-127)."""
    f32 = np.float32
    verts, norms = [], []
    shapes = [[]]  # triangulated corners (v, n)

    def fix(raw, count):
        return raw - 1 if raw > 0 else count + raw

    with open(path, "r") as fh:
        for line in fh:
            parts = line.split()
            if not parts or parts[0].startswith("#"):
                continue
            tag = parts[0]
            if tag in ("v", "vn"):
                (verts if tag == "v" else norms).append([f32(float(x)) for x in parts[1:4]])
            elif tag in ("o", "g"):
                if shapes[-1]:
                    shapes.append([])
            elif tag == "f":
                face = []
                for tok in parts[1:]:
                    sub = tok.split("/")
                    v = fix(int(sub[0]), len(verts))
                    n = fix(int(sub[2]), len(norms)) if len(sub) >= 3 and sub[2] else -1
                    face.append((v, n))
                if len(face) == 3:
                    shapes[-1] += face
                elif len(face) == 4:
                    p = [np.asarray(verts[c[0]], dtype=f32) for c in face]
                    a, b = p[2] - p[0], p[3] - p[1]
                    s02 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2]
                    s13 = b[0] * b[0] + b[1] * b[1] + b[2] * b[2]
                    order = (0, 1, 2, 0, 2, 3) if s02 < s13 else (0, 1, 3, 1, 2, 3)
                    shapes[-1] += [face[q] for q in order]
                elif len(face) > 4:
                    for q in range(1, len(face) - 1):
                        shapes[-1] += [face[0], face[q], face[q + 1]]
    if len(shapes) > 1 and not shapes[-1]:
        shapes.pop()
    if not verts:
        raise ValueError(f"File {path} does not contain a valid .obj mesh.")
    if not norms:
        raise ValueError(f"File {path} does not contain normals!")
    V = np.asarray(verts, dtype=f32)
    Nn = np.asarray(norms, dtype=f32)
    xyz = np.concatenate([V] * len(shapes))
    corners = [c for sh in shapes for c in sh]
    tri = np.asarray([c[0] for c in corners], dtype=np.uint32).reshape(-1, 3)
    nrm = np.zeros_like(xyz)
    for v, n in corners:
        if n >= 0:
            nrm[v] = nrm[v] + Nn[n]
    with np.errstate(divide="ignore", invalid="ignore"):
        d = nrm[:, 0] * nrm[:, 0] + nrm[:, 1] * nrm[:, 1] + nrm[:, 2] * nrm[:, 2]
        inv = f32(1.0) / np.sqrt(d)
        nrm = nrm * inv[:, None]
    t = tri.astype(np.int64)
    onb = np.zeros(xyz.shape[0], dtype=np.uint8)
    if t.size:
        a = np.concatenate([t[:, 0], t[:, 1], t[:, 2]])
        b = np.concatenate([t[:, 1], t[:, 2], t[:, 0]])
        key = np.minimum(a, b) * np.int64(xyz.shape[0]) + np.maximum(a, b)
        uniq, counts = np.unique(key, return_counts=True)
        once = uniq[counts == 1]
        onb[once // xyz.shape[0]] = 1
        onb[once % xyz.shape[0]] = 1
    return {"xyz": xyz, "tri": tri, "on_boundary": onb, "normals": nrm}


# Viridis as an order-6 polynomial per channel: the reference's default colour
# map (src/Utils/AssetStorage.cpp:141-152).
VIRIDIS = np.array([
    [0.274344, 0.004462, 0.331359],
    [0.108915, 1.397291, 1.388110],
    [-0.319631, 0.243490, 0.156419],
    [-4.629188, -5.882803, -19.646115],
    [6.181719, 14.388598, 57.442181],
    [4.876952, -13.955112, -66.125783],
    [-5.513165, 4.709245, 26.582180],
], dtype=np.float32)
OPEN, CLOSED, MIRRORED = 0, 1, 2  # enum class MeshType (include/Utils/Surface.hpp:25-29)


def viridis(t):
    """ColorMap::get_color (src/Utils/ColorMap.cpp:12-14) in float32."""
    t = np.float32(t)
    acc = VIRIDIS[6]
    for k in range(5, -1, -1):
        acc = VIRIDIS[k] + t * acc
    return acc


def ply_clip(xyz, tri, normals, values, vertex_extrusion, threshold, mesh_type=OPEN):
    """Surface::export_to_ply's clipping (src/Utils/Surface.cpp:146-255) in
    float32: positions, their values and faces.  Positions are keyed by value
    (glm::vec3 equality: -0.0 equals 0.0) in first-use order."""
    f = np.float32
    V = np.asarray(xyz, dtype=f).reshape(-1, 3)
    Nm = np.asarray(normals, dtype=f).reshape(-1, 3)
    val = np.asarray(values, dtype=f).reshape(-1)
    T = np.asarray(tri, dtype=np.int64).reshape(-1, 3)
    ext, thr = f(vertex_extrusion), f(threshold)
    positions, pvalues, faces = [], [], []
    index = {}

    def add(pos, value):
        key = tuple(float(z) + 0.0 for z in pos)  # + 0.0 maps -0.0 to 0.0
        if key not in index:
            index[key] = len(positions)
            positions.append(pos.astype(f))
            pvalues.append(f(value))
        return index[key]

    def mix(x, y, a):
        a = f(a)
        return x * (f(1.0) - a) + y * a

    def extruded(i):
        return (val[i] * max(f(0.0), ext)) * Nm[i] + V[i]

    def projected(i):
        scale = (thr - val[i] * ext) if mesh_type == MIRRORED else (thr * ext)
        return f(scale) * Nm[i] + V[i]

    for t in T:
        above = [bool(float(val[i] - thr) >= 1e-9) for i in t]
        count = sum(above)
        if count == 1:
            ai = above.index(True)
            clipped = [0, 0, 0]
            for i in range(3):
                if i != ai:
                    w = (thr - val[t[i]]) / (val[t[ai]] - val[t[i]])
                    clipped[i] = add(mix(extruded(t[i]), extruded(t[ai]), w), thr)
                else:
                    clipped[i] = add(extruded(t[i]), val[t[ai]])
            faces.append(tuple(clipped))
            if mesh_type != OPEN:
                faces.append(tuple(clipped[i] if i != ai else add(projected(t[i]), thr) for i in range(3)))
        elif count == 2:
            a1 = 0 if above[0] else 1
            a2 = 1 if (above[0] and above[1]) else 2
            b = 3 - (a1 + a2)
            p1, p2 = extruded(t[a1]), extruded(t[a2])
            q1 = mix(extruded(t[b]), p1, (thr - val[t[b]]) / (val[t[a1]] - val[t[b]]))
            q2 = mix(extruded(t[b]), p2, (thr - val[t[b]]) / (val[t[a2]] - val[t[b]]))
            i1, i2 = add(p1, val[t[a1]]), add(p2, val[t[a2]])
            j1, j2 = add(q1, thr), add(q2, thr)
            faces += [(j1, i1, j2), (j2, i1, i2)]
            if mesh_type != OPEN:
                k1 = add(projected(t[a1]), val[t[a1]])
                k2 = add(projected(t[a2]), val[t[a2]])
                faces += [(j1, k1, j2), (j2, k1, k2)]
        elif count == 3:
            faces.append(tuple(add(extruded(i), val[i]) for i in t))
            if mesh_type != OPEN:
                faces.append(tuple(add(projected(i), thr) for i in t))
    return (np.array(positions, dtype=f).reshape(-1, 3), np.array(pvalues, dtype=f),
            np.array(faces, dtype=np.int64).reshape(-1, 3))
