"""Write the FP64 golden fixtures used by the parity tests.

The vectors come from an implementation written independently of the oracle:
element matrices from the textbook barycentric-gradient formula
grad(lambda_i) = n x e_i / (2 area) (e_i = edge opposite vertex i), assembly
by scipy COO summation, and every time step solved with scipy's sparse direct
solver.  The oracle restates the reference's J^{-T} construction and Eigen's
CG instead, so agreement between the two pins both.

Meshes: the 64x64 structured unit square of the heat benchmark, the
reference's shipped assets/fem_meshes/icosphere.obj (10,242 vertices), whose
arrays are stored in the fixture so the tests do not need the reference tree,
and a jittered 24x24 lattice with per-element diffusivity and per-element,
unnormalised velocity.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spl

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "fea-visualizer_amd"))

from femhip.meshes import jittered_square, structured_square  # noqa: E402

sys.path.insert(0, ROOT)
from oracle import read_obj  # noqa: E402  (the reader's restatement, not the library's)

REF_ICO = "/root/reference/assets/fem_meshes/icosphere.obj"


def element_data(xyz, tri, velocity, normalize=True, velocity_e=None):
    X = xyz.astype(np.float64)
    a, b, c = X[tri[:, 0]], X[tri[:, 1]], X[tri[:, 2]]
    n = np.cross(b - a, c - a)
    ln = np.linalg.norm(n, axis=1)
    area = 0.5 * ln
    nh = n / ln[:, None]
    edges = [c - b, a - c, b - a]  # opposite vertex 0, 1, 2
    grads = np.stack([np.cross(nh, e) / (2 * area[:, None]) for e in edges], axis=1)  # [T,3,3]
    if velocity_e is None:
        v = np.asarray(velocity, dtype=np.float64)[None, :]
    else:
        v = np.asarray(velocity_e, dtype=np.float64).reshape(-1, 3)
    vt = v - (v * nh).sum(1, keepdims=True) * nh
    if normalize:
        vt = vt / np.linalg.norm(vt, axis=1, keepdims=True)
    return area, grads, vt


def assemble(mesh, bc, velocity=(1.0, 0.0, 0.0), normalize=True, velocity_e=None, kappa_e=None):
    xyz, tri, onb = mesh["xyz"], mesh["tri"].astype(np.int64), mesh["on_boundary"].astype(bool)
    nv = xyz.shape[0]
    idx = -np.ones(nv, dtype=np.int64)
    keep = np.ones(nv, bool) if bc == 1 else ~onb
    idx[keep] = np.arange(keep.sum())
    n = int(keep.sum())
    area, g, vt = element_data(xyz, tri, velocity, normalize, velocity_e)
    kap = np.ones(tri.shape[0]) if kappa_e is None else np.asarray(kappa_e, dtype=np.float64)
    rows, cols, kv, mv, av, pv = [], [], [], [], [], []
    for i in range(3):
        for j in range(3):
            bi, bj = onb[tri[:, i]], onb[tri[:, j]]
            keep_km = np.ones_like(bi) if bc == 1 else (~bi & ~bj)
            keep_a = ~bi & ~bj
            r, cc = idx[tri[:, i]], idx[tri[:, j]]
            kij = kap * area * (g[:, i] * g[:, j]).sum(1)
            mij = area / 12.0 * (2.0 if i == j else 1.0)
            aij = area / 3.0 * (vt * g[:, j]).sum(1)
            rows.append(r[keep_km]); cols.append(cc[keep_km])
            kv.append(kij[keep_km]); mv.append(np.broadcast_to(mij, kij.shape)[keep_km])
            av.append(np.where(keep_a, aij, 0.0)[keep_km])
    R = np.concatenate(rows); Cc = np.concatenate(cols)
    mats = {}
    for name, vals in (("K", kv), ("M", mv), ("A", av)):
        m = sp.coo_matrix((np.concatenate(vals), (R, Cc)), shape=(n, n)).tocsr()
        m.sum_duplicates()
        m.sort_indices()
        mats[name] = m
    # one shared pattern (explicit zeros kept by construction)
    pat = sp.coo_matrix((np.ones(len(R)), (R, Cc)), shape=(n, n)).tocsr()
    pat.sum_duplicates()
    pat.sort_indices()
    out = {"idx_map": idx.astype(np.int32), "row_ptr": pat.indptr.astype(np.int32),
           "col_idx": pat.indices.astype(np.int32)}
    for name, m in mats.items():
        vals = np.zeros(pat.nnz)
        lookup = {}
        mc = m.tocoo()
        d = dict(zip(zip(mc.row.tolist(), mc.col.tolist()), mc.data.tolist()))
        pc = pat.tocoo()
        vals = np.array([d.get((r, c), 0.0) for r, c in zip(pc.row.tolist(), pc.col.tolist())])
        out[name] = vals
        del lookup
    return out, mats, n


def heat_steps(mats, u, dt, c, nsteps):
    S = (mats["M"] / dt + c * mats["K"]).tocsc()
    lu = spl.splu(S)
    out = []
    for _ in range(nsteps):
        u = lu.solve((mats["M"] / dt) @ u)
        out.append(u.copy())
    return np.array(out)


def advdiff_steps(mats, u, dt, c, nsteps):
    S = (mats["M"] / dt + c * mats["K"] - mats["A"]).tocsc()
    lu = spl.splu(S)
    out = []
    for _ in range(nsteps):
        u = lu.solve((mats["M"] / dt) @ u)
        out.append(u.copy())
    return np.array(out)


def wave_steps(mats, u, v, dt, c, nsteps):
    S = (mats["M"] / dt + c * c * dt * mats["K"]).tocsc()
    lu = spl.splu(S)
    us, vs = [], []
    for _ in range(nsteps):
        v = lu.solve((mats["M"] / dt) @ v - c * c * (mats["K"] @ u))
        u = u + dt * v
        us.append(u.copy()); vs.append(v.copy())
    return np.array(us), np.array(vs)


def gs_steps(mats, u, v, dt, Du, Dv, F, k, nsteps, lumped):
    Su = spl.splu((mats["M"] / dt + Du * mats["K"]).tocsc())
    Sv = spl.splu((mats["M"] / dt + Dv * mats["K"]).tocsc())
    w = np.asarray(mats["M"].sum(1)).ravel() if lumped else np.ones(u.shape[0])
    us, vs = [], []
    for _ in range(nsteps):
        uvv = u * v * v
        bu = (mats["M"] / dt) @ u + w * (-uvv + F * (1 - u))
        bv = (mats["M"] / dt) @ v + w * (uvv - (F + k) * v)
        u, v = Su.solve(bu), Sv.solve(bv)
        us.append(u.copy()); vs.append(v.copy())
    return np.array(us), np.array(vs)


def patches(xyz, centres, radius):
    """1 inside any geodesic disc of `radius` (radians) around the unit
    `centres` on the sphere through the mesh's vertices, else 0."""
    X = xyz.astype(np.float64)
    X = X / np.linalg.norm(X, axis=1, keepdims=True)
    c = np.asarray(centres, dtype=np.float64)
    c = c / np.linalg.norm(c, axis=1, keepdims=True)
    return (np.arccos(np.clip(X @ c.T, -1.0, 1.0)) < radius).any(1).astype(np.float64)


def main():
    # ---- heat benchmark mesh: 64x64 structured square, Dirichlet
    sq = structured_square(64)
    asm, mats, n = assemble(sq, 0)
    x = sq["xyz"].astype(np.float64)
    u0_nv = np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.01)
    keep = asm["idx_map"] >= 0
    u0 = np.zeros(n); u0[asm["idx_map"][keep]] = u0_nv[keep]
    heat = heat_steps(mats, u0, 1e-3, 1.0, 3)
    adv = advdiff_steps(mats, u0, 1e-3, 0.25, 3)
    wu, wv = wave_steps(mats, u0, np.zeros(n), 0.05, 0.05, 3)
    np.savez_compressed(os.path.join(HERE, "square64.npz"), xyz=sq["xyz"], tri=sq["tri"],
                        on_boundary=sq["on_boundary"], u0=u0, heat=heat, advdiff=adv, wave_u=wu, wave_v=wv,
                        **asm)
    # ---- shipped icosphere (closed, 10,242 vertices)
    if os.path.exists(REF_ICO):
        ico = read_obj(REF_ICO)
        for bc in (1,):
            asm, mats, n = assemble(ico, bc)
            X = ico["xyz"].astype(np.float64)
            v0 = 0.25 * np.exp(-((X - np.array([0.0, 0.0, 1.0])) ** 2).sum(1) / 0.05)
            u0 = np.zeros(n)  # CPUSolver::clear_values leaves u = 0
            gu, gv = gs_steps(mats, u0, v0, 1e-3, 0.08, 0.04, 0.035, 0.06, 3, lumped=False)
            hu = heat_steps(mats, v0, 0.01, 0.05, 3)
            # the benchmark's Gray-Scott regime: dt = 0.5, u = 1, v = 0.25 on
            # six patches, reaction weighted by the lumped mass (row sums of M)
            cen = [(0, 0, 1), (0, 0, -1), (1, 0, 0), (-1, 0, 0), (0, 1, 0), (0.6, 0.6, 0.5)]
            lv0 = 0.25 * patches(ico["xyz"], cen, 0.15)
            lu0 = np.ones(n)
            lu, lv = gs_steps(mats, lu0, lv0, 0.5, 2e-4, 1e-4, 0.037, 0.06, 3, lumped=True)
            np.savez_compressed(os.path.join(HERE, "icosphere_shipped.npz"), xyz=ico["xyz"], tri=ico["tri"],
                                on_boundary=ico["on_boundary"], v0=v0, gs_u=gu, gs_v=gv, heat=hu,
                                lumped_v0=lv0, lumped_u=lu, lumped_v=lv, **asm)
    # ---- jittered lattice with per-element diffusivity and per-element,
    # unnormalised velocity (the differentiable workload's inputs), Dirichlet
    jm = jittered_square(24, seed=11)
    cen = jm["xyz"].astype(np.float64)[jm["tri"].astype(np.int64)].mean(axis=1)
    ve = np.stack([-(cen[:, 1] - 0.5), cen[:, 0] - 0.5, np.zeros(len(cen))], axis=1)
    rk = np.random.Generator(np.random.PCG64(17))
    kap = 1e-3 * np.exp(0.3 * rk.standard_normal(jm["tri"].shape[0]))
    asm, mats, n = assemble(jm, 0, normalize=False, velocity_e=ve, kappa_e=kap)
    X = jm["xyz"].astype(np.float64)
    keep = asm["idx_map"] >= 0
    u0 = np.exp(-((X[:, 0] - 0.3) ** 2 + (X[:, 1] - 0.5) ** 2) / (2 * 0.05 ** 2))[keep]
    adv = advdiff_steps(mats, u0, 2e-3, 1.0, 3)
    np.savez_compressed(os.path.join(HERE, "jittered_kappa.npz"), xyz=jm["xyz"], tri=jm["tri"],
                        on_boundary=jm["on_boundary"], velocity_e=ve, kappa_e=kap, u0=u0, advdiff=adv, **asm)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
