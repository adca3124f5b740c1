"""CPU tests that pin the oracle (the CPU restatement the GPU path is checked
against) to golden vectors, closed forms, invariants and brute force."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spl

import oracle
from femhip import meshes

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def load_gold(name):
    return dict(np.load(os.path.join(GOLD, name)))


def mesh_of(g):
    return {"xyz": g["xyz"], "tri": g["tri"], "on_boundary": g["on_boundary"]}


def brute_pattern(mesh, bc):
    tri = mesh["tri"].astype(np.int64)
    onb = mesh["on_boundary"].astype(bool)
    nv = mesh["xyz"].shape[0]
    keep = np.ones(nv, bool) if bc == 1 else ~onb
    idx = -np.ones(nv, dtype=np.int64)
    idx[keep] = np.arange(keep.sum())
    rows = [set() for _ in range(int(keep.sum()))]
    for t in tri:
        for a in t:
            for b in t:
                if bc == 1 or (not onb[a] and not onb[b]):
                    rows[idx[a]].add(int(idx[b]))
    rp = [0]
    ci = []
    for r in rows:
        ci += sorted(r)
        rp.append(len(ci))
    return idx.astype(np.int32), np.array(rp, np.int32), np.array(ci, np.int32)


@pytest.mark.parametrize("name,mk,bc", [
    ("square8", lambda: meshes.structured_square(8), 0),
    ("jitter12", lambda: meshes.jittered_square(12, seed=3), 0),
    ("jitter12N", lambda: meshes.jittered_square(12, seed=3), 1),
    ("ico2", lambda: meshes.icosphere(2), 1),
])
def test_pattern_brute_force(name, mk, bc):
    m = mk()
    s = oracle.System(m, bc=bc)
    idx, rp, ci = brute_pattern(m, bc)
    assert np.array_equal(s.idx_map, idx)
    assert np.array_equal(s.row_ptr, rp)
    assert np.array_equal(s.col_idx, ci)


def test_heat_mesh_counts_and_stencil():
    """64x64 structured square: N, nnz, 5-point stencil with stored zeros,
    M_ii = h^2/2 = 2^-13 exactly."""
    m = meshes.structured_square(64)
    s = oracle.System(m)
    assert (s.nv, s.ntri, s.n, s.nnz) == (4225, 8192, 3969, 27281)
    assert int(np.diff(s.row_ptr).max()) == 7
    assert set(np.round(s.K, 12).tolist()) <= {4.0, -1.0, 0.0}
    K = s.scipy("K")
    assert np.allclose(K.diagonal(), 4.0, atol=1e-12)
    # the diagonal-edge neighbours are stored explicit zeros
    assert np.sum(np.abs(s.K) < 1e-12) == s.nnz - 5 * s.n + 4 * 63
    M = s.scipy("M")
    assert np.allclose(M.diagonal(), 2.0 ** -13, rtol=4e-16, atol=0)
    assert np.allclose(s.M[np.abs(s.M - 2.0 ** -13) > 1e-9], 2.0 ** -12 / 12, rtol=4e-16, atol=0)


def test_golden_square_matrices_and_steps():
    g = load_gold("square64.npz")
    m = mesh_of(g)
    s = oracle.System(m, bc=0)
    assert np.array_equal(s.idx_map, g["idx_map"])
    assert np.array_equal(s.row_ptr, g["row_ptr"])
    assert np.array_equal(s.col_idx, g["col_idx"])
    for name in "KMA":
        ref = g[name]
        got = getattr(s, name)
        assert np.max(np.abs(got - ref)) <= 1e-12 * max(1e-30, np.max(np.abs(ref))), name
    u = g["u0"].copy()
    for k in range(3):
        _, u, _, _ = s.step(oracle.HEAT, [1e-3, 1.0], u, tol=1e-14)
        ref = g["heat"][k]
        assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-11
    u = g["u0"].copy()
    for k in range(3):
        _, u, _, _ = s.step(oracle.ADVDIFF, [1e-3, 0.25], u, krylov=oracle.BICGSTAB, tol=1e-14)
        ref = g["advdiff"][k]
        assert np.linalg.norm(u - ref) / np.linalg.norm(ref) < 1e-11
    u = g["u0"].copy()
    v = np.zeros_like(u)
    for k in range(3):
        _, u, v, _ = s.step(oracle.WAVE, [0.05, 0.05], u, v, tol=1e-14)
        assert np.linalg.norm(u - g["wave_u"][k]) / np.linalg.norm(g["wave_u"][k]) < 1e-11
        assert np.linalg.norm(v - g["wave_v"][k]) / np.linalg.norm(g["wave_v"][k]) < 1e-10


def test_golden_icosphere_shipped():
    g = load_gold("icosphere_shipped.npz")
    m = mesh_of(g)
    s = oracle.System(m, bc=1)
    assert (s.nv, s.ntri, s.nnz) == (10242, 20480, 71682)
    assert int(m["on_boundary"].sum()) == 0
    assert np.array_equal(s.row_ptr, g["row_ptr"])
    assert np.array_equal(s.col_idx, g["col_idx"])
    # K, M and the curved-surface advection A (velocity (1,0,0) projected on
    # each element's plane and normalised)
    for name in "KMA":
        assert np.max(np.abs(getattr(s, name) - g[name])) <= 1e-12 * np.max(np.abs(g[name]))
    # total area of the shipped icosphere (sum of all consistent-mass entries)
    assert abs(s.M.sum() - 12.5626) < 1e-3
    u = np.zeros(s.n)
    v = g["v0"].copy()
    prm = [1e-3, 0.08, 0.04, 0.035, 0.06, 0]
    for k in range(3):
        _, u, v, _ = s.step(oracle.GRAYSCOTT, prm, u, v, tol=1e-14)
        assert np.linalg.norm(u - g["gs_u"][k]) / np.linalg.norm(g["gs_u"][k]) < 1e-11
        assert np.linalg.norm(v - g["gs_v"][k]) / np.linalg.norm(g["gs_v"][k]) < 1e-11


def test_golden_icosphere_lumped_grayscott():
    """The benchmark's Gray-Scott regime (dt = 0.5, u = 1, patches of v = 0.25)
    with the reaction weighted by the lumped mass, three steps, against the
    direct-solve fixture."""
    g = load_gold("icosphere_shipped.npz")
    s = oracle.System(mesh_of(g), bc=1)
    u = np.ones(s.n)
    v = g["lumped_v0"].copy()
    prm = [0.5, 2e-4, 1e-4, 0.037, 0.06, 1]
    for k in range(3):
        _, u, v, _ = s.step(oracle.GRAYSCOTT, prm, u, v, tol=1e-14)
        assert np.linalg.norm(u - g["lumped_u"][k]) / np.linalg.norm(g["lumped_u"][k]) < 1e-11
        assert np.linalg.norm(v - g["lumped_v"][k]) / np.linalg.norm(g["lumped_v"][k]) < 1e-11
    # the lumped weighting changes the answer: the nodal one differs at step 1
    _, un, vn, _ = s.step(oracle.GRAYSCOTT, [0.5, 2e-4, 1e-4, 0.037, 0.06, 0], np.ones(s.n),
                          g["lumped_v0"].copy(), tol=1e-14)
    assert np.linalg.norm(un - g["lumped_u"][0]) / np.linalg.norm(g["lumped_u"][0]) > 1e-3


def test_golden_jittered_kappa_velocity():
    """Per-element diffusivity and per-element, unnormalised velocity: K, M,
    A and three backward-Euler advection-diffusion steps against the fixture
    built with the barycentric-gradient formula and a direct solver."""
    g = load_gold("jittered_kappa.npz")
    s = oracle.System(mesh_of(g), bc=0, velocity_e=g["velocity_e"], normalize=False, kappa_e=g["kappa_e"])
    assert np.array_equal(s.idx_map, g["idx_map"])
    assert np.array_equal(s.row_ptr, g["row_ptr"]) and np.array_equal(s.col_idx, g["col_idx"])
    for name in "KMA":
        assert np.max(np.abs(getattr(s, name) - g[name])) <= 1e-12 * np.max(np.abs(g[name]))
    u = g["u0"].copy()
    for k in range(3):
        _, u, _, _ = s.step(oracle.ADVDIFF, [2e-3, 1.0], u, krylov=oracle.BICGSTAB, tol=1e-14)
        assert np.linalg.norm(u - g["advdiff"][k]) / np.linalg.norm(g["advdiff"][k]) < 1e-11
    # a dropped kappa or a normalised velocity would not match
    s1 = oracle.System(mesh_of(g), bc=0, velocity_e=g["velocity_e"], normalize=False)
    assert np.max(np.abs(s1.K - g["K"])) > 1e-3 * np.max(np.abs(g["K"]))
    s2 = oracle.System(mesh_of(g), bc=0, velocity_e=g["velocity_e"], normalize=True, kappa_e=g["kappa_e"])
    assert np.max(np.abs(s2.A - g["A"])) > 1e-3 * np.max(np.abs(g["A"]))


def cot(a, b, c):
    """cotangent of the angle at a in triangle (a, b, c)"""
    u, w = b - a, c - a
    return np.dot(u, w) / np.linalg.norm(np.cross(u, w))


def test_stiffness_cotangent_closed_form():
    """K_ij = -1/2 cot(theta_k) for one arbitrary 3-D triangle (Neumann: all kept)."""
    rng = np.random.default_rng(0)
    X = rng.normal(size=(3, 3)).astype(np.float32)
    m = {"xyz": X, "tri": np.array([[0, 1, 2]], np.uint32), "on_boundary": np.zeros(3, np.uint8)}
    s = oracle.System(m, bc=1)
    K = s.scipy("K").toarray()
    P = X.astype(np.float64)
    ref = np.zeros((3, 3))
    for i, j, k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
        ref[i, j] = ref[j, i] = -0.5 * cot(P[k], P[i], P[j])
    for i in range(3):
        ref[i, i] = -ref[i].sum()
    assert np.allclose(K, ref, rtol=1e-12, atol=1e-12)
    area = 0.5 * np.linalg.norm(np.cross(P[1] - P[0], P[2] - P[0]))
    M = s.scipy("M").toarray()
    assert np.allclose(M, area / 12 * (np.ones((3, 3)) + np.eye(3)), rtol=1e-14)


def test_advection_closed_form_and_invariants():
    rng = np.random.default_rng(1)
    X = rng.normal(size=(3, 3)).astype(np.float32)
    m = {"xyz": X, "tri": np.array([[0, 1, 2]], np.uint32), "on_boundary": np.zeros(3, np.uint8)}
    vel = np.array([0.3, -1.2, 0.7])
    s = oracle.System(m, bc=1, velocity=vel)
    A = s.scipy("A").toarray()
    P = X.astype(np.float64)
    n = np.cross(P[1] - P[0], P[2] - P[0])
    area = 0.5 * np.linalg.norm(n)
    nh = n / np.linalg.norm(n)
    vt = vel - vel.dot(nh) * nh
    vt /= np.linalg.norm(vt)
    edges = [P[2] - P[1], P[0] - P[2], P[1] - P[0]]
    g = [np.cross(nh, e) / (2 * area) for e in edges]
    ref = np.array([[area / 3 * vt.dot(g[j]) for j in range(3)] for _ in range(3)])
    assert np.allclose(A, ref, rtol=1e-11, atol=1e-13)
    # the unnormalised variant keeps |v_t|
    s2 = oracle.System(m, bc=1, velocity=vel, normalize=False)
    vt2 = vel - vel.dot(nh) * nh
    ref2 = np.array([[area / 3 * vt2.dot(g[j]) for j in range(3)] for _ in range(3)])
    assert np.allclose(s2.scipy("A").toarray(), ref2, rtol=1e-11, atol=1e-13)


def test_row_sum_invariants():
    m = meshes.jittered_square(16, seed=5)
    s = oracle.System(m, bc=1)
    one = np.ones(s.n)
    assert np.max(np.abs(s.scipy("K") @ one)) < 1e-12
    assert abs(s.M.sum() - 1.0) < 1e-12  # unit square
    # A.1 = 0 on rows whose whole stencil passed the advection filter
    Ar = s.scipy("A") @ one
    inner = np.ones(s.n, bool)
    onb = m["on_boundary"].astype(bool)
    for t in m["tri"]:
        if onb[t].any():
            inner[t] = False
    assert np.max(np.abs(Ar[inner])) < 1e-12


def test_neumann_heat_conserves_mass():
    m = meshes.jittered_square(20, seed=7)
    s = oracle.System(m, bc=1)
    x = m["xyz"].astype(np.float64)
    u = np.exp(-((x[:, 0] - 0.3) ** 2 + (x[:, 1] - 0.6) ** 2) / 0.02)
    M = s.scipy("M")
    m0 = (M @ u).sum()
    _, u1, _, _ = s.run(oracle.HEAT, [1e-2, 0.1], 5, u.copy(), tol=1e-14)
    assert abs((M @ u1).sum() - m0) < 1e-11 * abs(m0)


def test_wave_reference_scheme_energy_decays():
    m = meshes.jittered_square(20, seed=8)
    s = oracle.System(m, bc=0)
    x = m["xyz"].astype(np.float64)
    u = s.gather(np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.01))
    v = np.zeros_like(u)
    K, M = s.scipy("K"), s.scipy("M")
    c = 0.5

    def energy(u, v):
        return 0.5 * v @ (M @ v) + 0.5 * c * c * u @ (K @ u)

    e_prev = energy(u, v)
    for _ in range(10):
        _, u, v, _ = s.step(oracle.WAVE, [0.01, c], u, v, tol=1e-14)
        e = energy(u, v)
        assert e <= e_prev * (1 + 1e-12)
        e_prev = e


def test_newmark_conserves_energy():
    m = meshes.jittered_square(20, seed=9)
    s = oracle.System(m, bc=0)
    x = m["xyz"].astype(np.float64)
    u = s.gather(np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / (2 * 0.05 ** 2)))
    v = np.zeros_like(u)
    c = 1.0
    a = s.newmark_init(c, u)
    K, M = s.scipy("K"), s.scipy("M")

    def energy(u, v):
        return 0.5 * v @ (M @ v) + 0.5 * c * c * u @ (K @ u)

    e0 = energy(u, v)
    for _ in range(20):
        _, u, v, a = s.step(oracle.NEWMARK, [2e-3, c, 0.25, 0.5], u, v, a, tol=1e-14)
    assert abs(energy(u, v) - e0) < 1e-9 * e0


def test_cg_and_bicgstab_match_direct():
    m = meshes.jittered_square(24, seed=11)
    s = oracle.System(m, bc=0)
    S = s.M / 1e-3 + 0.1 * s.K - s.A
    Ssp = sp.csr_matrix((S, s.col_idx, s.row_ptr), shape=(s.n, s.n))
    b = np.random.default_rng(2).normal(size=s.n)
    xd = spl.spsolve(Ssp.tocsc(), b)
    xb, _, err = s.bicgstab(S, b, tol=1e-14)
    assert np.linalg.norm(xb - xd) / np.linalg.norm(xd) < 1e-11
    Ssym = s.M / 1e-3 + 0.1 * s.K
    xd2 = spl.spsolve(sp.csr_matrix((Ssym, s.col_idx, s.row_ptr), shape=(s.n, s.n)).tocsc(), b)
    xc, its, _ = s.cg(Ssym, b, tol=1e-14)
    assert np.linalg.norm(xc - xd2) / np.linalg.norm(xd2) < 1e-11
    # Eigen's solve() with b = 0 returns x = 0 without iterating
    x0, its0, _ = s.cg(Ssym, np.zeros(s.n))
    assert its0 == 0 and not x0.any()


def test_adjoint_finite_differences():
    m = meshes.jittered_square(10, seed=12)
    T = m["tri"].shape[0]
    rng = np.random.default_rng(3)
    kap = 1e-2 * np.exp(0.3 * rng.normal(size=T))
    x = m["xyz"].astype(np.float64)
    cent = x[m["tri"]].mean(1)
    vel_e = np.stack([-(cent[:, 1] - 0.5), cent[:, 0] - 0.5, np.zeros(T)], 1)
    dt, nsteps = 2e-2, 6

    def run(kappa, u0):
        s = oracle.System(m, bc=0, velocity_e=vel_e, normalize=False, kappa_e=kappa)
        return s, s.adjoint(dt, nsteps, u0, target)

    s0 = oracle.System(m, bc=0, velocity_e=vel_e, normalize=False, kappa_e=kap)
    u0 = s0.gather(np.exp(-((x[:, 0] - 0.4) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.02))
    target = s0.gather(np.exp(-((x[:, 0] - 0.6) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.03))
    _, (L0, uT, gu, gk) = run(kap, u0)
    h = 1e-6
    for seed in range(2):
        d = np.random.default_rng(10 + seed).normal(size=s0.n)
        Lp = run(kap, u0 + h * d)[1][0]
        Lm = run(kap, u0 - h * d)[1][0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - gu @ d) < 1e-6 * max(1.0, abs(fd))
        dk = np.random.default_rng(20 + seed).normal(size=T) * kap
        Lp = run(kap + h * dk, u0)[1][0]
        Lm = run(kap - h * dk, u0)[1][0]
        fd = (Lp - Lm) / (2 * h)
        assert abs(fd - gk @ dk) < 1e-6 * max(1e-3, abs(fd))


def test_port_matches_fp32_restatement():
    """The timed FP32 port (CSC, re-formed matrices) advances like the oracle."""
    m = meshes.structured_square(32)
    sd = oracle.System(m, bc=0)
    x = m["xyz"].astype(np.float64)
    u0 = sd.gather(np.exp(-((x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.01))
    its, up, _ = oracle.port_run(sd, oracle.HEAT, [1e-3, 1.0], 5, u0)
    _, ud, _, _ = sd.run(oracle.HEAT, [1e-3, 1.0], 5, u0.copy(), tol=1e-14)
    assert its > 0
    assert np.linalg.norm(up - ud) / np.linalg.norm(ud) < 1e-5
    uv = np.zeros(sd.n)
    vv = u0.copy()
    _, pu, pv = oracle.port_run(sd, oracle.GRAYSCOTT, [1e-3, 0.08, 0.04, 0.035, 0.06, 0], 3, uv, vv)
    _, ou, ov, _ = sd.run(oracle.GRAYSCOTT, [1e-3, 0.08, 0.04, 0.035, 0.06, 0], 3, uv.copy(), vv.copy(), tol=1e-14)
    assert np.linalg.norm(pv - ov) / np.linalg.norm(ov) < 1e-5
    assert np.linalg.norm(pu - ou) / np.linalg.norm(ou) < 1e-5


TUBE = "/root/reference/assets/fem_meshes/tube.obj"


@pytest.mark.skipif(not os.path.exists(TUBE), reason="reference assets not mounted")
def test_shipped_tube_counts():
    """tube.obj: 4,704 vertices, 4,608 quads -> 9,216 triangles, 192 boundary
    vertices; Dirichlet N = 4,512, nnz 31,200; Neumann nnz 32,544.  (The quad
    split rule of tinyobjloader is unpinned here; the counts do not depend on
    the diagonal chosen.)"""
    m = oracle.read_obj(TUBE)
    s = oracle.System(m, bc=0)
    assert (s.nv, s.ntri, int(m["on_boundary"].sum()), s.n, s.nnz) == (4704, 9216, 192, 4512, 31200)
    assert oracle.System(m, bc=1).nnz == 32544
    idx, rp, ci = brute_pattern(m, 0)
    assert np.array_equal(s.row_ptr, rp) and np.array_equal(s.col_idx, ci)


def test_smooth_normals_oracle_closed_forms():
    """Flat mesh, zero values: each vertex sums 2 x (area of its incident
    triangles) along +z (counter-clockwise triangles).  Icosphere with zero
    values: every summed normal points outward.  Extrusion clamps at 1."""
    m = meshes.jittered_square(9, seed=3)
    x = m["xyz"].astype(np.float64)
    t = m["tri"].astype(np.int64)
    nrm = np.tile([0.0, 0.0, 1.0], (x.shape[0], 1))
    out = oracle.smooth_normals(x, t, nrm, np.zeros(x.shape[0]), 5.0)
    e1, e2 = x[t[:, 1]] - x[t[:, 0]], x[t[:, 2]] - x[t[:, 0]]
    area2 = e1[:, 0] * e2[:, 1] - e1[:, 1] * e2[:, 0]
    expect = np.zeros(x.shape[0])
    for k in range(3):
        np.add.at(expect, t[:, k], area2)
    assert np.all(area2 > 0)
    assert np.allclose(out[:, 2], expect, rtol=1e-5, atol=1e-7) and np.allclose(out[:, :2], 0.0, atol=1e-7)
    # uniform extrusion of a flat mesh along +z is a translation: same normals
    moved = oracle.smooth_normals(x, t, nrm, np.full(x.shape[0], 0.3), 2.0)
    assert np.allclose(moved, out, rtol=1e-5, atol=1e-7)
    # values beyond 1/extrusion clamp to a unit move: same as value 1/extrusion
    big = oracle.smooth_normals(x, t, nrm, np.full(x.shape[0], 7.0), 2.0)
    assert np.allclose(big, out, rtol=1e-5, atol=1e-7)
    s = meshes.icosphere(2)
    p = s["xyz"].astype(np.float64)
    n_out = oracle.smooth_normals(p, s["tri"], p / np.linalg.norm(p, axis=1, keepdims=True), np.zeros(p.shape[0]), 1.0)
    assert np.all((n_out * p).sum(axis=1) > 0)


def test_ply_clip_oracle_cases():
    """The restatement of Surface::export_to_ply's clipping on one triangle:
    the three clipping cases, the closed cap, nothing above the threshold,
    and Viridis's endpoints."""
    xyz = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    nrm = np.tile(np.float32([0, 0, 1]), (3, 1))
    tri = np.array([[0, 1, 2]])
    # all above: one triangle, positions extruded by value * extrusion
    pos, val, f = oracle.ply_clip(xyz, tri, nrm, [1.0, 2.0, 3.0], 0.5, 0.0)
    assert f.tolist() == [[0, 1, 2]] and np.allclose(pos[:, 2], [0.5, 1.0, 1.5])
    # one above (vertex 0 at 1, others 0, threshold 0.5): the cut points sit half way
    pos, val, f = oracle.ply_clip(xyz, tri, nrm, [1.0, 0.0, 0.0], 1.0, 0.5)
    assert len(f) == 1 and len(pos) == 3
    assert np.allclose(sorted(map(tuple, pos.tolist())), sorted([(0, 0, 1), (0.5, 0, 0.5), (0, 0.5, 0.5)]))
    assert np.allclose(sorted(val.tolist()), [0.5, 0.5, 1.0])
    # two above: a quad split in two triangles; CLOSED adds the projected cap (2 more faces)
    pos, val, f = oracle.ply_clip(xyz, tri, nrm, [1.0, 1.0, 0.0], 1.0, 0.5)
    assert len(f) == 2 and len(pos) == 4
    assert f.tolist() == [[2, 0, 3], [3, 0, 1]]
    pos_c, _, f_c = oracle.ply_clip(xyz, tri, nrm, [1.0, 1.0, 0.0], 1.0, 0.5, oracle.CLOSED)
    assert len(f_c) == 4 and len(pos_c) == 6
    # closed cap: the kept vertices projected to threshold * extrusion above the plane
    assert np.allclose(pos_c[4:, 2], 0.5)
    # none above: nothing
    pos, val, f = oracle.ply_clip(xyz, tri, nrm, [0.0, 0.0, 0.0], 1.0, 0.5)
    assert len(f) == 0 and len(pos) == 0
    # viridis endpoints: dark purple at 0, yellow at 1
    c0, c1 = oracle.viridis(0.0), oracle.viridis(1.0)
    assert c0[2] > c0[1] and c1[0] > 0.9 and c1[1] > 0.85


def test_read_obj_oracle_hand_built(tmp_path):
    """The OBJ restatement on a hand-built file: a triangle with v//n
    corners, a quad split along its shorter diagonal either way, a pentagon
    fanned, negative indices, two objects (the second repeats the vertex
    list), normals summed per corner and normalised, edge-count boundary."""
    text = "\n".join([
        "# hand-built", "o first",
        "v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0", "v 2 0 0", "v 3 0.2 0",
        "vn 0 0 1", "vn 0 0 2",
        "f 1//1 2//1 3//2",          # triangle; vertex 3 gets (0,0,1) + (0,0,2) below
        "f 2//1 5//1 6//1 3//2",     # quad: |6-2|^2 = 4.04 > |3-5|^2 = 2 -> [0,1,3],[1,2,3]
        "o second",
        "f -6//-2 -4//-2 -3//-1",    # negative indices: vertices 1, 3, 4
        "f 1//1 3//1 4//1 5//1 6//1",  # pentagon, fanned from vertex 1
    ]) + "\n"
    path = tmp_path / "hand.obj"
    path.write_text(text)
    m = oracle.read_obj(str(path))
    assert m["xyz"].shape == (12, 3)          # two objects: the vertex list twice
    assert np.array_equal(m["xyz"][:6], m["xyz"][6:])
    assert m["tri"].tolist() == [[0, 1, 2], [1, 4, 2], [4, 5, 2], [0, 2, 3], [0, 2, 3], [0, 3, 4], [0, 4, 5]]
    # every used vertex's normal is +z; unused copies are NaN
    assert np.allclose(m["normals"][:6], [0, 0, 1])
    assert np.all(np.isnan(m["normals"][6:]))
    assert m["on_boundary"][6:].sum() == 0


def test_gl_step_oracle_pins():
    """The GL-semantics restatement: with enough iterations its warm-started,
    unpreconditioned CG reaches the direct solution of the same system; with
    the reference's 10 iterations it stops early on a fine mesh; its
    Gray-Scott RHS sums the reaction over the row's slots."""
    import scipy.sparse.linalg as sla
    m = meshes.jittered_square(12, seed=2)
    s = oracle.System(m, bc=0)
    u0 = s.gather(np.exp(-((m["xyz"][:, 0] - 0.5) ** 2 + (m["xyz"][:, 1] - 0.5) ** 2) / 0.02))
    dt, c = 1e-2, 1.0
    S = (s.scipy("M") / dt + c * s.scipy("K")).tocsc()
    exact = sla.spsolve(S, s.scipy("M") @ u0 / dt)
    u, _, rr = oracle.gl_step(s, oracle.HEAT, [dt, c], u0, np.zeros(s.n), max_it=500)
    assert np.linalg.norm(u - exact) <= 1e-5 * np.linalg.norm(exact)
    u10, _, rr10 = oracle.gl_step(s, oracle.HEAT, [dt, c], u0, np.zeros(s.n), max_it=2)
    assert rr10 > rr and np.linalg.norm(u10 - exact) > np.linalg.norm(u - exact)
    # Gray-Scott from u = v = 0: the u right-hand side is F x (slots of the
    # row), and one CG step from x = 0 is alpha b with alpha = b.b / b.S b;
    # the v system then has a zero right-hand side
    F = 0.035
    gu, gv, _ = oracle.gl_step(s, oracle.GRAYSCOTT, [1e-3, 0.08, 0.04, F, 0.06], np.zeros(s.n), np.zeros(s.n),
                               max_it=1)
    b = F * np.diff(s.row_ptr).astype(np.float64)
    Su = (s.scipy("M") / 1e-3 + 0.08 * s.scipy("K"))
    alpha = (b @ b) / (b @ (Su @ b))
    assert np.allclose(gu, alpha * b, rtol=1e-5) and not np.any(gv)
