"""The library's surface input and output (host code, no GPU):
fem_surface_from_obj against the reader's restatement in the oracle and
the golden arrays, fem_export_ply against the clipping restatement.  Both
sides run in float32 with the reference's operation order, so arrays and
written numbers compare bit for bit."""
import os

import numpy as np
import pytest

import oracle
from femhip import export, meshes

GOLD = os.path.join(os.path.dirname(__file__), "golden")
REF = "/root/reference/assets/fem_meshes"


def same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.shape == b.shape and a.dtype == b.dtype and np.array_equal(a, b, equal_nan=a.dtype.kind == "f")


def check_reader(path):
    got = meshes.read_obj(path)
    ref = oracle.read_obj(path)
    for key in ("xyz", "tri", "on_boundary", "normals"):
        assert same(got[key], ref[key]), key
    return got


def hand_built(tmp_path):
    text = "\n".join([
        "# hand-built", "o first",
        "v 0 0 0", "v 1 0 0", "v 1 1 0", "v 0 1 0", "v 2 0 0", "v 3 0.2 0",
        "vn 0 0 1", "vn 0 0 2",
        "f 1//1 2//1 3//2",
        "f 2//1 5//1 6//1 3//2",
        "o second",
        "f -6//-2 -4//-2 -3//-1",
        "f 1//1 3//1 4//1 5//1 6//1",
    ]) + "\n"
    p = tmp_path / "hand.obj"
    p.write_text(text)
    return str(p)


def test_reader_hand_built(tmp_path):
    m = check_reader(hand_built(tmp_path))
    assert m["tri"].tolist() == [[0, 1, 2], [1, 4, 2], [4, 5, 2], [0, 2, 3], [0, 2, 3], [0, 3, 4], [0, 4, 5]]


def write_obj(path, xyz, tri, quads=None, fmt="{:.9g}"):
    """An OBJ with per-face normals (v//n corners), optional quads."""
    X = np.asarray(xyz, dtype=np.float64)
    with open(path, "w") as f:
        for p in X:
            f.write("v " + " ".join(fmt.format(c) for c in p) + "\n")
        faces = [list(t) for t in tri] + ([list(q) for q in quads] if quads is not None else [])
        for k, fc in enumerate(faces):
            a, b, c = X[fc[0]], X[fc[1]], X[fc[2]]
            n = np.cross(b - a, c - a)
            f.write("vn " + " ".join(fmt.format(c) for c in n / np.linalg.norm(n)) + "\n")
        for k, fc in enumerate(faces):
            f.write("f " + " ".join(f"{i + 1}//{k + 1}" for i in fc) + "\n")


def test_reader_synthetic_jittered_and_quads(tmp_path):
    """A jittered lattice written with 9-digit coordinates (not exactly
    representable), plus a strip of quads of both diagonal orders."""
    m = meshes.jittered_square(20, seed=9)
    rng = np.random.default_rng(0)
    xyz = m["xyz"].astype(np.float64) + rng.normal(scale=1e-3, size=m["xyz"].shape)
    n1 = 21
    quads = [[j * n1 + i, j * n1 + i + 1, (j + 1) * n1 + i + 1, (j + 1) * n1 + i] for j in range(3) for i in range(5)]
    p = str(tmp_path / "jit.obj")
    write_obj(p, xyz, m["tri"][:300], quads)
    got = check_reader(p)
    assert got["tri"].shape == (300 + 2 * len(quads), 3)


def test_reader_errors(tmp_path):
    from femhip import _lib
    with pytest.raises(_lib.FemError, match="A valid file was not provided"):
        meshes.read_obj(str(tmp_path / "missing.obj"))
    p = tmp_path / "nonormals.obj"
    p.write_text("v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n")
    with pytest.raises(_lib.FemError, match="does not contain normals"):
        meshes.read_obj(str(p))
    p = tmp_path / "empty.obj"
    p.write_text("# nothing\n")
    with pytest.raises(_lib.FemError, match="does not contain a valid .obj mesh"):
        meshes.read_obj(str(p))
    p = tmp_path / "bad.obj"
    p.write_text("v 0 0 0\nvn 0 0 1\nf 1//1 2//1 7//1\n")
    with pytest.raises(_lib.FemError, match="out of range"):
        meshes.read_obj(str(p))


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "icosphere.obj")), reason="reference assets not mounted")
def test_reader_shipped_icosphere_matches_golden():
    m = check_reader(os.path.join(REF, "icosphere.obj"))
    g = dict(np.load(os.path.join(GOLD, "icosphere_shipped.npz")))
    for key in ("xyz", "tri", "on_boundary"):
        assert same(m[key], g[key]), key
    # per-face normals summed at each vertex point outward on the sphere
    assert np.all((m["normals"] * m["xyz"]).sum(1) > 0.99 * np.linalg.norm(m["xyz"], axis=1))


@pytest.mark.skipif(not os.path.exists(os.path.join(REF, "tube.obj")), reason="reference assets not mounted")
def test_reader_shipped_tube():
    m = check_reader(os.path.join(REF, "tube.obj"))
    assert (m["xyz"].shape[0], m["tri"].shape[0], int(m["on_boundary"].sum())) == (4704, 9216, 192)


def read_ply_exact(path):
    """Header lines, float32 vertex rows and faces of an ASCII PLY."""
    with open(path) as f:
        lines = f.read().splitlines()
    h = lines.index("end_header") + 1
    nv = int(lines[2].split()[-1])
    nf = int(lines[9].split()[-1])
    rows = np.array([[np.float32(z) for z in l.split()] for l in lines[h:h + nv]], dtype=np.float32).reshape(-1, 6)
    faces = np.array([[int(z) for z in l.split()] for l in lines[h + nv:h + nv + nf]], dtype=np.int64).reshape(-1, 4)
    assert len(lines) == h + nv + nf
    return lines[:h], rows, faces


def check_export(tmp_path, xyz, tri, nrm, vals, ext, thr, mt):
    p = str(tmp_path / f"out{mt}.ply")
    export.export_ply(p, xyz, tri, nrm, vals, ext, thr, mt)
    head, rows, faces = read_ply_exact(p)
    pos, pval, fref = oracle.ply_clip(xyz, tri, nrm, vals, ext, thr, mt)
    assert head == ["ply", "format ascii 1.0", f"element vertex {len(pos)}", "property float x", "property float y",
                    "property float z", "property float red", "property float green", "property float blue",
                    f"element face {len(fref)}", "property list uchar uint vertex_indices", "end_header"]
    assert same(rows[:, :3], pos)
    col = np.array([oracle.viridis(v) for v in pval], dtype=np.float32).reshape(-1, 3)
    assert same(rows[:, 3:], col)
    assert np.all(faces[:, 0] == 3) and np.array_equal(faces[:, 1:], fref)
    return len(pos), len(fref)


@pytest.mark.parametrize("mt", [oracle.OPEN, oracle.CLOSED, oracle.MIRRORED])
def test_export_matches_oracle(tmp_path, mt):
    """An icosphere with a random field clipped at 0.3: every clipping case
    occurs; positions, colours and faces equal the restatement's."""
    m = meshes.icosphere(3)
    x = m["xyz"].astype(np.float32)
    nrm = x / np.linalg.norm(x, axis=1, keepdims=True)
    vals = np.random.default_rng(4).uniform(0.0, 1.0, x.shape[0]).astype(np.float32)
    nv, nf = check_export(tmp_path, x, m["tri"], nrm, vals, 0.25, 0.3, mt)
    assert 0 < nv and 0 < nf


def test_export_merges_signed_zero(tmp_path):
    """glm::vec3 equality merges -0.0 and 0.0: two triangles whose shared
    corner is (0, 0, -0.0) in one and (0, 0, 0.0) in the other write it once."""
    xyz = np.array([[0, 0, -0.0], [1, 0, 0], [0, 1, 0], [0, 0, 0.0], [-1, 0, 0]], np.float32)
    nrm = np.array([[0, 0, -1]] + [[0, 0, 1]] * 4, np.float32)
    tri = np.array([[0, 1, 2], [3, 2, 4]])
    vals = np.array([0.0, 1.0, 1.0, 0.0, 1.0], np.float32)
    nv, nf = check_export(tmp_path, xyz, tri, nrm, vals, 0.5, -1.0, oracle.OPEN)
    assert (nv, nf) == (4, 2)
