"""The C++ Solver adapter (include/femhip_solver.hpp) through its headless
driver (tests/cpp/solver_driver.cpp), which replays the reference's frame
loop: gather surface values -> advance one step -> instability check ->
scatter (Application.cpp:530-540, CPUSolver.cpp:13-37).

The surface carries float32 values (Surface::values), so the driver's state
is rounded to float32 between frames exactly as in the reference; the oracle
comparison applies the same rounding."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from femhip import build as fb
from femhip import meshes

DRIVER = fb.DRIVER


def write_input(path, mesh, bc, values):
    with open(path, "wb") as f:
        np.array([mesh["xyz"].shape[0], mesh["tri"].shape[0], bc], dtype=np.int32).tofile(f)
        np.ascontiguousarray(mesh["xyz"], dtype=np.float32).tofile(f)
        np.ascontiguousarray(mesh["tri"], dtype=np.uint32).tofile(f)
        np.ascontiguousarray(mesh["on_boundary"], dtype=np.uint8).tofile(f)
        np.ascontiguousarray(values, dtype=np.float32).tofile(f)


def read_output(path, nv):
    raw = np.fromfile(path, dtype=np.uint8)
    head = raw[:8].view(np.int32)
    return int(head[0]), int(head[1]), raw[8:8 + 4 * nv].view(np.float32)


def test_adapter_compiles_links_and_reports_errors(tmp_path):
    """The header compiles warning-free against the C ABI and the driver links
    every symbol it uses; without a usable device the library's error is
    surfaced as an exception message and a non-zero exit."""
    exe = str(tmp_path / "solver_driver")
    cmd = fb.driver_command(out=exe) + ["-Werror"]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    res = subprocess.run([exe], capture_output=True, text=True)
    assert res.returncode == 2 and "usage" in res.stderr
    res = subprocess.run([exe, str(tmp_path / "missing.bin"), str(tmp_path / "o.bin"), "heat", "1", "0.01", "1"],
                         capture_output=True, text=True)
    assert res.returncode == 1 and "cannot open input" in res.stderr


def run_driver(tmp_path, mesh, bc, values, equation, steps, *params):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is missing: run __graft_entry__.build() first")
    fin, fout = tmp_path / "in.bin", tmp_path / "out.bin"
    write_input(fin, mesh, bc, values)
    args = [p if isinstance(p, str) else repr(float(p)) for p in params]
    res = subprocess.run([DRIVER, str(fin), str(fout), equation, str(steps)] + args,
                         capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    return read_output(fout, mesh["xyz"].shape[0])


def oracle_frames(s, equation, prm, steps, field0, state, switch=None, repaint=False):
    """The reference loop with float32 surface values between frames;
    switch = (frame, system): from that frame on, step with the other
    system (re-assembled matrices); repaint: at that frame the solver's state
    is cleared and the surface painted again with field0 (a boundary-condition
    change: clear_solver, then a new unknown numbering)."""
    u, v = state
    surf = field0.astype(np.float32)
    for f in range(steps):
        if switch is not None and f == switch[0]:
            s = switch[1]
            if repaint:
                u, v = np.zeros(s.n), np.zeros(s.n)
                surf = field0.astype(np.float32)
        x = s.gather(surf.astype(np.float64))
        if equation == oracle.GRAYSCOTT:
            v = x
        else:
            u = x
        _, u, v, _ = s.step(equation, prm, u, v, tol=1e-14)
        surf = s.scatter(v if equation == oracle.GRAYSCOTT else u).astype(np.float32)
    return surf


@pytest.mark.gpu
@pytest.mark.parametrize("eq", ["heat", "wave"])
def test_driver_frame_loop_matches_oracle(gpu, tmp_path, eq):
    m = meshes.structured_square(20)
    s = oracle.System(m, bc=0)
    x = m["xyz"].astype(np.float64)
    vals = np.exp(-((x[:, 0] - 0.4) ** 2 + (x[:, 1] - 0.55) ** 2) / 0.02)
    steps = 8
    prm = [1e-3, 1.0] if eq == "heat" else [5e-3, 0.8]
    code = oracle.HEAT if eq == "heat" else oracle.WAVE
    print("case driver", eq, flush=True)
    frames, unstable, got = run_driver(tmp_path, m, 0, vals, eq, steps, *prm)
    assert frames == steps and unstable == 0
    ref = oracle_frames(s, code, prm, steps, vals, (np.zeros(s.n), np.zeros(s.n)))
    onb = m["on_boundary"].astype(bool)
    assert np.all(got[onb] == 0.0)
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)


@pytest.mark.gpu
@pytest.mark.parametrize("eq", ["heat", "wave"])
@pytest.mark.parametrize("bcs", [(0, 1), (1, 0)])
def test_driver_boundary_condition_switch(gpu, tmp_path, eq, bcs):
    """A boundary-condition change after 5 of 10 frames (Application.cpp:313-315:
    clear_solver, then update_boundary_conditions, which re-numbers the
    unknowns): the adapter rebuilds its context and solver, the surface is
    painted again, and the frames continue with the other unknown count,
    against the oracle doing the same."""
    m = meshes.jittered_square(20, seed=9)
    x = m["xyz"].astype(np.float64)
    vals = np.exp(-((x[:, 0] - 0.3) ** 2 + (x[:, 1] - 0.45) ** 2) / 0.03)
    prm = [1e-3, 1.0] if eq == "heat" else [5e-3, 0.8]
    code = oracle.HEAT if eq == "heat" else oracle.WAVE
    print("case driver bc switch", eq, bcs, flush=True)
    frames, unstable, got = run_driver(tmp_path, m, bcs[0], vals, eq, 10, *prm, "--bc-at", 5, bcs[1])
    assert frames == 10 and unstable == 0
    s1, s2 = oracle.System(m, bc=bcs[0]), oracle.System(m, bc=bcs[1])
    assert s1.n != s2.n
    ref = oracle_frames(s1, code, prm, 10, vals, (np.zeros(s1.n), np.zeros(s1.n)), switch=(5, s2), repaint=True)
    onb = m["on_boundary"].astype(bool)
    if bcs[1] == 0:
        assert np.all(got[onb] == 0.0)
    else:
        assert np.any(got[onb] != 0.0)
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    # staying on the first condition gives another field
    stay = oracle_frames(s1, code, prm, 10, vals, (np.zeros(s1.n), np.zeros(s1.n)), switch=(5, s1), repaint=True)
    assert np.linalg.norm(got - stay) > 1e-4 * np.linalg.norm(ref)


@pytest.mark.gpu
def test_driver_grayscott_nodal_instability_clears(gpu, tmp_path):
    """Reference Gray-Scott (nodal reaction) at dt = 0.5 on a fine sphere
    blows up; the loop sees has_numerical_instability(), clears and stops."""
    m = meshes.icosphere(4)
    vals = np.where(np.arange(m["xyz"].shape[0]) % 11 == 0, 0.25, 0.0)
    frames, unstable, got = run_driver(tmp_path, m, 1, vals, "grayscott", 20, 0.5, 2e-4, 1e-4, 0.037, 0.06)
    assert unstable == 1 and frames < 20
    assert not np.any(got)


@pytest.mark.gpu
def test_driver_advdiff_velocity_change(gpu, tmp_path):
    """Advection-diffusion through the adapter (BiCGSTAB), with the velocity
    changed and the matrices re-assembled after 5 of 10 frames, against the
    oracle stepping with the matching matrices."""
    m = meshes.jittered_square(24, seed=8)
    x = m["xyz"].astype(np.float64)
    vals = np.exp(-((x[:, 0] - 0.4) ** 2 + (x[:, 1] - 0.5) ** 2) / 0.02)
    v2 = (-0.3, 0.9, 0.0)
    prm = [2e-3, 0.05]
    print("case driver advdiff velocity change", flush=True)
    frames, unstable, got = run_driver(tmp_path, m, 0, vals, "advdiff", 10, *prm, "--velocity-at", 5, *v2)
    assert frames == 10 and unstable == 0
    s1 = oracle.System(m, bc=0, velocity=(1.0, 0.0, 0.0))
    s2 = oracle.System(m, bc=0, velocity=v2)
    ref = oracle_frames(s1, oracle.ADVDIFF, prm, 10, vals, (np.zeros(s1.n), np.zeros(s1.n)), switch=(5, s2))
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
    # without the re-assembly the result differs
    stale = oracle_frames(s1, oracle.ADVDIFF, prm, 10, vals, (np.zeros(s1.n), np.zeros(s1.n)))
    assert np.linalg.norm(got - stale) > 1e-3 * np.linalg.norm(ref)


@pytest.mark.gpu
def test_driver_grayscott_lumped_stable(gpu, tmp_path):
    """Gray-Scott with the lumped weighting at dt = 0.5 stays stable through
    the adapter; the surface carries v (CPUSolver.cpp:130,144) and u stays on
    the device, starting from 0 as after clear_values (CPUSolver.cpp:54-59)."""
    m = meshes.icosphere(3)
    s = oracle.System(m, bc=1)
    vals = np.where(np.arange(m["xyz"].shape[0]) % 7 == 0, 0.25, 0.0)
    prm = [0.5, 2e-4, 1e-4, 0.037, 0.06, 1]
    print("case driver grayscott lumped", flush=True)
    frames, unstable, got = run_driver(tmp_path, m, 1, vals, "grayscott", 12, *prm)
    assert frames == 12 and unstable == 0
    ref = oracle_frames(s, oracle.GRAYSCOTT, prm, 12, vals, (np.zeros(s.n), np.zeros(s.n)))
    assert np.linalg.norm(got - ref) <= 1e-5 * np.linalg.norm(ref)
