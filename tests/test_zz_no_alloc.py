"""fem_ctx_solve_noalloc (the allocation-free solve) against fem_ctx_solve,
through the C ABI: two calls in a row, the second on a smaller system, each
bit-identical to the allocating entry point (same kernels, same plan, same
reduction order) and within 1e-8 of the oracle's FP64 solve; the scratch
limit is refused with a clear error before any device work."""
import ctypes as C

import numpy as np
import pytest

import oracle
from femhip import FEMContext, SolverOptions, _lib, meshes

pytestmark = pytest.mark.gpu


def solve(lib, ctx, fn, coef, b, opts):
    o = opts.to_c()
    k = 1 if b.ndim == 1 else b.shape[1]
    o.nrhs = k
    o.warm_start = 0
    bb = np.ascontiguousarray(b, dtype=np.float64)
    x = np.full_like(bb, np.nan)
    cf = np.ascontiguousarray(coef, dtype=np.float64)
    st = _lib.fem_step_stats()
    rc = getattr(lib, fn)(ctx.handle, _lib.ptr(cf, C.c_double), C.byref(o), _lib.ptr(bb, C.c_double),
                          _lib.ptr(x, C.c_double), C.byref(st))
    return rc, x, st


CASES = [
    # (mesh, bc, coef, nrhs, options): the second of each pair is smaller
    ("jitter96_heat", lambda: meshes.jittered_square(96, seed=5), 0, (1000.0, 1.0, 0.0), 1, {}),
    ("jitter40_heat", lambda: meshes.jittered_square(40, seed=6), 0, (1000.0, 1.0, 0.0), 1, {}),
    ("jitter64_advdiff_4rhs", lambda: meshes.jittered_square(64, seed=7), 0, (500.0, 0.1, -1.0), 4, {}),
    ("ico3_fp64_16rhs", lambda: meshes.icosphere(3), 1, (2.0, 1e-3, 0.0), 16,
     {"precision": _lib.FEM_PREC_FP64, "tol": 1e-12}),
]


def test_noalloc_matches_solve_twice_second_smaller(gpu):
    lib = _lib.load()
    for pair in ((CASES[0], CASES[1]), (CASES[2], CASES[3])):
        for name, mk, bc, coef, k, kw in pair:
            print("case noalloc", name, flush=True)
            m = mk()
            ctx = FEMContext(m, bc, velocity=(0.4, 0.7, 0.0), normalize_velocity=False)
            s = oracle.System(m, bc=bc, velocity=(0.4, 0.7, 0.0), normalize=False)
            rng = np.random.default_rng(11)
            b = rng.standard_normal((s.n, k)) if k > 1 else rng.standard_normal(s.n)
            opts = SolverOptions(**kw)
            rc0, x0, st0 = solve(lib, ctx, "fem_ctx_solve", coef, b, opts)
            rc1, x1, st1 = solve(lib, ctx, "fem_ctx_solve_noalloc", coef, b, opts)
            assert rc0 == _lib.FEM_OK and rc1 == _lib.FEM_OK, (rc0, rc1, _lib.last_error())
            assert np.array_equal(x0, x1), name
            assert (st0.iterations, st0.sweeps) == (st1.iterations, st1.sweeps)
            S = coef[0] * s.M + coef[1] * s.K + coef[2] * s.A
            bc_ = b.reshape(s.n, -1)
            xr = x1.reshape(s.n, -1)
            for j in range(bc_.shape[1]):
                if coef[2] != 0.0:
                    ref, _, _ = s.bicgstab(S, bc_[:, j], tol=1e-14)
                else:
                    ref, _, _ = s.cg(S, bc_[:, j], tol=1e-14)
                assert np.linalg.norm(xr[:, j] - ref) <= 1e-8 * np.linalg.norm(ref)
            ctx.close()


def test_noalloc_refuses_beyond_scratch(gpu):
    lib = _lib.load()
    m = meshes.jittered_square(700, seed=1)  # ~490,000 unknowns: beyond the 32 MiB scratch
    ctx = FEMContext(m, 0)
    b = np.ones(ctx.num_unknowns())
    rc, x, _ = solve(lib, ctx, "fem_ctx_solve_noalloc", (1.0, 1e-3, 0.0), b, SolverOptions())
    assert rc == _lib.FEM_ERR_INVALID
    assert "bytes of scratch" in _lib.last_error()
    assert np.all(np.isnan(x))  # nothing was written
    ctx.close()
