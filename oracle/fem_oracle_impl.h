/*
 * fem_oracle_impl.h -- type-generic body of the oracle, included twice by
 * fem_oracle.c with REAL = double (SFX _d) and REAL = float (SFX _f).
 * TEST INFRASTRUCTURE ONLY (see fem_oracle.h).
 */
#define CAT2(a, b) a##b
#define CAT(a, b) CAT2(a, b)
#define FN(name) CAT(name, SFX)

/* Element geometry: surface gradients of the three P1 basis functions, area
 * and |J| = |(b-a)x(c-a)|.  FEMContext.cpp:73-95 (stiffness) and :159-184
 * (advection): J = [b-a, c-a, n], grad_i = J^{-T} g_i with reference
 * gradients g = {(-1,-1,0),(1,0,0),(0,1,0)}.  J^{-1} by cofactors. */
static void FN(elem_geom)(const float *xyz, const uint32_t *t, REAL g[3][3], REAL *area,
                          REAL *jdet, REAL nrm[3]) {
    REAL a[3], b[3], c[3], e1[3], e2[3], cr[3];
    for (int k = 0; k < 3; k++) {
        a[k] = (REAL)xyz[3 * (size_t)t[0] + k];
        b[k] = (REAL)xyz[3 * (size_t)t[1] + k];
        c[k] = (REAL)xyz[3 * (size_t)t[2] + k];
        e1[k] = b[k] - a[k];
        e2[k] = c[k] - a[k];
    }
    cr[0] = e1[1] * e2[2] - e1[2] * e2[1];
    cr[1] = e1[2] * e2[0] - e1[0] * e2[2];
    cr[2] = e1[0] * e2[1] - e1[1] * e2[0];
    REAL len = SQRT(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
    for (int k = 0; k < 3; k++) nrm[k] = cr[k] / len;
    *area = (REAL)0.5 * len;
    *jdet = len;
    /* m = J (rows x,y,z; columns e1, e2, n) */
    REAL m00 = e1[0], m01 = e2[0], m02 = nrm[0];
    REAL m10 = e1[1], m11 = e2[1], m12 = nrm[1];
    REAL m20 = e1[2], m21 = e2[2], m22 = nrm[2];
    REAL c00 = m11 * m22 - m12 * m21;
    REAL c01 = m12 * m20 - m10 * m22;
    REAL c02 = m10 * m21 - m11 * m20;
    REAL det = m00 * c00 + m01 * c01 + m02 * c02;
    /* rows 0 and 1 of J^{-1} are the gradients of phi_1 and phi_2 */
    g[1][0] = c00 / det;
    g[1][1] = (m02 * m21 - m01 * m22) / det;
    g[1][2] = (m01 * m12 - m02 * m11) / det;
    g[2][0] = c01 / det;
    g[2][1] = (m00 * m22 - m02 * m20) / det;
    g[2][2] = (m02 * m10 - m00 * m12) / det;
    for (int k = 0; k < 3; k++) g[0][k] = -g[1][k] - g[2][k];
}

static int32_t FN(find_col)(const int32_t *ci, int32_t lo, int32_t hi, int32_t col) {
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (ci[mid] < col) lo = mid + 1; else hi = mid;
    }
    return lo;
}

/* FEMContext::assemble_{stiffness,mass,advection}_matrix restated onto the
 * shared CSR pattern.  Summation per entry follows triplet order (element k,
 * then i, then j), as setFromTriplets does. */
void FN(orc_assemble)(int32_t nv, int32_t ntri, const float *xyz, const uint32_t *tri,
                      const uint8_t *onb, int32_t bc, const int32_t *idx_map,
                      const int32_t *rp, const int32_t *ci, const double *velocity,
                      const double *vel_e, int32_t normalize, const double *kappa_e, REAL *K,
                      REAL *M, REAL *A) {
    (void)nv;
    int32_t n = 0;
    for (int32_t i = 0; i < nv; i++) if (idx_map[i] >= n) n = idx_map[i] + 1;
    int64_t nnz = rp[n];
    for (int64_t s = 0; s < nnz; s++) { K[s] = 0; M[s] = 0; A[s] = 0; }
    for (int32_t k = 0; k < ntri; k++) {
        const uint32_t *t = tri + 3 * (size_t)k;
        REAL g[3][3], area, jdet, nrm[3];
        FN(elem_geom)(xyz, t, g, &area, &jdet, nrm);
        REAL kap = kappa_e ? (REAL)kappa_e[k] : (REAL)1;
        /* mass: (|J|/24) * [[2,1,1],[1,2,1],[1,1,2]]  FEMContext.cpp:126-132 */
        REAL mo = (REAL)((double)jdet / 24.0);
        /* advection velocity: tangential projection, optionally normalised
         * FEMContext.cpp:179-180 */
        REAL vel[3];
        for (int q = 0; q < 3; q++) vel[q] = (REAL)(vel_e ? vel_e[3 * (size_t)k + q] : velocity[q]);
        REAL nn = nrm[0] * nrm[0] + nrm[1] * nrm[1] + nrm[2] * nrm[2];
        REAL vd = (vel[0] * nrm[0] + vel[1] * nrm[1] + vel[2] * nrm[2]) / nn;
        REAL tv[3];
        for (int q = 0; q < 3; q++) tv[q] = vel[q] - vd * nrm[q];
        if (normalize) {
            REAL z = tv[0] * tv[0] + tv[1] * tv[1] + tv[2] * tv[2];
            if (z > 0) { REAL s = SQRT(z); for (int q = 0; q < 3; q++) tv[q] = tv[q] / s; }
        }
        REAL jd6 = jdet / (REAL)6;
        for (int i = 0; i < 3; i++) {
            for (int j = 0; j < 3; j++) {
                int bi = onb[t[i]] != 0, bj = onb[t[j]] != 0;
                int keep_km = (bc == 1) || (!bi && !bj);   /* :99, :136 */
                int keep_a = !bi && !bj;                    /* :188 regardless of BC */
                if (!keep_km && !keep_a) continue;
                int32_t r = idx_map[t[i]], c = idx_map[t[j]];
                int32_t s = FN(find_col)(ci, rp[r], rp[r + 1], c);
                if (keep_km) {
                    REAL kij = area * (g[i][0] * g[j][0] + g[i][1] * g[j][1] + g[i][2] * g[j][2]);
                    K[s] += kap * kij;
                    M[s] += (i == j) ? (REAL)2 * mo : mo;
                }
                if (keep_a) {
                    A[s] += jd6 * (tv[0] * g[j][0] + tv[1] * g[j][1] + tv[2] * g[j][2]);
                }
            }
        }
    }
}

/* y = S x on CSR */
static void FN(csr_mv)(int32_t n, const int32_t *rp, const int32_t *ci, const REAL *S,
                       const REAL *x, REAL *y) {
    for (int32_t r = 0; r < n; r++) {
        REAL acc = 0;
        for (int32_t s = rp[r]; s < rp[r + 1]; s++) acc += S[s] * x[ci[s]];
        y[r] = acc;
    }
}

static REAL FN(dot)(int32_t n, const REAL *a, const REAL *b) {
    REAL acc = 0;
    for (int32_t i = 0; i < n; i++) acc += a[i] * b[i];
    return acc;
}

/* Eigen DiagonalPreconditioner::factorize: 1/a_jj, or 1 when absent/zero. */
static void FN(jacobi)(int32_t n, const int32_t *rp, const int32_t *ci, const REAL *S, REAL *inv) {
    for (int32_t r = 0; r < n; r++) {
        inv[r] = 1;
        for (int32_t s = rp[r]; s < rp[r + 1]; s++)
            if (ci[s] == r && S[s] != 0) inv[r] = (REAL)1 / S[s];
    }
}

/* Eigen conjugate_gradient() restated (ConjugateGradient.h). */
int32_t FN(orc_cg)(int32_t n, const int32_t *rp, const int32_t *ci, const REAL *S, const REAL *b,
                   REAL *x, REAL tol, int32_t maxit, REAL *err) {
    REAL *r = (REAL *)malloc(sizeof(REAL) * (size_t)n * 4 + 64);
    REAL *p = r + n, *z = p + n, *tmp = z + n;
    REAL *inv = (REAL *)malloc(sizeof(REAL) * (size_t)n + 64);
    FN(jacobi)(n, rp, ci, S, inv);
    FN(csr_mv)(n, rp, ci, S, x, tmp);
    for (int32_t i = 0; i < n; i++) r[i] = b[i] - tmp[i];
    REAL rhs2 = FN(dot)(n, b, b);
    int32_t it = 0;
    if (rhs2 == 0) {
        for (int32_t i = 0; i < n; i++) x[i] = 0;
        *err = 0;
        goto done;
    }
    {
        REAL thr = tol * tol * rhs2;
        if (thr < REALMIN) thr = REALMIN;
        REAL res2 = FN(dot)(n, r, r);
        if (res2 < thr) { *err = SQRT(res2 / rhs2); goto done; }
        for (int32_t i = 0; i < n; i++) p[i] = inv[i] * r[i];
        REAL absnew = FN(dot)(n, r, p);
        while (it < maxit) {
            FN(csr_mv)(n, rp, ci, S, p, tmp);
            REAL alpha = absnew / FN(dot)(n, p, tmp);
            for (int32_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * tmp[i]; }
            res2 = FN(dot)(n, r, r);
            if (res2 < thr) break;
            for (int32_t i = 0; i < n; i++) z[i] = inv[i] * r[i];
            REAL absold = absnew;
            absnew = FN(dot)(n, r, z);
            REAL beta = absnew / absold;
            for (int32_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
            it++;
        }
        *err = SQRT(res2 / rhs2);
    }
done:
    free(r);
    free(inv);
    return it;
}

/* Right-Jacobi-preconditioned BiCGSTAB; stop when |r|^2 < tol^2 |b|^2.
 * Restarts (fresh shadow residual) on breakdown. */
static int32_t FN(bicgstab)(int32_t n, const int32_t *rp, const int32_t *ci, const REAL *S,
                            const REAL *b, REAL *x, REAL tol, int32_t maxit, REAL *err) {
    REAL *r = (REAL *)malloc(sizeof(REAL) * (size_t)n * 8 + 64);
    REAL *r0 = r + n, *p = r0 + n, *v = p + n, *s = v + n, *t = s + n, *ph = t + n, *sh = ph + n;
    REAL *inv = (REAL *)malloc(sizeof(REAL) * (size_t)n + 64);
    FN(jacobi)(n, rp, ci, S, inv);
    FN(csr_mv)(n, rp, ci, S, x, t);
    for (int32_t i = 0; i < n; i++) r[i] = b[i] - t[i];
    REAL rhs2 = FN(dot)(n, b, b);
    int32_t it = 0;
    if (rhs2 == 0) { for (int32_t i = 0; i < n; i++) x[i] = 0; *err = 0; goto done; }
    {
        REAL thr = tol * tol * rhs2;
        if (thr < REALMIN) thr = REALMIN;
        REAL res2 = FN(dot)(n, r, r);
        REAL rho = 1, alpha = 1, omega = 1;
        for (int32_t i = 0; i < n; i++) { r0[i] = r[i]; p[i] = 0; v[i] = 0; }
        REAL r0n2 = res2;
        while (res2 >= thr && it < maxit) {
            REAL rho_new = FN(dot)(n, r0, r);
            if (FABS(rho_new) < (REAL)1e-30 * r0n2 || omega == 0) {
                /* restart with the current residual as shadow */
                for (int32_t i = 0; i < n; i++) { r0[i] = r[i]; p[i] = 0; v[i] = 0; }
                r0n2 = FN(dot)(n, r, r);
                rho = alpha = omega = 1;
                rho_new = r0n2;
            }
            REAL beta = (rho_new / rho) * (alpha / omega);
            rho = rho_new;
            for (int32_t i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
            for (int32_t i = 0; i < n; i++) ph[i] = inv[i] * p[i];
            FN(csr_mv)(n, rp, ci, S, ph, v);
            alpha = rho / FN(dot)(n, r0, v);
            for (int32_t i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
            REAL s2 = FN(dot)(n, s, s);
            if (s2 < thr) {
                for (int32_t i = 0; i < n; i++) { x[i] += alpha * ph[i]; r[i] = s[i]; }
                res2 = s2;
                it++;
                break;
            }
            for (int32_t i = 0; i < n; i++) sh[i] = inv[i] * s[i];
            FN(csr_mv)(n, rp, ci, S, sh, t);
            REAL tt = FN(dot)(n, t, t);
            omega = tt > 0 ? FN(dot)(n, t, s) / tt : 0;
            for (int32_t i = 0; i < n; i++) {
                x[i] += alpha * ph[i] + omega * sh[i];
                r[i] = s[i] - omega * t[i];
            }
            res2 = FN(dot)(n, r, r);
            it++;
        }
        *err = SQRT(res2 / rhs2);
    }
done:
    free(r);
    free(inv);
    return it;
}

static int32_t FN(solve)(int32_t krylov, int32_t n, const int32_t *rp, const int32_t *ci,
                         const REAL *S, const REAL *b, REAL *x, double tol, int32_t maxit) {
    REAL err;
    if (maxit <= 0) maxit = 2 * n; /* Eigen default maxIterations = 2*cols */
    if (krylov == 1) return FN(bicgstab)(n, rp, ci, S, b, x, (REAL)tol, maxit, &err);
    return FN(orc_cg)(n, rp, ci, S, b, x, (REAL)tol, maxit, &err);
}

/* CPUSolver::advance_time restated (CPUSolver.cpp:64-147), plus Newmark. */
int32_t FN(orc_step)(int32_t n, const int32_t *rp, const int32_t *ci, const REAL *K,
                     const REAL *M, const REAL *A, int32_t eq, const double *prm, REAL *u,
                     REAL *v, REAL *acc, int32_t krylov, double tol, int32_t maxit,
                     int32_t warm) {
    int64_t nnz = rp[n];
    REAL *S = (REAL *)malloc(sizeof(REAL) * (size_t)(nnz + 1));
    REAL *S2 = (REAL *)malloc(sizeof(REAL) * (size_t)(nnz + 1));
    REAL *b = (REAL *)malloc(sizeof(REAL) * (size_t)n * 4 + 64);
    REAL *x = b + n, *t1 = x + n, *t2 = t1 + n;
    REAL dt = (REAL)prm[0];
    int32_t its = 0;
    switch (eq) {
    case 0: /* heat :72-84  (M/dt + cK) u+ = (M/dt) u */
    case 2: /* advection-diffusion :90-102  (M/dt + cK - A) u+ = (M/dt) u */ {
        REAL c = (REAL)prm[1];
        for (int64_t s = 0; s < nnz; s++) {
            S[s] = M[s] / dt + c * K[s];
            if (eq == 2 && A) S[s] -= A[s];
            S2[s] = M[s] / dt;
        }
        FN(csr_mv)(n, rp, ci, S2, u, b);
        for (int32_t i = 0; i < n; i++) x[i] = warm ? u[i] : 0;
        its = FN(solve)(krylov, n, rp, ci, S, b, x, tol, maxit);
        for (int32_t i = 0; i < n; i++) u[i] = x[i];
    } break;
    case 1: /* wave :108-121  (M/dt + c^2 dt K) v+ = (M/dt) v - c^2 K u ; u += dt v+ */ {
        REAL c = (REAL)prm[1];
        for (int64_t s = 0; s < nnz; s++) {
            S[s] = M[s] / dt + c * c * K[s] * dt;
            S2[s] = M[s] / dt;
        }
        FN(csr_mv)(n, rp, ci, S2, v, t1);
        FN(csr_mv)(n, rp, ci, K, u, t2);
        for (int32_t i = 0; i < n; i++) b[i] = t1[i] - c * c * t2[i];
        for (int32_t i = 0; i < n; i++) x[i] = warm ? v[i] : 0;
        its = FN(solve)(krylov, n, rp, ci, S, b, x, tol, maxit);
        for (int32_t i = 0; i < n; i++) { v[i] = x[i]; u[i] = u[i] + v[i] * dt; }
    } break;
    case 3: /* Gray-Scott :127-145, both RHS from the old u, v */ {
        REAL Du = (REAL)prm[1], Dv = (REAL)prm[2], F = (REAL)prm[3], kk = (REAL)prm[4];
        int lumped = prm[5] != 0;
        REAL *bu = b, *bv = t2;
        REAL *ml = (REAL *)malloc(sizeof(REAL) * (size_t)n + 64);
        for (int32_t r = 0; r < n; r++) {
            REAL s_ = 0;
            for (int32_t s = rp[r]; s < rp[r + 1]; s++) s_ += M[s];
            ml[r] = lumped ? s_ : (REAL)1;
        }
        for (int64_t s = 0; s < nnz; s++) S2[s] = M[s] / dt;
        FN(csr_mv)(n, rp, ci, S2, u, bu);
        FN(csr_mv)(n, rp, ci, S2, v, bv);
        for (int32_t i = 0; i < n; i++) {
            REAL uvv = u[i] * (v[i] * v[i]);
            REAL fu = -uvv + F * ((REAL)1 - u[i]);
            REAL fv = uvv - (F + kk) * v[i];
            bu[i] = bu[i] + ml[i] * fu;
            bv[i] = bv[i] + ml[i] * fv;
        }
        free(ml);
        for (int64_t s = 0; s < nnz; s++) S[s] = M[s] / dt + Du * K[s];
        for (int32_t i = 0; i < n; i++) x[i] = warm ? u[i] : 0;
        its = FN(solve)(krylov, n, rp, ci, S, bu, x, tol, maxit);
        for (int32_t i = 0; i < n; i++) u[i] = x[i];
        for (int64_t s = 0; s < nnz; s++) S[s] = M[s] / dt + Dv * K[s];
        for (int32_t i = 0; i < n; i++) x[i] = warm ? v[i] : 0;
        its += FN(solve)(krylov, n, rp, ci, S, bv, x, tol, maxit);
        for (int32_t i = 0; i < n; i++) v[i] = x[i];
    } break;
    case 4: /* Newmark-beta on M a + c^2 K u = 0 (no reference counterpart):
               (M + beta dt^2 c^2 K) a+ = -c^2 K (u + dt v + (1/2-beta) dt^2 a) */ {
        REAL c = (REAL)prm[1], beta = (REAL)prm[2], gam = (REAL)prm[3];
        REAL c2 = c * c;
        for (int64_t s = 0; s < nnz; s++) S[s] = M[s] + beta * dt * dt * c2 * K[s];
        for (int32_t i = 0; i < n; i++)
            t1[i] = u[i] + dt * v[i] + ((REAL)0.5 - beta) * dt * dt * acc[i];
        FN(csr_mv)(n, rp, ci, K, t1, b);
        for (int32_t i = 0; i < n; i++) b[i] = -c2 * b[i];
        for (int32_t i = 0; i < n; i++) x[i] = warm ? acc[i] : 0;
        its = FN(solve)(krylov, n, rp, ci, S, b, x, tol, maxit);
        for (int32_t i = 0; i < n; i++) {
            u[i] = t1[i] + beta * dt * dt * x[i];
            v[i] = v[i] + dt * (((REAL)1 - gam) * acc[i] + gam * x[i]);
            acc[i] = x[i];
        }
    } break;
    default:
        its = -1;
    }
    free(S);
    free(S2);
    free(b);
    return its;
}

#undef FN
#undef CAT
#undef CAT2
