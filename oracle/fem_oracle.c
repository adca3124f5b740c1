/*
 * fem_oracle.c -- CPU restatement of the reference FEM hot path.
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library, and the timed
 * CPU baseline ("port").  See fem_oracle.h for what each routine restates.
 * Built by oracle/Makefile with the reference's release flags (-O3 -DNDEBUG,
 * no -march, no OpenMP; reference CMakeLists.txt).
 */
#include "fem_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define REAL double
#define SFX _d
#define SQRT sqrt
#define FABS fabs
#define REALMIN DBL_MIN
#include "fem_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS
#undef REALMIN

#define REAL float
#define SFX _f
#define SQRT sqrtf
#define FABS fabsf
#define REALMIN FLT_MIN
#include "fem_oracle_impl.h"
#undef REAL
#undef SFX
#undef SQRT
#undef FABS
#undef REALMIN

/* FEMContext::update_boundary_conditions (FEMContext.cpp:224-240). */
int32_t orc_idx_map(int32_t nv, const uint8_t *onb, int32_t bc, int32_t *idx_map) {
    int32_t idx = 0;
    for (int32_t i = 0; i < nv; i++) {
        idx_map[i] = -1;
        if (bc == ORC_NEUMANN || !onb[i]) idx_map[i] = idx++;
    }
    return idx;
}

static int cmp_i32(const void *a, const void *b) {
    int32_t x = *(const int32_t *)a, y = *(const int32_t *)b;
    return (x > y) - (x < y);
}

/* Pattern of setFromTriplets over the stiffness triplets (FEMContext.cpp:97-112):
 * per row the sorted set of columns, explicit zeros kept. */
int64_t orc_pattern(int32_t nv, int32_t ntri, const uint32_t *tri, const uint8_t *onb, int32_t bc,
                    const int32_t *idx_map, int32_t n, int32_t *rp, int32_t *ci, int64_t cap) {
    (void)nv;
    int64_t *cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    for (int32_t k = 0; k < ntri; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                uint32_t a = tri[3 * (size_t)k + i], b = tri[3 * (size_t)k + j];
                if (bc == ORC_NEUMANN || (!onb[a] && !onb[b])) cnt[idx_map[a] + 1]++;
            }
    for (int32_t r = 0; r < n; r++) cnt[r + 1] += cnt[r];
    int64_t total = cnt[n];
    int32_t *buf = (int32_t *)malloc(sizeof(int32_t) * (size_t)(total + 1));
    int64_t *pos = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    memcpy(pos, cnt, sizeof(int64_t) * ((size_t)n + 1));
    for (int32_t k = 0; k < ntri; k++)
        for (int i = 0; i < 3; i++)
            for (int j = 0; j < 3; j++) {
                uint32_t a = tri[3 * (size_t)k + i], b = tri[3 * (size_t)k + j];
                if (bc == ORC_NEUMANN || (!onb[a] && !onb[b]))
                    buf[pos[idx_map[a]]++] = idx_map[b];
            }
    int64_t nnz = 0;
    rp[0] = 0;
    for (int32_t r = 0; r < n; r++) {
        int64_t lo = cnt[r], hi = cnt[r + 1];
        qsort(buf + lo, (size_t)(hi - lo), sizeof(int32_t), cmp_i32);
        for (int64_t s = lo; s < hi; s++) {
            if (s > lo && buf[s] == buf[s - 1]) continue;
            if (nnz >= cap) { nnz = -1; goto out; }
            ci[nnz++] = buf[s];
        }
        rp[r + 1] = (int32_t)nnz;
    }
out:
    free(cnt);
    free(buf);
    free(pos);
    return nnz;
}

void orc_spmv_d(int32_t n, const int32_t *rp, const int32_t *ci, const double *K, const double *M,
                const double *A, double aK, double aM, double aA, const double *x, double *y) {
    for (int32_t r = 0; r < n; r++) {
        double acc = 0;
        for (int32_t s = rp[r]; s < rp[r + 1]; s++) {
            double val = 0;
            if (aM != 0) val += aM * M[s];
            if (aK != 0) val += aK * K[s];
            if (aA != 0 && A) val += aA * A[s];
            acc += val * x[ci[s]];
        }
        y[r] = acc;
    }
}

int32_t orc_bicgstab_d(int32_t n, const int32_t *rp, const int32_t *ci, const double *S,
                       const double *b, double *x, double tol, int32_t maxit, double *err) {
    return bicgstab_d(n, rp, ci, S, b, x, tol, maxit, err);
}

int32_t orc_newmark_init_d(int32_t n, const int32_t *rp, const int32_t *ci, const double *K,
                           const double *M, double c, const double *u0, double *a0, double tol) {
    double *b = (double *)malloc(sizeof(double) * (size_t)n + 64);
    csr_mv_d(n, rp, ci, K, u0, b);
    for (int32_t i = 0; i < n; i++) { b[i] = -c * c * b[i]; a0[i] = 0; }
    double err;
    int32_t it = orc_cg_d(n, rp, ci, M, b, a0, tol, 4 * n, &err);
    free(b);
    return it;
}

/* Position of (c, r) for every CSR entry (r, c) of a structurally symmetric pattern. */
static void transpose_perm(int32_t n, const int32_t *rp, const int32_t *ci, int64_t *perm) {
    for (int32_t r = 0; r < n; r++)
        for (int32_t s = rp[r]; s < rp[r + 1]; s++) {
            int32_t c = ci[s];
            perm[s] = find_col_d(ci, rp[c], rp[c + 1], r);
        }
}

double orc_adjoint_d(int32_t nv, int32_t ntri, const float *xyz, const uint32_t *tri,
                     const int32_t *idx_map, int32_t n, const int32_t *rp, const int32_t *ci,
                     const double *Kk, const double *M, const double *A, double dt, int32_t nsteps,
                     const double *u0, const double *target, double tol, double *uT,
                     double *grad_u0, double *grad_kappa) {
    (void)nv;
    int64_t nnz = rp[n];
    double *S = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
    double *ST = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
    double *B = (double *)malloc(sizeof(double) * (size_t)(nnz + 1));
    int64_t *perm = (int64_t *)malloc(sizeof(int64_t) * (size_t)(nnz + 1));
    double *U = (double *)malloc(sizeof(double) * (size_t)n * (size_t)(nsteps + 1) + 64);
    double *rhs = (double *)malloc(sizeof(double) * (size_t)n * 3 + 64);
    double *lam = rhs + n, *mu = lam + n;
    double err;
    for (int64_t s = 0; s < nnz; s++) { B[s] = M[s] / dt; S[s] = B[s] + Kk[s] - A[s]; }
    transpose_perm(n, rp, ci, perm);
    for (int64_t s = 0; s < nnz; s++) ST[s] = S[perm[s]];
    memcpy(U, u0, sizeof(double) * (size_t)n);
    for (int32_t t = 0; t < nsteps; t++) {
        double *un = U + (size_t)n * t, *un1 = U + (size_t)n * (t + 1);
        csr_mv_d(n, rp, ci, B, un, rhs);
        memcpy(un1, un, sizeof(double) * (size_t)n);
        bicgstab_d(n, rp, ci, S, rhs, un1, tol, 20 * n, &err);
    }
    double *ul = U + (size_t)n * nsteps;
    double loss = 0;
    for (int32_t i = 0; i < n; i++) {
        double d = ul[i] - target[i];
        loss += d * d;
        lam[i] = 2 * d;
        uT[i] = ul[i];
    }
    for (int32_t e = 0; e < ntri; e++) grad_kappa[e] = 0;
    for (int32_t t = nsteps - 1; t >= 0; t--) {
        /* mu = S^{-T} lam_{t+1} ; lam_t = B mu ; dL/dkappa_e -= mu^T K_e u_{t+1} */
        for (int32_t i = 0; i < n; i++) mu[i] = 0;
        bicgstab_d(n, rp, ci, ST, lam, mu, tol, 20 * n, &err);
        const double *un1 = U + (size_t)n * (t + 1);
        for (int32_t e = 0; e < ntri; e++) {
            const uint32_t *tt = tri + 3 * (size_t)e;
            double g[3][3], area, jdet, nrm[3];
            elem_geom_d(xyz, tt, g, &area, &jdet, nrm);
            double acc = 0;
            for (int a = 0; a < 3; a++) {
                int32_t ra = idx_map[tt[a]];
                if (ra < 0) continue;
                for (int b = 0; b < 3; b++) {
                    int32_t rb = idx_map[tt[b]];
                    if (rb < 0) continue;
                    double kab = area * (g[a][0] * g[b][0] + g[a][1] * g[b][1] + g[a][2] * g[b][2]);
                    acc += mu[ra] * kab * un1[rb];
                }
            }
            grad_kappa[e] -= acc;
        }
        csr_mv_d(n, rp, ci, B, mu, lam);
    }
    memcpy(grad_u0, lam, sizeof(double) * (size_t)n);
    free(S); free(ST); free(B); free(perm); free(U); free(rhs);
    return loss;
}

/* ---------------- reference-cost port (FP32) ---------------- */

typedef struct { int32_t n; int32_t *p; int32_t *i; float *v; } csc_f;

/* C = a*X + b*Y (+ c*Z) by per-column union merge into freshly allocated
 * storage, as Eigen assigns a sparse sum expression. */
static csc_f port_sum(int32_t n, float a, const int32_t *xp, const int32_t *xi, const float *xv,
                      float b, const int32_t *yp, const int32_t *yi, const float *yv, float c,
                      const int32_t *zp, const int32_t *zi, const float *zv) {
    csc_f C;
    C.n = n;
    int64_t cap = (int64_t)xp[n] + yp[n] + (zp ? zp[n] : 0);
    C.p = (int32_t *)malloc(sizeof(int32_t) * ((size_t)n + 1));
    C.i = (int32_t *)malloc(sizeof(int32_t) * (size_t)(cap + 1));
    C.v = (float *)malloc(sizeof(float) * (size_t)(cap + 1));
    int32_t w = 0;
    C.p[0] = 0;
    for (int32_t col = 0; col < n; col++) {
        int32_t sx = xp[col], ex = xp[col + 1], sy = yp[col], ey = yp[col + 1];
        int32_t sz = zp ? zp[col] : 0, ez = zp ? zp[col + 1] : 0;
        while (sx < ex || sy < ey || sz < ez) {
            int32_t r = INT32_MAX;
            if (sx < ex && xi[sx] < r) r = xi[sx];
            if (sy < ey && yi[sy] < r) r = yi[sy];
            if (sz < ez && zi[sz] < r) r = zi[sz];
            float val = 0;
            if (sx < ex && xi[sx] == r) val += a * xv[sx++];
            if (sy < ey && yi[sy] == r) val += b * yv[sy++];
            if (sz < ez && zi[sz] == r) val += c * zv[sz++];
            C.i[w] = r;
            C.v[w] = val;
            w++;
        }
        C.p[col + 1] = w;
    }
    return C;
}

static void port_free(csc_f *C) { free(C->p); free(C->i); free(C->v); }

/* res = scale * X * rhs, column scatter (Eigen ColMajor sparse * dense). */
static void port_mv(int32_t n, const int32_t *p, const int32_t *ii, const float *v, float scale,
                    const float *rhs, float *res) {
    for (int32_t r = 0; r < n; r++) res[r] = 0;
    for (int32_t j = 0; j < n; j++) {
        float rj = rhs[j];
        for (int32_t s = p[j]; s < p[j + 1]; s++) res[ii[s]] += (scale * v[s]) * rj;
    }
}

static float port_dot(int32_t n, const float *a, const float *b) {
    float acc[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    int32_t i = 0;
    for (; i + 8 <= n; i += 8)
        for (int q = 0; q < 8; q++) acc[q] += a[i + q] * b[i + q];
    float s = ((acc[0] + acc[4]) + (acc[1] + acc[5])) + ((acc[2] + acc[6]) + (acc[3] + acc[7]));
    for (; i < n; i++) s += a[i] * b[i];
    return s;
}

/* ConjugateGradient<SparseMatrix<float>, Lower|Upper>::compute + solve. */
static int32_t port_cg(const csc_f *A, const float *b, float *x) {
    int32_t n = A->n;
    float *invd = (float *)malloc(sizeof(float) * (size_t)n + 64);
    for (int32_t j = 0; j < n; j++) {
        int32_t s = A->p[j];
        while (s < A->p[j + 1] && A->i[s] != j) s++;
        invd[j] = (s < A->p[j + 1] && A->v[s] != 0) ? 1.0f / A->v[s] : 1.0f;
    }
    float *r = (float *)malloc(sizeof(float) * (size_t)n * 4 + 64);
    float *p = r + n, *z = p + n, *tmp = z + n;
    int32_t maxit = 2 * n, it = 0;
    float tol = FLT_EPSILON;
    for (int32_t i = 0; i < n; i++) { x[i] = 0; r[i] = b[i]; }
    float rhs2 = port_dot(n, b, b);
    if (rhs2 == 0) goto done;
    {
        float thr = tol * tol * rhs2;
        if (thr < FLT_MIN) thr = FLT_MIN;
        float res2 = port_dot(n, r, r);
        if (res2 < thr) goto done;
        for (int32_t i = 0; i < n; i++) p[i] = invd[i] * r[i];
        float absnew = port_dot(n, r, p);
        while (it < maxit) {
            port_mv(n, A->p, A->i, A->v, 1.0f, p, tmp);
            float alpha = absnew / port_dot(n, p, tmp);
            for (int32_t i = 0; i < n; i++) { x[i] += alpha * p[i]; r[i] -= alpha * tmp[i]; }
            res2 = port_dot(n, r, r);
            if (res2 < thr) break;
            for (int32_t i = 0; i < n; i++) z[i] = invd[i] * r[i];
            float absold = absnew;
            absnew = port_dot(n, r, z);
            float beta = absnew / absold;
            for (int32_t i = 0; i < n; i++) p[i] = z[i] + beta * p[i];
            it++;
        }
    }
done:
    free(invd);
    free(r);
    return it;
}

int64_t orc_port_run_f(int32_t n, const int32_t *kp, const int32_t *ki, const float *kv,
                       const int32_t *mp, const int32_t *mi, const float *mv, const int32_t *ap,
                       const int32_t *ai, const float *av, int32_t eq, const double *prm, float *u,
                       float *v, int32_t nsteps) {
    int64_t total = 0;
    float dt = (float)prm[0];
    float *b = (float *)malloc(sizeof(float) * (size_t)n * 4 + 64);
    float *b2 = b + n, *t1 = b2 + n, *x = t1 + n;
    float *ml = NULL;
    if (eq == ORC_GRAYSCOTT && prm[5] != 0) {
        /* lumped-mass reaction weighting (documented deviation): row sums of M */
        ml = (float *)calloc((size_t)n + 16, sizeof(float));
        for (int32_t j = 0; j < n; j++)
            for (int32_t s = mp[j]; s < mp[j + 1]; s++) ml[mi[s]] += mv[s];
    }
    for (int32_t step = 0; step < nsteps; step++) {
        switch (eq) {
        case ORC_HEAT:
        case ORC_ADVDIFF: {
            float c = (float)prm[1];
            csc_f A = port_sum(n, 1.0f / dt, mp, mi, mv, c, kp, ki, kv, -1.0f,
                               eq == ORC_ADVDIFF ? ap : NULL, ai, av);
            port_mv(n, mp, mi, mv, 1.0f / dt, u, b);
            total += port_cg(&A, b, x);
            memcpy(u, x, sizeof(float) * (size_t)n);
            port_free(&A);
        } break;
        case ORC_WAVE: {
            float c = (float)prm[1];
            csc_f A = port_sum(n, 1.0f / dt, mp, mi, mv, c * c * dt, kp, ki, kv, 0, NULL, NULL, NULL);
            port_mv(n, mp, mi, mv, 1.0f / dt, v, b);
            port_mv(n, kp, ki, kv, c * c, u, t1);
            for (int32_t i = 0; i < n; i++) b[i] -= t1[i];
            total += port_cg(&A, b, x);
            for (int32_t i = 0; i < n; i++) { v[i] = x[i]; u[i] = u[i] + v[i] * dt; }
            port_free(&A);
        } break;
        case ORC_GRAYSCOTT: {
            float Du = (float)prm[1], Dv = (float)prm[2], F = (float)prm[3], kk = (float)prm[4];
            csc_f Au = port_sum(n, 1.0f / dt, mp, mi, mv, Du, kp, ki, kv, 0, NULL, NULL, NULL);
            port_mv(n, mp, mi, mv, 1.0f / dt, u, b);
            csc_f Av = port_sum(n, 1.0f / dt, mp, mi, mv, Dv, kp, ki, kv, 0, NULL, NULL, NULL);
            port_mv(n, mp, mi, mv, 1.0f / dt, v, b2);
            for (int32_t i = 0; i < n; i++) {
                float w = ml ? ml[i] : 1.0f;
                float uvv = u[i] * (v[i] * v[i]);
                b[i] += w * (-uvv + F * (1.0f - u[i]));
                b2[i] += w * (uvv - (F + kk) * v[i]);
            }
            total += port_cg(&Au, b, x);
            memcpy(u, x, sizeof(float) * (size_t)n);
            total += port_cg(&Av, b2, x);
            memcpy(v, x, sizeof(float) * (size_t)n);
            port_free(&Au);
            port_free(&Av);
        } break;
        default:
            break;
        }
    }
    free(b);
    free(ml);
    return total;
}

int32_t orc_unstable_f(int32_t n, const float *u, const float *v) {
    for (int32_t i = 0; i < n; i++)
        if (u[i] > 1e4f || v[i] > 1e4f) return 1;
    return 0;
}
