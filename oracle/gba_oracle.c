/*
 * gba_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * CPU restatement of the reference LM bundle-adjustment solve
 * (src/lm.cu, include/gba/schur.cuh, src/bal_reader.cpp).  It is the checker
 * for the HIP product path and, timed on host cores, the "port" CPU baseline.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Parity status of each function (what pins it) is listed in the tests:
 *   residuals      -> golden vectors from the reference's own bal_reader.cpp
 *                     compiled under oracle/_ref, closed-form cases
 *   Jacobian       -> reference FD cases (tests/test_vs_ceres.cpp:34-39) and
 *                     complex-step derivatives of an independent residual
 *   linearisation  -> dense J^T J / -J^T r built by brute force
 *   batched blocks -> numpy/LAPACK on the reference's test protocol
 *   Schur step     -> full damped normal equations solved densely
 *   LM control     -> parity unpinned against the reference binary (CUDA +
 *                     Eigen, not buildable here); pinned by invariants only.
 */
#define _GNU_SOURCE
#include "gba_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* small helpers                                                        */
/* ------------------------------------------------------------------ */

static double now_s(void) {
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* damping diagonal: x if x > 0 else 1e-12 (src/lm.cu:294-296, :590, :612, :635) */
static inline double damp_diag(double x) { return x > 0.0 ? x : 1e-12; }

typedef void (*task_fn)(void* ctx, int t, int nt);
typedef struct {
    task_fn fn;
    void* ctx;
    int t, nt;
} task_arg;

static void* task_tramp(void* p) {
    task_arg* a = (task_arg*)p;
    a->fn(a->ctx, a->t, a->nt);
    return NULL;
}

/* Runs fn(ctx, t, nt) for t in [0, nt) on nt threads (inline when nt == 1). */
static void parallel_run(int nt, task_fn fn, void* ctx) {
    if (nt <= 1) {
        fn(ctx, 0, 1);
        return;
    }
    pthread_t* th = (pthread_t*)malloc(sizeof(pthread_t) * (size_t)nt);
    task_arg* args = (task_arg*)malloc(sizeof(task_arg) * (size_t)nt);
    for (int t = 0; t < nt; ++t) {
        args[t].fn = fn;
        args[t].ctx = ctx;
        args[t].t = t;
        args[t].nt = nt;
        if (t > 0) pthread_create(&th[t], NULL, task_tramp, &args[t]);
    }
    fn(ctx, 0, nt);
    for (int t = 1; t < nt; ++t) pthread_join(th[t], NULL);
    free(th);
    free(args);
}

static void range_of(long n, int t, int nt, long* b, long* e) {
    long chunk = (n + nt - 1) / nt;
    *b = (long)t * chunk;
    *e = *b + chunk;
    if (*b > n) *b = n;
    if (*e > n) *e = n;
}

/* CSR of edges grouped by key (ascending edge index inside each group). */
static void build_csr(int nkeys, int O, const int* key, int** off_out, int** idx_out) {
    int* off = (int*)calloc((size_t)nkeys + 1, sizeof(int));
    int* idx = (int*)malloc(sizeof(int) * (size_t)(O > 0 ? O : 1));
    for (int e = 0; e < O; ++e) off[key[e] + 1]++;
    for (int k = 0; k < nkeys; ++k) off[k + 1] += off[k];
    int* fill = (int*)malloc(sizeof(int) * (size_t)(nkeys > 0 ? nkeys : 1));
    for (int k = 0; k < nkeys; ++k) fill[k] = off[k];
    for (int e = 0; e < O; ++e) idx[fill[key[e]]++] = e;
    free(fill);
    *off_out = off;
    *idx_out = idx;
}

/* ------------------------------------------------------------------ */
/* camera model                                                         */
/* ------------------------------------------------------------------ */

/* src/bal_reader.cpp:100-128 */
void orc_snavely_residual(const double* cam, const double* X, const double* obs, double* r) {
    const double* a = cam;
    const double* t = cam + 3;
    double theta = sqrt(a[0] * a[0] + a[1] * a[1] + a[2] * a[2]);
    double P[3];
    if (theta > 1e-12) {
        double w0 = a[0] / theta, w1 = a[1] / theta, w2 = a[2] / theta;
        double c = cos(theta), s = sin(theta);
        double dot = w0 * X[0] + w1 * X[1] + w2 * X[2];
        double cr0 = w1 * X[2] - w2 * X[1];
        double cr1 = w2 * X[0] - w0 * X[2];
        double cr2 = w0 * X[1] - w1 * X[0];
        P[0] = X[0] * c + cr0 * s + w0 * dot * (1.0 - c) + t[0];
        P[1] = X[1] * c + cr1 * s + w1 * dot * (1.0 - c) + t[1];
        P[2] = X[2] * c + cr2 * s + w2 * dot * (1.0 - c) + t[2];
    } else {
        P[0] = X[0] + t[0];
        P[1] = X[1] + t[1];
        P[2] = X[2] + t[2];
    }
    double px = -P[0] / P[2], py = -P[1] / P[2];
    double r2 = px * px + py * py;
    double dist = 1.0 + cam[7] * r2 + cam[8] * r2 * r2;
    r[0] = cam[6] * dist * px - obs[0];
    r[1] = cam[6] * dist * py - obs[1];
}

/* src/bal_reader.cpp:130-139 */
double orc_total_squared_error(int O, const double* cams, const double* pts, const int* oc,
                               const int* op, const double* px) {
    double sum = 0.0, r[2];
    for (int e = 0; e < O; ++e) {
        orc_snavely_residual(cams + 9 * (size_t)oc[e], pts + 3 * (size_t)op[e], px + 2 * (size_t)e, r);
        sum += r[0] * r[0] + r[1] * r[1];
    }
    return sum;
}

/* Rotated point R(a) X: Rodrigues when theta^2 > 1e-24, else X + a x X
 * (src/lm.cu:62-76). */
static void rotate_point(const double* a, const double* X, double* RX) {
    double th2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    if (th2 > 1e-24) {
        double th = sqrt(th2);
        double w0 = a[0] / th, w1 = a[1] / th, w2 = a[2] / th;
        double c = cos(th), s = sin(th);
        double cr0 = w1 * X[2] - w2 * X[1];
        double cr1 = w2 * X[0] - w0 * X[2];
        double cr2 = w0 * X[1] - w1 * X[0];
        double dot = w0 * X[0] + w1 * X[1] + w2 * X[2];
        RX[0] = X[0] * c + cr0 * s + w0 * dot * (1.0 - c);
        RX[1] = X[1] * c + cr1 * s + w1 * dot * (1.0 - c);
        RX[2] = X[2] * c + cr2 * s + w2 * dot * (1.0 - c);
    } else {
        RX[0] = X[0] + (a[1] * X[2] - a[2] * X[1]);
        RX[1] = X[1] + (a[2] * X[0] - a[0] * X[2]);
        RX[2] = X[2] + (a[0] * X[1] - a[1] * X[0]);
    }
}

/* src/lm.cu:58-82 */
void orc_edge_residual(const double* cam, const double* X, const double* obs, double* r) {
    double RX[3];
    rotate_point(cam, X, RX);
    double P0 = RX[0] + cam[3], P1 = RX[1] + cam[4], P2 = RX[2] + cam[5];
    double u = -P0 / P2, v = -P1 / P2;
    double r2 = u * u + v * v;
    double dist = 1.0 + cam[7] * r2 + cam[8] * r2 * r2;
    r[0] = cam[6] * dist * u - obs[0];
    r[1] = cam[6] * dist * v - obs[1];
}

/* src/lm.cu:93-191.  Main branch: R = cI + s[w]x + (1-c)ww^T,
 * dRX/dtheta = -sX + c(w x X) + s(w.X)w, A = dRX/dw = -s[X]x + (1-c)((w.X)I + wX^T),
 * dRX/da = (A - (Aw)w^T)/theta + (dRX/dtheta) w^T.  Small branch (theta^2 <=
 * 1e-24): RX = X + a x X, R = I, dRX/da = -[X]x. */
void orc_edge_residual_jacobian(const double* cam, const double* X, const double* obs, double* r,
                                double* Jc, double* Jp) {
    orc_edge_residual(cam, X, obs, r);
    const double* a = cam;
    double th2 = a[0] * a[0] + a[1] * a[1] + a[2] * a[2];
    double R[9], Dra[9], RX[3];
    if (th2 > 1e-24) {
        double th = sqrt(th2);
        double w[3] = {a[0] / th, a[1] / th, a[2] / th};
        double c = cos(th), s = sin(th);
        double wxX[3] = {w[1] * X[2] - w[2] * X[1], w[2] * X[0] - w[0] * X[2],
                         w[0] * X[1] - w[1] * X[0]};
        double q = w[0] * X[0] + w[1] * X[1] + w[2] * X[2];
        for (int i = 0; i < 3; ++i) RX[i] = X[i] * c + wxX[i] * s + w[i] * q * (1.0 - c);
        /* skew(w) and skew(X) */
        double Kw[9] = {0.0, -w[2], w[1], w[2], 0.0, -w[0], -w[1], w[0], 0.0};
        double KX[9] = {0.0, -X[2], X[1], X[2], 0.0, -X[0], -X[1], X[0], 0.0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                R[3 * i + j] = (i == j ? c : 0.0) + s * Kw[3 * i + j] + (1.0 - c) * w[i] * w[j];
        double dth[3];
        for (int i = 0; i < 3; ++i) dth[i] = -s * X[i] + c * wxX[i] + s * q * w[i];
        double A[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                A[3 * i + j] = -s * KX[3 * i + j] + (1.0 - c) * ((i == j ? q : 0.0) + w[i] * X[j]);
        double Aw[3];
        for (int i = 0; i < 3; ++i) Aw[i] = A[3 * i] * w[0] + A[3 * i + 1] * w[1] + A[3 * i + 2] * w[2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                Dra[3 * i + j] = (A[3 * i + j] - Aw[i] * w[j]) / th + dth[i] * w[j];
    } else {
        RX[0] = X[0] + (a[1] * X[2] - a[2] * X[1]);
        RX[1] = X[1] + (a[2] * X[0] - a[0] * X[2]);
        RX[2] = X[2] + (a[0] * X[1] - a[1] * X[0]);
        for (int i = 0; i < 9; ++i) R[i] = (i % 4 == 0) ? 1.0 : 0.0;
        double nKX[9] = {0.0, X[2], -X[1], -X[2], 0.0, X[0], X[1], -X[0], 0.0};
        memcpy(Dra, nKX, sizeof nKX);
    }
    double P[3] = {RX[0] + cam[3], RX[1] + cam[4], RX[2] + cam[5]};
    double u = -P[0] / P[2], v = -P[1] / P[2];
    double r2 = u * u + v * v;
    double f = cam[6], k1 = cam[7], k2 = cam[8];
    double dist = 1.0 + k1 * r2 + k2 * r2 * r2;
    double L = k1 + 2.0 * k2 * r2;
    double D00 = f * (dist + 2.0 * L * u * u);
    double D01 = 2.0 * f * L * u * v;
    double D11 = f * (dist + 2.0 * L * v * v);
    double iz = 1.0 / P[2];
    double M[6] = {-iz, 0.0, P[0] * iz * iz, 0.0, -iz, P[1] * iz * iz};
    double DM[6];
    for (int j = 0; j < 3; ++j) {
        DM[j] = D00 * M[j] + D01 * M[3 + j];
        DM[3 + j] = D01 * M[j] + D11 * M[3 + j];
    }
    for (int row = 0; row < 2; ++row) {
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += DM[3 * row + k] * Dra[3 * k + j];
            Jc[9 * row + j] = acc;
            Jc[9 * row + 3 + j] = DM[3 * row + j];
        }
        double pr = row == 0 ? u : v;
        Jc[9 * row + 6] = dist * pr;
        Jc[9 * row + 7] = f * r2 * pr;
        Jc[9 * row + 8] = f * r2 * r2 * pr;
        for (int j = 0; j < 3; ++j) {
            double acc = 0.0;
            for (int k = 0; k < 3; ++k) acc += DM[3 * row + k] * R[3 * k + j];
            Jp[3 * row + j] = acc;
        }
    }
}

/* ------------------------------------------------------------------ */
/* linearisation (src/lm.cu:205-252) and cost (:255-281)               */
/* ------------------------------------------------------------------ */

/* Huber loss on s = |r|^2, as the reference's Ceres harness applies
 * HuberLoss(delta) (benchmarks/bench_ceres.cpp:120-131): rho(s) = s for
 * s <= delta^2, else 2 delta sqrt(s) - delta^2; delta <= 0 means none.  The
 * robustified Gauss-Newton model (Triggs, McLauchlan, Hartley, Fitzgibbon,
 * "Bundle Adjustment - A Modern Synthesis", 2000, sec. 4.3) scales r and J by
 * sqrt(rho'(s)) plus a rank-one curvature term that vanishes when
 * rho''(s) <= 0; for Huber rho'' is 0 inside and negative outside, so the
 * model is the plain IRLS scaling. */
static double huber_rho(double s, double delta) {
    return (delta > 0.0 && s > delta * delta) ? 2.0 * delta * sqrt(s) - delta * delta : s;
}
static double huber_sqrt_drho(double s, double delta) {
    return (delta > 0.0 && s > delta * delta) ? sqrt(delta / sqrt(s)) : 1.0;
}

typedef struct {
    int C, P, O;
    const double *cams, *pts, *px;
    const int *oc, *op;
    double* J;         /* 26 per edge: r(2) Jc(18) Jp(6), loss-weighted */
    double* rho;       /* rho(|r|^2) per edge */
    unsigned char* ok; /* finite flag per edge */
    double *U, *V, *W, *bc, *bp;
    const int *coff, *cidx, *poff, *pidx;
    double huber;
} lin_ctx;

static void lin_edges(void* vctx, int t, int nt) {
    lin_ctx* c = (lin_ctx*)vctx;
    long b, e;
    range_of(c->O, t, nt, &b, &e);
    for (long k = b; k < e; ++k) {
        double* Jk = c->J + 26 * k;
        double *r = Jk, *Jc = Jk + 2, *Jp = Jk + 20;
        orc_edge_residual_jacobian(c->cams + 9 * (size_t)c->oc[k], c->pts + 3 * (size_t)c->op[k],
                                   c->px + 2 * k, r, Jc, Jp);
        int fin = isfinite(r[0]) && isfinite(r[1]);
        for (int j = 0; j < 18; ++j) fin = fin && isfinite(Jc[j]);
        for (int j = 0; j < 6; ++j) fin = fin && isfinite(Jp[j]);
        c->ok[k] = (unsigned char)fin;
        if (fin) {
            const double sq = r[0] * r[0] + r[1] * r[1];
            const double w = huber_sqrt_drho(sq, c->huber);
            c->rho[k] = huber_rho(sq, c->huber);
            for (int j = 0; j < 26; ++j) Jk[j] *= w;
        }
        double* We = c->W + 27 * k;
        if (!fin) { /* the reference leaves W of a skipped edge stale (lm.cu:222-225);
                       here it is zero, so the edge contributes nothing (lm.cu:19-21) */
            memset(We, 0, 27 * sizeof(double));
            continue;
        }
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 3; ++j) We[3 * i + j] = Jc[i] * Jp[j] + Jc[9 + i] * Jp[3 + j];
    }
}

static void lin_cams(void* vctx, int t, int nt) {
    lin_ctx* c = (lin_ctx*)vctx;
    long b, e;
    range_of(c->C, t, nt, &b, &e);
    for (long cam = b; cam < e; ++cam) {
        double* U = c->U + 81 * cam;
        double* bc = c->bc + 9 * cam;
        for (int k = c->coff[cam]; k < c->coff[cam + 1]; ++k) {
            int ed = c->cidx[k];
            if (!c->ok[ed]) continue;
            const double* Jk = c->J + 26 * (size_t)ed;
            const double *r = Jk, *Jc = Jk + 2;
            for (int i = 0; i < 9; ++i) {
                for (int j = 0; j <= i; ++j) U[9 * i + j] += Jc[i] * Jc[j] + Jc[9 + i] * Jc[9 + j];
                bc[i] += -(Jc[i] * r[0] + Jc[9 + i] * r[1]);
            }
        }
    }
}

static void lin_pts(void* vctx, int t, int nt) {
    lin_ctx* c = (lin_ctx*)vctx;
    long b, e;
    range_of(c->P, t, nt, &b, &e);
    for (long p = b; p < e; ++p) {
        double* V = c->V + 9 * p;
        double* bp = c->bp + 3 * p;
        for (int k = c->poff[p]; k < c->poff[p + 1]; ++k) {
            int ed = c->pidx[k];
            if (!c->ok[ed]) continue;
            const double* Jk = c->J + 26 * (size_t)ed;
            const double *r = Jk, *Jp = Jk + 20;
            for (int i = 0; i < 3; ++i) {
                for (int j = 0; j <= i; ++j) V[3 * i + j] += Jp[i] * Jp[j] + Jp[3 + i] * Jp[3 + j];
                bp[i] += -(Jp[i] * r[0] + Jp[3 + i] * r[1]);
            }
        }
    }
}

void orc_linearize(int C, int P, int O, const double* cams, const double* pts, const int* oc,
                   const int* op, const double* px, double* U, double* V, double* W, double* bc,
                   double* bp, double* sse, unsigned long long* anomalies, int nthreads) {
    orc_linearize_loss(C, P, O, cams, pts, oc, op, px, 0.0, U, V, W, bc, bp, sse, anomalies, nthreads);
}

void orc_linearize_loss(int C, int P, int O, const double* cams, const double* pts, const int* oc,
                        const int* op, const double* px, double huber_delta, double* U, double* V,
                        double* W, double* bc, double* bp, double* sse, unsigned long long* anomalies,
                        int nthreads) {
    if (nthreads < 1) nthreads = 1;
    lin_ctx c;
    c.huber = huber_delta;
    c.rho = (double*)malloc(sizeof(double) * (size_t)(O > 0 ? O : 1));
    c.C = C;
    c.P = P;
    c.O = O;
    c.cams = cams;
    c.pts = pts;
    c.px = px;
    c.oc = oc;
    c.op = op;
    c.J = (double*)malloc(sizeof(double) * 26 * (size_t)(O > 0 ? O : 1));
    c.ok = (unsigned char*)malloc((size_t)(O > 0 ? O : 1));
    c.U = U;
    c.V = V;
    c.W = W;
    c.bc = bc;
    c.bp = bp;
    int *coff, *cidx, *poff, *pidx;
    build_csr(C, O, oc, &coff, &cidx);
    build_csr(P, O, op, &poff, &pidx);
    c.coff = coff;
    c.cidx = cidx;
    c.poff = poff;
    c.pidx = pidx;
    memset(U, 0, sizeof(double) * 81 * (size_t)C);
    memset(V, 0, sizeof(double) * 9 * (size_t)P);
    memset(bc, 0, sizeof(double) * 9 * (size_t)C);
    memset(bp, 0, sizeof(double) * 3 * (size_t)P);
    parallel_run(nthreads, lin_edges, &c);
    parallel_run(nthreads < C ? nthreads : (C > 0 ? C : 1), lin_cams, &c);
    parallel_run(nthreads, lin_pts, &c);
    double s = 0.0;
    unsigned long long an = 0;
    for (long k = 0; k < O; ++k) {
        if (!c.ok[k]) {
            ++an;
            continue;
        }
        s += c.rho[k];
    }
    *sse = s;
    *anomalies = an;
    free(c.rho);
    free(c.J);
    free(c.ok);
    free(coff);
    free(cidx);
    free(poff);
    free(pidx);
}

typedef struct {
    int O;
    const double *cams, *pts, *px;
    const int *oc, *op;
    double* r2; /* per-edge rho(squared residual), NaN marks anomaly */
    double huber;
} cost_ctx;

static void cost_edges(void* vctx, int t, int nt) {
    cost_ctx* c = (cost_ctx*)vctx;
    long b, e;
    range_of(c->O, t, nt, &b, &e);
    for (long k = b; k < e; ++k) {
        double r[2];
        orc_edge_residual(c->cams + 9 * (size_t)c->oc[k], c->pts + 3 * (size_t)c->op[k], c->px + 2 * k, r);
        c->r2[k] = (isfinite(r[0]) && isfinite(r[1])) ? huber_rho(r[0] * r[0] + r[1] * r[1], c->huber) : NAN;
    }
}

double orc_cost(int O, const double* cams, const double* pts, const int* oc, const int* op,
                const double* px, unsigned long long* anomalies, int nthreads) {
    return orc_cost_loss(O, cams, pts, oc, op, px, 0.0, anomalies, nthreads);
}

double orc_cost_loss(int O, const double* cams, const double* pts, const int* oc, const int* op,
                     const double* px, double huber_delta, unsigned long long* anomalies, int nthreads) {
    cost_ctx c = {O, cams, pts, px, oc, op, (double*)malloc(sizeof(double) * (size_t)(O > 0 ? O : 1)),
                  huber_delta};
    parallel_run(nthreads < 1 ? 1 : nthreads, cost_edges, &c);
    double s = 0.0;
    unsigned long long an = 0;
    for (long k = 0; k < O; ++k) {
        if (isnan(c.r2[k])) {
            ++an;
            continue;
        }
        s += c.r2[k];
    }
    free(c.r2);
    if (anomalies) *anomalies = an;
    return s;
}

/* ------------------------------------------------------------------ */
/* batched blocks (include/gba/schur.cuh)                              */
/* ------------------------------------------------------------------ */

/* include/gba/schur.cuh:23-38 (Cholesky-Crout, pivot must be > 0) */
int orc_chol9_factor(const double* A, double* L) {
    for (int i = 0; i < 9; ++i) {
        for (int j = 0; j <= i; ++j) {
            double s = A[9 * i + j];
            for (int k = 0; k < j; ++k) s -= L[9 * i + k] * L[9 * j + k];
            if (i == j) {
                if (!(s > 0.0)) return 0;
                L[9 * i + i] = sqrt(s);
            } else {
                L[9 * i + j] = s / L[9 * j + j];
            }
        }
        for (int j = i + 1; j < 9; ++j) L[9 * i + j] = 0.0;
    }
    return 1;
}

/* include/gba/schur.cuh:41-53 */
void orc_chol9_solve(const double* L, const double* b, double* x) {
    double y[9];
    for (int i = 0; i < 9; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= L[9 * i + k] * y[k];
        y[i] = s / L[9 * i + i];
    }
    for (int i = 8; i >= 0; --i) {
        double s = y[i];
        for (int k = i + 1; k < 9; ++k) s -= L[9 * k + i] * x[k];
        x[i] = s / L[9 * i + i];
    }
}

/* include/gba/schur.cuh:66-96 (A^-1 = L^-T L^-1, full 3x3 written) */
int orc_inverse3_spd(const double* A, double* Ainv) {
    double L[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j <= i; ++j) {
            double s = A[3 * i + j];
            for (int k = 0; k < j; ++k) s -= L[3 * i + k] * L[3 * j + k];
            if (i == j) {
                if (!(s > 0.0)) return 0;
                L[3 * i + i] = sqrt(s);
            } else {
                L[3 * i + j] = s / L[3 * j + j];
            }
        }
    double Li[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int i = 0; i < 3; ++i) {
        Li[3 * i + i] = 1.0 / L[3 * i + i];
        for (int j = 0; j < i; ++j) {
            double s = 0.0;
            for (int k = j; k < i; ++k) s -= L[3 * i + k] * Li[3 * k + j];
            Li[3 * i + j] = s / L[3 * i + i];
        }
    }
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += Li[3 * k + i] * Li[3 * k + j];
            Ainv[3 * i + j] = s;
        }
    return 1;
}

/* include/gba/schur.cuh:104-109 */
void orc_batch_chol9_factor_solve(const double* A, const double* b, double* x, int* ok, int n) {
    for (int i = 0; i < n; ++i) {
        double L[81];
        ok[i] = orc_chol9_factor(A + 81 * (size_t)i, L);
        if (ok[i]) orc_chol9_solve(L, b + 9 * (size_t)i, x + 9 * (size_t)i);
    }
}

/* include/gba/schur.cuh:112-116 */
void orc_batch_inverse3_spd(const double* A, double* Ainv, int* ok, int n) {
    for (int i = 0; i < n; ++i) ok[i] = orc_inverse3_spd(A + 9 * (size_t)i, Ainv + 9 * (size_t)i);
}

/* src/lm.cu:289-302 */
long orc_damp_invert_v(int P, const double* V, double lambda, double* Vinv) {
    long bad = 0;
    for (long p = 0; p < P; ++p) {
        const double* Vp = V + 9 * p;
        double A[9] = {Vp[0] + lambda * damp_diag(Vp[0]), Vp[3], Vp[6],
                       Vp[3], Vp[4] + lambda * damp_diag(Vp[4]), Vp[7],
                       Vp[6], Vp[7], Vp[8] + lambda * damp_diag(Vp[8])};
        if (!orc_inverse3_spd(A, Vinv + 9 * p)) ++bad;
    }
    return bad;
}

/* ------------------------------------------------------------------ */
/* dense Cholesky (stand-in for Eigen::LLT<MatrixXd, Lower>, lm.cu:598) */
/* single-threaded, blocked right-looking, column-major                 */
/* ------------------------------------------------------------------ */

#define CHOL_NB 64

/* unblocked lower Cholesky of the kb x kb block at A (leading dim n) */
static int chol_unblocked(int kb, double* A, int n) {
    for (int j = 0; j < kb; ++j) {
        double d = A[j + (size_t)j * n];
        for (int k = 0; k < j; ++k) d -= A[j + (size_t)k * n] * A[j + (size_t)k * n];
        if (!(d > 0.0)) return 0;
        d = sqrt(d);
        A[j + (size_t)j * n] = d;
        for (int i = j + 1; i < kb; ++i) {
            double s = A[i + (size_t)j * n];
            for (int k = 0; k < j; ++k) s -= A[i + (size_t)k * n] * A[j + (size_t)k * n];
            A[i + (size_t)j * n] = s / d;
        }
    }
    return 1;
}

int orc_dense_cholesky(int n, double* A) {
    double* pack = (double*)malloc(sizeof(double) * (size_t)CHOL_NB * (size_t)(n > 0 ? n : 1));
    int ok = 1;
    for (int k0 = 0; k0 < n && ok; k0 += CHOL_NB) {
        int kb = n - k0 < CHOL_NB ? n - k0 : CHOL_NB;
        double* Akk = A + k0 + (size_t)k0 * n;
        if (!chol_unblocked(kb, Akk, n)) {
            ok = 0;
            break;
        }
        int m = n - k0 - kb; /* rows below the panel */
        if (m <= 0) break;
        double* A21 = A + (k0 + kb) + (size_t)k0 * n;
        /* TRSM: A21 <- A21 L_kk^{-T}, column by column */
        for (int j = 0; j < kb; ++j) {
            double* cj = A21 + (size_t)j * n;
            for (int q = 0; q < j; ++q) {
                double l = Akk[j + (size_t)q * n];
                const double* cq = A21 + (size_t)q * n;
                for (int i = 0; i < m; ++i) cj[i] -= cq[i] * l;
            }
            double inv = 1.0 / Akk[j + (size_t)j * n];
            for (int i = 0; i < m; ++i) cj[i] *= inv;
        }
        /* pack the panel k-major: pack[k*m + i] = A21(i, k) */
        for (int k = 0; k < kb; ++k) memcpy(pack + (size_t)k * m, A21 + (size_t)k * n, sizeof(double) * (size_t)m);
        /* SYRK: A22 (lower) -= A21 A21^T with an 8x4 register tile */
        double* A22 = A + (k0 + kb) + (size_t)(k0 + kb) * n;
        for (int jb = 0; jb < m; jb += 4) {
            int jw = m - jb < 4 ? m - jb : 4;
            for (int ib = jb; ib < m; ib += 8) {
                int iw = m - ib < 8 ? m - ib : 8;
                double acc[4][8];
                memset(acc, 0, sizeof acc);
                if (iw == 8 && jw == 4) {
                    for (int k = 0; k < kb; ++k) {
                        const double* pk = pack + (size_t)k * m;
                        for (int jj = 0; jj < 4; ++jj) {
                            double bj = pk[jb + jj];
                            for (int ii = 0; ii < 8; ++ii) acc[jj][ii] += pk[ib + ii] * bj;
                        }
                    }
                } else {
                    for (int k = 0; k < kb; ++k) {
                        const double* pk = pack + (size_t)k * m;
                        for (int jj = 0; jj < jw; ++jj) {
                            double bj = pk[jb + jj];
                            for (int ii = 0; ii < iw; ++ii) acc[jj][ii] += pk[ib + ii] * bj;
                        }
                    }
                }
                for (int jj = 0; jj < jw; ++jj) {
                    double* col = A22 + (size_t)(jb + jj) * n;
                    for (int ii = 0; ii < iw; ++ii)
                        if (ib + ii >= jb + jj) col[ib + ii] -= acc[jj][ii];
                }
            }
        }
    }
    free(pack);
    return ok;
}

void orc_dense_cholesky_solve(int n, const double* L, double* b) {
    for (int j = 0; j < n; ++j) { /* forward: L y = b (column sweep) */
        double y = b[j] / L[j + (size_t)j * n];
        b[j] = y;
        const double* col = L + (size_t)j * n;
        for (int i = j + 1; i < n; ++i) b[i] -= col[i] * y;
    }
    for (int j = n - 1; j >= 0; --j) { /* backward: L^T x = y (dot products) */
        const double* col = L + (size_t)j * n;
        double s = b[j];
        for (int i = j + 1; i < n; ++i) s -= col[i] * b[i];
        b[j] = s / col[j];
    }
}

/* ------------------------------------------------------------------ */
/* Schur complement step (src/lm.cu:483-643)                           */
/* ------------------------------------------------------------------ */

typedef struct {
    int C, P, O, dim, nworkers;
    const int *oc, *op, *poff, *pidx;
    const double *W, *Vinv, *bp;
    double *T, *Spriv, *rpriv, *S;
} schur_ctx;

/* Phase 1: T_e = W_e V'^-1_{p(e)} (lm.cu:492-514) */
static void schur_phase1(void* vctx, int t, int nt) {
    schur_ctx* c = (schur_ctx*)vctx;
    long b, e;
    range_of(c->O, t, nt, &b, &e);
    for (long k = b; k < e; ++k) {
        const double* We = c->W + 27 * k;
        const double* Vi = c->Vinv + 9 * (size_t)c->op[k];
        double* Te = c->T + 27 * k;
        for (int i = 0; i < 9; ++i)
            for (int j = 0; j < 3; ++j)
                Te[3 * i + j] = We[3 * i] * Vi[j] + We[3 * i + 1] * Vi[3 + j] + We[3 * i + 2] * Vi[6 + j];
    }
}

/* Phase 2: per-worker private S / rhs over a strided point subset
 * (lm.cu:515-559); S is column-major, only blocks with cf <= ce written. */
static void schur_phase2(void* vctx, int t, int nt) {
    schur_ctx* c = (schur_ctx*)vctx;
    (void)nt;
    size_t dim = (size_t)c->dim;
    double* Sp = c->Spriv + (size_t)t * dim * dim;
    double* rp = c->rpriv + (size_t)t * dim;
    for (int p = t; p < c->P; p += c->nworkers) {
        int k0 = c->poff[p], k1 = c->poff[p + 1];
        if (k0 == k1) continue;
        const double* bpp = c->bp + 3 * (size_t)p;
        for (int a = k0; a < k1; ++a) {
            int e = c->pidx[a];
            int ce = c->oc[e];
            const double* Te = c->T + 27 * (size_t)e;
            double* rseg = rp + 9 * (size_t)ce;
            for (int i = 0; i < 9; ++i) rseg[i] -= Te[3 * i] * bpp[0] + Te[3 * i + 1] * bpp[1] + Te[3 * i + 2] * bpp[2];
            for (int bb = k0; bb < k1; ++bb) {
                int f = c->pidx[bb];
                int cf = c->oc[f];
                if (cf > ce) continue;
                const double* Wf = c->W + 27 * (size_t)f;
                for (int j = 0; j < 9; ++j) {
                    double* col = Sp + (9 * (size_t)cf + j) * dim + 9 * (size_t)ce;
                    for (int i = 0; i < 9; ++i)
                        col[i] -= Te[3 * i] * Wf[3 * j] + Te[3 * i + 1] * Wf[3 * j + 1] + Te[3 * i + 2] * Wf[3 * j + 2];
                }
            }
        }
    }
}

/* Merge of the private copies (lm.cu:560-579) */
static void schur_merge(void* vctx, int t, int nt) {
    schur_ctx* c = (schur_ctx*)vctx;
    size_t total = (size_t)c->dim * (size_t)c->dim;
    long b, e;
    range_of((long)total, t, nt, &b, &e);
    for (long i = b; i < e; ++i) {
        double s = 0.0;
        for (int w = 0; w < c->nworkers; ++w) s += c->Spriv[(size_t)w * total + i];
        c->S[i] = s;
    }
}

static int schur_step_impl(int C, int P, int O, const int* oc, const int* op, const int* poff,
                           const int* pidx, const double* U, const double* V, const double* W,
                           const double* bc, const double* bp, double lambda, double* Vinv,
                           double* T, double* Spriv, double* rpriv, int nworkers, double* S,
                           double* rhs, double* dc, double* dp, double* denom, int nthreads,
                           double* t_dampinv, double* t_assembly, double* t_solve,
                           double* S_copy, double* rhs_copy) {
    double ta = now_s();
    long bad = orc_damp_invert_v(P, V, lambda, Vinv);
    if (t_dampinv) *t_dampinv += now_s() - ta;
    if (bad) return ORC_TRIAL_NONSPD_V;
    ta = now_s();
    int dim = 9 * C;
    memset(Spriv, 0, sizeof(double) * (size_t)nworkers * dim * dim);
    memset(rpriv, 0, sizeof(double) * (size_t)nworkers * dim);
    schur_ctx c;
    c.C = C;
    c.P = P;
    c.O = O;
    c.dim = dim;
    c.nworkers = nworkers;
    c.oc = oc;
    c.op = op;
    c.poff = poff;
    c.pidx = pidx;
    c.W = W;
    c.Vinv = Vinv;
    c.bp = bp;
    c.T = T;
    c.Spriv = Spriv;
    c.rpriv = rpriv;
    c.S = S;
    parallel_run(nthreads, schur_phase1, &c);
    parallel_run(nworkers, schur_phase2, &c);
    parallel_run(nthreads, schur_merge, &c);
    for (int i = 0; i < dim; ++i) { /* lm.cu:580-584 */
        double s = bc[i];
        for (int w = 0; w < nworkers; ++w) s += rpriv[(size_t)w * dim + i];
        rhs[i] = s;
    }
    for (int cam = 0; cam < C; ++cam) /* lm.cu:585-594 */
        for (int i = 0; i < 9; ++i) {
            double dg = U[81 * (size_t)cam + 9 * i + i];
            for (int j = 0; j <= i; ++j) {
                double v = U[81 * (size_t)cam + 9 * i + j];
                if (i == j) v += lambda * damp_diag(dg);
                S[(9 * (size_t)cam + i) + (9 * (size_t)cam + j) * (size_t)dim] += v;
            }
        }
    if (t_assembly) *t_assembly += now_s() - ta;
    if (S_copy) memcpy(S_copy, S, sizeof(double) * (size_t)dim * dim);
    if (rhs_copy) memcpy(rhs_copy, rhs, sizeof(double) * (size_t)dim);
    ta = now_s();
    if (!orc_dense_cholesky(dim, S)) {
        if (t_solve) *t_solve += now_s() - ta;
        return ORC_TRIAL_LLT_FAIL;
    }
    memcpy(dc, rhs, sizeof(double) * (size_t)dim);
    orc_dense_cholesky_solve(dim, S, dc);
    /* back-substitution and gain denominator (lm.cu:606-643) */
    double den = 0.0;
    for (int cam = 0; cam < C; ++cam)
        for (int i = 0; i < 9; ++i) {
            double dg = U[81 * (size_t)cam + 9 * i + i];
            double d = dc[9 * cam + i];
            den += d * (lambda * damp_diag(dg) * d + bc[9 * cam + i]);
        }
    for (int p = 0; p < P; ++p) {
        double m0 = bp[3 * (size_t)p], m1 = bp[3 * (size_t)p + 1], m2 = bp[3 * (size_t)p + 2];
        for (int a = poff[p]; a < poff[p + 1]; ++a) {
            int e = pidx[a];
            const double* We = W + 27 * (size_t)e;
            const double* dce = dc + 9 * (size_t)oc[e];
            for (int i = 0; i < 9; ++i) {
                m0 -= We[3 * i] * dce[i];
                m1 -= We[3 * i + 1] * dce[i];
                m2 -= We[3 * i + 2] * dce[i];
            }
        }
        const double* Vi = Vinv + 9 * (size_t)p;
        double d0 = Vi[0] * m0 + Vi[1] * m1 + Vi[2] * m2;
        double d1 = Vi[3] * m0 + Vi[4] * m1 + Vi[5] * m2;
        double d2 = Vi[6] * m0 + Vi[7] * m1 + Vi[8] * m2;
        dp[3 * (size_t)p] = d0;
        dp[3 * (size_t)p + 1] = d1;
        dp[3 * (size_t)p + 2] = d2;
        const double* Vp = V + 9 * (size_t)p;
        den += d0 * (lambda * damp_diag(Vp[0]) * d0 + bp[3 * (size_t)p]) +
               d1 * (lambda * damp_diag(Vp[4]) * d1 + bp[3 * (size_t)p + 1]) +
               d2 * (lambda * damp_diag(Vp[8]) * d2 + bp[3 * (size_t)p + 2]);
    }
    *denom = den;
    if (t_solve) *t_solve += now_s() - ta;
    if (!(den > 0.0)) return ORC_TRIAL_DENOM;
    return ORC_TRIAL_OK;
}

static int workers_for(int nthreads, int C, int P) {
    long dim = 9L * C;
    long bymem = (16L << 20) / (dim * dim * 8 > 0 ? dim * dim * 8 : 1);
    if (bymem < 1) bymem = 1;
    long w = nthreads;
    if (P < w) w = P;
    if (bymem < w) w = bymem;
    return w < 1 ? 1 : (int)w;
}

int orc_schur_step(int C, int P, int O, const int* oc, const int* op, const double* U,
                   const double* V, const double* W, const double* bc, const double* bp,
                   double lambda, double* dc, double* dp, double* denom, double* S_out,
                   double* rhs_out, int nthreads) {
    int nt = nthreads < 1 ? 1 : (nthreads < C ? nthreads : C);
    if (nt < 1) nt = 1;
    int nworkers = workers_for(nt, C, P);
    int dim = 9 * C;
    int *poff, *pidx;
    build_csr(P, O, op, &poff, &pidx);
    double* Vinv = (double*)malloc(sizeof(double) * 9 * (size_t)P);
    double* T = (double*)malloc(sizeof(double) * 27 * (size_t)(O > 0 ? O : 1));
    double* Spriv = (double*)malloc(sizeof(double) * (size_t)nworkers * dim * dim);
    double* rpriv = (double*)malloc(sizeof(double) * (size_t)nworkers * dim);
    double* S = (double*)malloc(sizeof(double) * (size_t)dim * dim);
    double* rhs = (double*)malloc(sizeof(double) * (size_t)dim);
    *denom = 0.0;
    int st = schur_step_impl(C, P, O, oc, op, poff, pidx, U, V, W, bc, bp, lambda, Vinv, T, Spriv,
                             rpriv, nworkers, S, rhs, dc, dp, denom, nt, NULL, NULL, NULL, S_out,
                             rhs_out);
    free(poff);
    free(pidx);
    free(Vinv);
    free(T);
    free(Spriv);
    free(rpriv);
    free(S);
    free(rhs);
    return st;
}

/* ------------------------------------------------------------------ */
/* full LM solve (src/lm.cu:371-724)                                    */
/* ------------------------------------------------------------------ */

/* ORC_MUTATE = k (k > 0) builds a copy of the oracle with one LM control rule
 * deliberately broken; tests/test_oracle_lm_schedule.py compiles each copy
 * and checks that its schedule checks reject it.  The product build and the
 * oracle proper use ORC_MUTATE = 0.
 *   1 nu not reset per outer iteration   2 nu never doubles
 *   3 accept clamp 1/2 instead of 1/3     4 accept factor from rho - 1
 *   5 textbook gain ratio                 6-9 no escalation after a non-SPD V,
 *   LLT failure, bad denominator, non-finite trial   10 no escalation after a
 *   plain rejection   11 relative decrease over sse, not max(sse, 1)
 *   12 nu doubled before it scales lambda 13 one extra inner trial */
#ifndef ORC_MUTATE
#define ORC_MUTATE 0
#endif

int orc_lm_solve(int C, int P, int O, double* cams, double* pts, const int* oc, const int* op,
                 const double* px, const orc_options* opts, orc_result* res, double* trace,
                 int max_trace, int* n_trace, int nthreads) {
    double t0 = now_s();
    memset(res, 0, sizeof *res);
    if (n_trace) *n_trace = 0;
    int hw = nthreads < 1 ? 1 : nthreads;
    int nt = C < hw ? C : hw; /* lm.cu:382-383 */
    if (nt < 1) nt = 1;
    int nworkers = workers_for(nt, C, P); /* lm.cu:449-451 */
    int dim = 9 * C;
    size_t sC = (size_t)C, sP = (size_t)P, sO = (size_t)(O > 0 ? O : 1);
    double* U = (double*)malloc(sizeof(double) * 81 * sC);
    double* V = (double*)malloc(sizeof(double) * 9 * sP);
    double* Vinv = (double*)malloc(sizeof(double) * 9 * sP);
    double* W = (double*)calloc(27 * sO, sizeof(double));
    double* bc = (double*)malloc(sizeof(double) * 9 * sC);
    double* bp = (double*)malloc(sizeof(double) * 3 * sP);
    double* T = (double*)malloc(sizeof(double) * 27 * sO);
    double* Spriv = (double*)malloc(sizeof(double) * (size_t)nworkers * dim * dim);
    double* rpriv = (double*)malloc(sizeof(double) * (size_t)nworkers * dim);
    double* S = (double*)malloc(sizeof(double) * (size_t)dim * dim);
    double* rhs = (double*)malloc(sizeof(double) * (size_t)dim);
    double* dc = (double*)malloc(sizeof(double) * (size_t)dim);
    double* dp = (double*)malloc(sizeof(double) * 3 * sP);
    double* tc = (double*)malloc(sizeof(double) * 9 * sC);
    double* tp = (double*)malloc(sizeof(double) * 3 * sP);
    int *poff, *pidx;
    build_csr(P, O, op, &poff, &pidx);

    unsigned long long counter0 = 0, an = 0;
    double sse = 0.0, ta;
    ta = now_s();
    orc_linearize_loss(C, P, O, cams, pts, oc, op, px, opts->huber_delta, U, V, W, bc, bp, &sse, &an, hw);
    res->t_linearize += now_s() - ta;
    counter0 += an;
    res->initial_cost = sse / 2.0;
    double lambda = opts->initial_lambda;
    int converged = 0, failed = 0;
    double rel_decrease = 0.0;
#if ORC_MUTATE == 1
    double nu = 2.0; /* planted: nu not reset per outer iteration */
#endif
    while (res->iterations < opts->max_iterations && !converged && !failed) {
#if ORC_MUTATE != 1
        double nu = 2.0; /* lm.cu:461 */
#endif
        int accepted = 0;
        for (int trial = 0; trial < opts->max_inner_trials + (ORC_MUTATE == 13); ++trial) {
            unsigned long long before0 = counter0;
            double denom = 0.0, lam_used = lambda;
            res->trials++;
            int st = schur_step_impl(C, P, O, oc, op, poff, pidx, U, V, W, bc, bp, lambda, Vinv, T,
                                     Spriv, rpriv, nworkers, S, rhs, dc, dp, &denom, nt,
                                     &res->t_dampinv, &res->t_assembly, &res->t_solve, NULL,
                                     NULL);
            int code = -1;
            double tcost = NAN, rho = NAN;
            if (st == ORC_TRIAL_NONSPD_V) code = ORC_TR_NONSPD_V;
            else if (st == ORC_TRIAL_LLT_FAIL) code = ORC_TR_LLT_FAIL;
            else if (st == ORC_TRIAL_DENOM) code = ORC_TR_DENOM;
            if (code < 0) {
                ta = now_s();
                for (size_t i = 0; i < 9 * sC; ++i) tc[i] = cams[i] + dc[i];
                for (size_t i = 0; i < 3 * sP; ++i) tp[i] = pts[i] + dp[i];
                unsigned long long tan_ = 0;
                double tsse = orc_cost_loss(O, tc, tp, oc, op, px, opts->huber_delta, &tan_, hw);
                counter0 += tan_;
                res->t_trial_cost += now_s() - ta;
                tcost = tsse / 2.0;
                if (counter0 != before0 || !isfinite(tsse)) {
                    code = ORC_TR_NONFINITE;
                } else {
#if ORC_MUTATE == 5
                    rho = (sse - tsse) / denom; /* planted: textbook ratio */
#else
                    rho = (sse - tsse) / (0.5 * denom); /* lm.cu:672 */
#endif
                    if (rho > 0.0 && tsse < sse) {      /* lm.cu:678-686 */
                        memcpy(cams, tc, sizeof(double) * 9 * sC);
                        memcpy(pts, tp, sizeof(double) * 3 * sP);
#if ORC_MUTATE == 11
                        rel_decrease = (sse - tsse) / sse; /* planted: no max(sse, 1) */
#else
                        rel_decrease = (sse - tsse) / (sse > 1.0 ? sse : 1.0);
#endif
                        sse = tsse;
#if ORC_MUTATE == 4
                        double fac = 1.0 - pow(rho - 1.0, 3.0); /* planted: textbook rho in the update */
#else
                        double fac = 1.0 - pow(2.0 * rho - 1.0, 3.0); /* lm.cu:683 */
#endif
#if ORC_MUTATE == 3
                        lambda *= (fac > 0.5 ? fac : 0.5); /* planted: wrong lower clamp */
#else
                        lambda *= (fac > 1.0 / 3.0 ? fac : 1.0 / 3.0);
#endif
                        accepted = 1;
                        code = ORC_TR_ACCEPT;
                    } else {
                        code = ORC_TR_REJECT;
                    }
                }
            }
            if (trace && n_trace && *n_trace < max_trace) {
                double* tr = trace + ORC_TRACE_STRIDE * (size_t)(*n_trace);
                tr[0] = res->iterations;
                tr[1] = lam_used;
                tr[2] = tcost;
                tr[3] = rho;
                tr[4] = code;
                tr[5] = nu;
                tr[6] = denom;
                (*n_trace)++;
            }
            if (opts->verbose)
                fprintf(stderr, "  iter %d trial: lambda %.3g cost %.9g rho %.3g code %d\n",
                        res->iterations, lambda, tcost, rho, code);
            if (accepted) break;
            /* every rejection path escalates the same way (lm.cu:478-479,
             * 600-601, 640-641, 667-668, 687-688) */
#if ORC_MUTATE >= 6 && ORC_MUTATE <= 10
            {
                /* planted: one failure path leaves lambda and nu alone */
                static const int skip_code[5] = {ORC_TR_NONSPD_V, ORC_TR_LLT_FAIL, ORC_TR_DENOM,
                                                 ORC_TR_NONFINITE, ORC_TR_REJECT};
                if (code == skip_code[ORC_MUTATE - 6]) continue;
            }
#endif
#if ORC_MUTATE == 12
            nu *= 2.0; /* planted: nu doubled before it scales lambda */
            lambda *= nu;
#elif ORC_MUTATE == 2
            lambda *= nu; /* planted: nu never doubles */
#else
            lambda *= nu;
            nu *= 2.0;
#endif
        }
        if (!accepted) {
            failed = 1;
            break;
        }
        ++res->iterations;
        if (rel_decrease < opts->rel_decrease_tol) {
            converged = 1;
            break;
        }
        /* re-linearise at the accepted parameters (lm.cu:700-702); the
         * current cost stays the trial cost (lm.cu:682), not this SSE */
        double lin_sse;
        ta = now_s();
        orc_linearize_loss(C, P, O, cams, pts, oc, op, px, opts->huber_delta, U, V, W, bc, bp, &lin_sse, &an,
                           hw);
        res->t_linearize += now_s() - ta;
        counter0 += an;
    }
    res->final_cost = sse / 2.0;
    res->anomalies = counter0;
    res->termination = converged ? ORC_CONVERGENCE : (failed ? ORC_FAILURE : ORC_MAX_ITERATIONS);
    res->total_time_s = now_s() - t0;
    free(U);
    free(V);
    free(Vinv);
    free(W);
    free(bc);
    free(bp);
    free(T);
    free(Spriv);
    free(rpriv);
    free(S);
    free(rhs);
    free(dc);
    free(dp);
    free(tc);
    free(tp);
    free(poff);
    free(pidx);
    return 0;
}
