/*
 * FP64 differentiable rollout and its reverse pass — TEST INFRASTRUCTURE ONLY.
 * See mpm_adjoint.h for scope, the reference lines each part follows, and
 * what pins these functions.
 *
 * The reverse pass stores every state of the rollout (CPU memory is ample at
 * the sizes the tests use) and recomputes each step's grid from its state.
 */
#include "mpm_adjoint.h"
#include "mpm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* reference/solver.ts:266-283 */
#define CUSHION_CELLS 8.0
#define MAX_INCREMENT 0.25
#define MAX_SUBSTEPS 64

enum { IX = 0, IY = 1, IVX = 2, IVY = 3, IC = 4, IF = 8 };
#define SW ORC_DIFF_STATE

/* Stencil of one particle: base cell, weights and their x/y derivatives. */
typedef struct {
    long bx, by;
    double fx, fy;
    double wx[3], wy[3];
    double gx[3], gy[3]; /* d wx / d x, d wy / d y */
} stn_t;

/* updateParticleStencil (solver.ts:2017-2034) and the stencil range check of
 * P2G (solver.ts:430-437). */
static int make_stn(const orc_diff_config* c, double x, double y, stn_t* s) {
    const double dx = c->cell_size;
    const double gx = x / dx, gy = y / dx;
    s->bx = (long)floor(gx - 0.5);
    s->by = (long)floor(gy - 0.5);
    if (!isfinite(gx) || !isfinite(gy) || s->bx < -1 || s->by < -1 || s->bx + 2 > c->cells_x + 1 ||
        s->by + 2 > c->cells_y + 1)
        return ORC_EINVALID;
    const double fx = gx - (double)s->bx, fy = gy - (double)s->by;
    s->fx = fx;
    s->fy = fy;
    s->wx[0] = 0.5 * (1.5 - fx) * (1.5 - fx);
    s->wx[1] = 0.75 - (fx - 1) * (fx - 1);
    s->wx[2] = 0.5 * (fx - 0.5) * (fx - 0.5);
    s->wy[0] = 0.5 * (1.5 - fy) * (1.5 - fy);
    s->wy[1] = 0.75 - (fy - 1) * (fy - 1);
    s->wy[2] = 0.5 * (fy - 0.5) * (fy - 0.5);
    s->gx[0] = (fx - 1.5) / dx;
    s->gx[1] = -2 * (fx - 1) / dx;
    s->gx[2] = (fx - 0.5) / dx;
    s->gy[0] = (fy - 1.5) / dx;
    s->gy[1] = -2 * (fy - 1) / dx;
    s->gy[2] = (fy - 0.5) / dx;
    return ORC_OK;
}

/* Fixed-corotated P F^T = 2 mu (F - R) F^T + lambda J (J - 1) I: the snow
 * branch of constitutiveStressTimesFt (solver.ts:2171-2203) with hardening
 * factor exp(10 (1 - 1)) = 1. */
static int corot(const double* F, double lam, double mu, double* K) {
    const double a = F[0], b = F[1], c = F[2], d = F[3];
    const double J = a * d - b * c;
    const double x = a + d, y = c - b, r = hypot(x, y);
    if (!(J > 0) || !(r > 0) || !isfinite(J)) return ORC_EDETERMINANT;
    const double co = x / r, si = y / r;
    const double m0 = a - co, m1 = b + si, m2 = c - si, m3 = d - co;
    const double s2 = 2 * mu, vol = lam * J * (J - 1);
    K[0] = s2 * (m0 * a + m1 * b) + vol;
    K[1] = s2 * (m0 * c + m1 * d);
    K[2] = s2 * (m2 * a + m3 * b);
    K[3] = s2 * (m2 * c + m3 * d) + vol;
    return ORC_OK;
}

/* Reverse of corot: accumulates dL/dF, dL/dlambda, dL/dmu from dL/dK. */
static void corot_bwd(const double* F, double lam, double mu, const double* Kb, double* Fb, double* lamb,
                      double* mub) {
    const double a = F[0], b = F[1], c = F[2], d = F[3];
    const double J = a * d - b * c;
    const double x = a + d, y = c - b, r2 = x * x + y * y, r = sqrt(r2);
    const double co = x / r, si = y / r;
    const double m0 = a - co, m1 = b + si, m2 = c - si, m3 = d - co;
    const double s2 = 2 * mu;
    /* mu and lambda */
    *mub += 2 * (Kb[0] * (m0 * a + m1 * b) + Kb[1] * (m0 * c + m1 * d) + Kb[2] * (m2 * a + m3 * b) +
                 Kb[3] * (m2 * c + m3 * d));
    const double tr = Kb[0] + Kb[3];
    *lamb += J * (J - 1) * tr;
    /* M = F - R: dL/dM = 2 mu Kb F */
    const double mb0 = s2 * (Kb[0] * a + Kb[1] * c), mb1 = s2 * (Kb[0] * b + Kb[1] * d);
    const double mb2 = s2 * (Kb[2] * a + Kb[3] * c), mb3 = s2 * (Kb[2] * b + Kb[3] * d);
    /* F through M F^T's right factor: 2 mu Kb^T M, and through M itself */
    Fb[0] += s2 * (Kb[0] * m0 + Kb[2] * m2) + mb0;
    Fb[1] += s2 * (Kb[0] * m1 + Kb[2] * m3) + mb1;
    Fb[2] += s2 * (Kb[1] * m0 + Kb[3] * m2) + mb2;
    Fb[3] += s2 * (Kb[1] * m1 + Kb[3] * m3) + mb3;
    /* volume term: d[J (J - 1)] = (2J - 1) cof(F) : dF */
    const double jv = lam * (2 * J - 1) * tr;
    Fb[0] += jv * d;
    Fb[1] -= jv * c;
    Fb[2] -= jv * b;
    Fb[3] += jv * a;
    /* R = rot(theta), dR/dtheta = [-s -c; c -s], dL/dR = -dL/dM */
    const double thb = mb0 * si + mb1 * co - mb2 * co + mb3 * si;
    const double xb = -thb * y / r2, yb = thb * x / r2;
    Fb[0] += xb;
    Fb[3] += xb;
    Fb[2] += yb;
    Fb[1] -= yb;
}

/* applyWallCushion factor (solver.ts:2542-2551): depends on the sign of v. */
static double cushion_factor(double v, double coord, double extent) {
    const double dist = v < 0 ? coord : extent - coord;
    const double cushion = fmin(CUSHION_CELLS, extent / 4);
    return fmax(0.0, fmin(1.0, dist / cushion));
}
/* applyWallTangentialFriction factor (solver.ts:2552-2559). */
static double friction_factor(double coord, double extent, double friction) {
    if (friction == 0) return 1.0;
    const double dist = fmax(0.0, fmin(coord, extent - coord));
    const double cushion = fmin(CUSHION_CELLS, extent / 4);
    return 1 - friction * (1 - fmax(0.0, fmin(1.0, dist / cushion)));
}

typedef struct {
    long W, H;
    double *m, *qx, *qy, *vhx, *vhy, *vx, *vy;
} grid_t;

static int grid_alloc(const orc_diff_config* c, grid_t* g) {
    g->W = c->cells_x + 3;
    g->H = c->cells_y + 3;
    const size_t nn = (size_t)(g->W * g->H);
    double* base = (double*)calloc(nn * 7, sizeof(double));
    if (!base) return ORC_EINVALID;
    g->m = base;
    g->qx = base + nn;
    g->qy = base + 2 * nn;
    g->vhx = base + 3 * nn;
    g->vhy = base + 4 * nn;
    g->vx = base + 5 * nn;
    g->vy = base + 6 * nn;
    return ORC_OK;
}

/* P2G (solver.ts:418-479) and the solid-phase grid update
 * (solver.ts:1507-1524 velocity, 2542-2559 walls). */
static int grid_pass(const orc_diff_config* c, int32_t n, const double* s, const double* lam, const double* mu,
                     grid_t* g) {
    const size_t nn = (size_t)(g->W * g->H);
    memset(g->m, 0, sizeof(double) * nn * 3);
    const double dx = c->cell_size, pm = c->particle_mass;
    const double ss = -c->dt * c->particle_volume * (4.0 / (dx * dx));
    for (int32_t p = 0; p < n; ++p) {
        const double* q = s + (size_t)p * SW;
        stn_t st;
        if (make_stn(c, q[IX], q[IY], &st)) return ORC_EINVALID;
        double K[4];
        if (corot(q + IF, lam[p], mu[p], K)) return ORC_EDETERMINANT;
        double A[4];
        for (int k = 0; k < 4; ++k) A[k] = q[IC + k] * pm + K[k] * ss;
        for (int oy = 0; oy < 3; ++oy)
            for (int ox = 0; ox < 3; ++ox) {
                const double w = st.wx[ox] * st.wy[oy];
                const long key = (st.by + oy + 1) * g->W + (st.bx + ox + 1);
                const double dpx = (ox - st.fx) * dx, dpy = (oy - st.fy) * dx;
                g->m[key] += w * pm;
                g->qx[key] += w * (pm * q[IVX] + A[0] * dpx + A[1] * dpy);
                g->qy[key] += w * (pm * q[IVY] + A[2] * dpx + A[3] * dpy);
            }
    }
    for (long sy = 0; sy < g->H; ++sy)
        for (long sx = 0; sx < g->W; ++sx) {
            const long k = sy * g->W + sx;
            const double xi = (double)(sx - 1), yi = (double)(sy - 1);
            double vhx = 0, vhy = 0;
            if (g->m[k] > 0) {
                vhx = g->qx[k] / g->m[k] + c->dt * c->gravity_x;
                vhy = g->qy[k] / g->m[k] + c->dt * c->gravity_y;
            }
            g->vhx[k] = vhx;
            g->vhy[k] = vhy;
            g->vx[k] = (vhx * cushion_factor(vhx, xi, c->cells_x)) * friction_factor(yi, c->cells_y, c->wall_friction);
            g->vy[k] = (vhy * cushion_factor(vhy, yi, c->cells_y)) * friction_factor(xi, c->cells_x, c->wall_friction);
            if (!isfinite(g->vx[k]) || !isfinite(g->vy[k])) return ORC_ENUMERIC;
        }
    return ORC_OK;
}

static void mat_mul(const double* a, const double* b, double* r) {
    const double r0 = a[0] * b[0] + a[1] * b[2], r1 = a[0] * b[1] + a[1] * b[3];
    const double r2 = a[2] * b[0] + a[3] * b[2], r3 = a[2] * b[1] + a[3] * b[3];
    r[0] = r0;
    r[1] = r1;
    r[2] = r2;
    r[3] = r3;
}

/* Substep count of updateDeformationGradient (solver.ts:2561-2592). */
static int substeps(const double* C, double dt) {
    const double row = fmax(fabs(C[0]) + fabs(C[1]), fabs(C[2]) + fabs(C[3]));
    const double inc = dt * row;
    if (!isfinite(inc)) return -1;
    const double s = fmax(1.0, ceil(inc / MAX_INCREMENT));
    return s > MAX_SUBSTEPS ? -1 : (int)s;
}

/* confineAxis (solver.ts:2528-2540): 1 when the axis is clamped. */
static int clamped(double x, double v, double maximum) {
    return x < 0 || (x == 0 && v < 0) || x > maximum || (x == maximum && v > 0);
}

/* G2P (solver.ts:610-773) for one particle: gathered velocity and C. */
static void gather(const orc_diff_config* c, const grid_t* g, const stn_t* st, double* v, double* C) {
    const double dx = c->cell_size, gs = 4.0 / (dx * dx);
    double vx = 0, vy = 0, c0 = 0, c1 = 0, c2 = 0, c3 = 0;
    for (int oy = 0; oy < 3; ++oy)
        for (int ox = 0; ox < 3; ++ox) {
            const double w = st->wx[ox] * st->wy[oy];
            const long key = (st->by + oy + 1) * g->W + (st->bx + ox + 1);
            const double dpx = (ox - st->fx) * dx, dpy = (oy - st->fy) * dx;
            const double nvx = g->vx[key], nvy = g->vy[key];
            vx += w * nvx;
            vy += w * nvy;
            c0 += w * nvx * dpx;
            c1 += w * nvx * dpy;
            c2 += w * nvy * dpx;
            c3 += w * nvy * dpy;
        }
    v[0] = vx;
    v[1] = vy;
    C[0] = c0 * gs;
    C[1] = c1 * gs;
    C[2] = c2 * gs;
    C[3] = c3 * gs;
}

static int g2p_pass(const orc_diff_config* c, int32_t n, const double* s, const grid_t* g, double* out) {
    const double Wd = c->cells_x * c->cell_size, Hd = c->cells_y * c->cell_size;
    for (int32_t p = 0; p < n; ++p) {
        const double* q = s + (size_t)p * SW;
        double* o = out + (size_t)p * SW;
        stn_t st;
        if (make_stn(c, q[IX], q[IY], &st)) return ORC_EINVALID;
        double v[2], C[4];
        gather(c, g, &st, v, C);
        const int ns = substeps(C, c->dt);
        if (ns < 0) return ORC_ESUBSTEP;
        const double h = c->dt / ns;
        const double U[4] = {1 + h * C[0], h * C[1], h * C[2], 1 + h * C[3]};
        double F[4] = {q[IF], q[IF + 1], q[IF + 2], q[IF + 3]};
        for (int k = 0; k < ns; ++k) mat_mul(U, F, F);
        if (!(F[0] * F[3] - F[1] * F[2] > 0)) return ORC_EDETERMINANT;
        const double nx = q[IX] + c->dt * v[0], ny = q[IY] + c->dt * v[1];
        if (!isfinite(nx) || !isfinite(ny)) return ORC_ENUMERIC;
        const int cx = clamped(nx, v[0], Wd), cy = clamped(ny, v[1], Hd);
        o[IX] = cx ? (nx < 0 || (nx == 0 && v[0] < 0) ? 0 : Wd) : nx;
        o[IY] = cy ? (ny < 0 || (ny == 0 && v[1] < 0) ? 0 : Hd) : ny;
        o[IVX] = cx ? 0 : v[0];
        o[IVY] = cy ? 0 : v[1];
        for (int k = 0; k < 4; ++k) {
            o[IC + k] = C[k];
            o[IF + k] = F[k];
        }
    }
    return ORC_OK;
}

static int check_config(const orc_diff_config* c, int32_t n) {
    if (!c || n < 0 || !(c->dt > 0) || c->cells_x < 1 || c->cells_y < 1 || !(c->cell_size > 0) ||
        !(c->particle_mass > 0) || !(c->particle_volume > 0) || c->wall_friction < 0 || c->wall_friction > 1 ||
        c->steps < 0)
        return ORC_EINVALID;
    return ORC_OK;
}

int orc_diff_step(const orc_diff_config* c, int32_t n, const double* in, const double* lambda, const double* mu,
                  double* out) {
    int e = check_config(c, n);
    if (e) return e;
    grid_t g;
    if (grid_alloc(c, &g)) return ORC_EINVALID;
    e = grid_pass(c, n, in, lambda, mu, &g);
    if (!e) e = g2p_pass(c, n, in, &g, out);
    free(g.m);
    return e;
}

int orc_diff_forward(const orc_diff_config* c, int32_t n, const double* state0, const double* lambda,
                     const double* mu, double* state_final) {
    int e = check_config(c, n);
    if (e) return e;
    grid_t g;
    if (grid_alloc(c, &g)) return ORC_EINVALID;
    const size_t bytes = sizeof(double) * (size_t)n * SW;
    double* tmp = (double*)malloc(bytes ? bytes : 1);
    if (!tmp) {
        free(g.m);
        return ORC_EINVALID;
    }
    memcpy(state_final, state0, bytes);
    for (int32_t t = 0; t < c->steps && !e; ++t) {
        e = grid_pass(c, n, state_final, lambda, mu, &g);
        if (!e) e = g2p_pass(c, n, state_final, &g, tmp);
        if (!e) memcpy(state_final, tmp, bytes);
    }
    free(tmp);
    free(g.m);
    return e;
}

/* Reverse of F' = U^s F with U = I + h C': returns dL/dF in Fb and
 * dL/dU = sum_k (U^T)^k G F^T (U^T)^(s-1-k) in Ub, for G = dL/dF'. */
static void substep_bwd(const double* U, const double* F, int ns, const double* G, double* Fb, double* Ub) {
    const double Ut[4] = {U[0], U[2], U[1], U[3]};
    const double Ft[4] = {F[0], F[2], F[1], F[3]};
    double Q[4] = {Ft[0], Ft[1], Ft[2], Ft[3]};
    double T[4];
    mat_mul(G, Q, T);
    for (int j = 1; j < ns; ++j) {
        double GQ[4];
        mat_mul(Q, Ut, Q);
        mat_mul(Ut, T, T);
        mat_mul(G, Q, GQ);
        for (int k = 0; k < 4; ++k) T[k] += GQ[k];
    }
    for (int k = 0; k < 4; ++k) {
        Ub[k] = T[k];
        Fb[k] = G[k];
    }
    for (int j = 0; j < ns; ++j) mat_mul(Ut, Fb, Fb);
}

/* One reverse step: state s of step t, its grid g, dL/d(state t+1) in
 * ab_next -> dL/d(state t) in ab; lambda/mu gradients accumulate. */
static int step_bwd(const orc_diff_config* c, int32_t n, const double* s, const double* lam, const double* mu,
                    const grid_t* g, const double* ab_next, double* ab, double* glam, double* gmu,
                    double* node_b, int64_t* confined) {
    const size_t nn = (size_t)(g->W * g->H);
    double* vbx = node_b;
    double* vby = node_b + nn;
    double* qbx = node_b + 2 * nn;
    double* qby = node_b + 3 * nn;
    double* mb = node_b + 4 * nn;
    memset(node_b, 0, sizeof(double) * nn * 5);
    const double dx = c->cell_size, gs = 4.0 / (dx * dx), dt = c->dt;
    const double Wd = c->cells_x * dx, Hd = c->cells_y * dx;

    /* G2P and position update, reversed: scatter dL/d(node velocity). */
    for (int32_t p = 0; p < n; ++p) {
        const double* q = s + (size_t)p * SW;
        const double* an = ab_next + (size_t)p * SW;
        double* a = ab + (size_t)p * SW;
        stn_t st;
        if (make_stn(c, q[IX], q[IY], &st)) return ORC_EINVALID;
        double v[2], C[4];
        gather(c, g, &st, v, C);
        const int ns = substeps(C, dt);
        if (ns < 0) return ORC_ESUBSTEP;
        const double h = dt / ns;
        const double U[4] = {1 + h * C[0], h * C[1], h * C[2], 1 + h * C[3]};
        const double nx = q[IX] + dt * v[0], ny = q[IY] + dt * v[1];
        const int cx = clamped(nx, v[0], Wd), cy = clamped(ny, v[1], Hd);
        if (confined) *confined += cx + cy;
        const double vgb[2] = {cx ? 0 : an[IVX] + dt * an[IX], cy ? 0 : an[IVY] + dt * an[IY]};
        double xb[2] = {cx ? 0 : an[IX], cy ? 0 : an[IY]};
        double Fb[4], Ub[4];
        substep_bwd(U, q + IF, ns, an + IF, Fb, Ub);
        double Csb[4]; /* dL/d(sum w v dp^T) = gs * dL/dC' */
        for (int k = 0; k < 4; ++k) Csb[k] = gs * (an[IC + k] + h * Ub[k]);
        for (int oy = 0; oy < 3; ++oy)
            for (int ox = 0; ox < 3; ++ox) {
                const double w = st.wx[ox] * st.wy[oy];
                const long key = (st.by + oy + 1) * g->W + (st.bx + ox + 1);
                const double dpx = (ox - st.fx) * dx, dpy = (oy - st.fy) * dx;
                const double nvx = g->vx[key], nvy = g->vy[key];
                vbx[key] += w * (vgb[0] + Csb[0] * dpx + Csb[1] * dpy);
                vby[key] += w * (vgb[1] + Csb[2] * dpx + Csb[3] * dpy);
                const double wb = nvx * vgb[0] + nvy * vgb[1] + nvx * (Csb[0] * dpx + Csb[1] * dpy) +
                                  nvy * (Csb[2] * dpx + Csb[3] * dpy);
                const double dpbx = w * (Csb[0] * nvx + Csb[2] * nvy), dpby = w * (Csb[1] * nvx + Csb[3] * nvy);
                xb[0] += wb * st.gx[ox] * st.wy[oy] - dpbx;
                xb[1] += wb * st.wx[ox] * st.gy[oy] - dpby;
            }
        memset(a, 0, sizeof(double) * SW);
        a[IX] = xb[0];
        a[IY] = xb[1];
        for (int k = 0; k < 4; ++k) a[IF + k] = Fb[k];
    }

    /* Grid update, reversed. */
    for (long sy = 0; sy < g->H; ++sy)
        for (long sx = 0; sx < g->W; ++sx) {
            const long k = sy * g->W + sx;
            if (!(g->m[k] > 0)) continue;
            const double xi = (double)(sx - 1), yi = (double)(sy - 1);
            const double ax = cushion_factor(g->vhx[k], xi, c->cells_x) * friction_factor(yi, c->cells_y, c->wall_friction);
            const double ay = cushion_factor(g->vhy[k], yi, c->cells_y) * friction_factor(xi, c->cells_x, c->wall_friction);
            const double hbx = ax * vbx[k], hby = ay * vby[k];
            qbx[k] = hbx / g->m[k];
            qby[k] = hby / g->m[k];
            mb[k] = -(hbx * (g->vhx[k] - dt * c->gravity_x) + hby * (g->vhy[k] - dt * c->gravity_y)) / g->m[k];
        }

    /* P2G and stress, reversed: gather dL/d(node mass, momentum). */
    const double pm = c->particle_mass;
    const double ss = -dt * c->particle_volume * (4.0 / (dx * dx));
    for (int32_t p = 0; p < n; ++p) {
        const double* q = s + (size_t)p * SW;
        double* a = ab + (size_t)p * SW;
        stn_t st;
        make_stn(c, q[IX], q[IY], &st);
        double K[4], A[4], Ab[4] = {0, 0, 0, 0}, vb[2] = {0, 0};
        if (corot(q + IF, lam[p], mu[p], K)) return ORC_EDETERMINANT;
        for (int k = 0; k < 4; ++k) A[k] = q[IC + k] * pm + K[k] * ss;
        for (int oy = 0; oy < 3; ++oy)
            for (int ox = 0; ox < 3; ++ox) {
                const double w = st.wx[ox] * st.wy[oy];
                const long key = (st.by + oy + 1) * g->W + (st.bx + ox + 1);
                const double dpx = (ox - st.fx) * dx, dpy = (oy - st.fy) * dx;
                const double gx = qbx[key], gy = qby[key];
                const double momx = pm * q[IVX] + A[0] * dpx + A[1] * dpy;
                const double momy = pm * q[IVY] + A[2] * dpx + A[3] * dpy;
                const double wb = pm * mb[key] + gx * momx + gy * momy;
                vb[0] += w * pm * gx;
                vb[1] += w * pm * gy;
                Ab[0] += w * gx * dpx;
                Ab[1] += w * gx * dpy;
                Ab[2] += w * gy * dpx;
                Ab[3] += w * gy * dpy;
                const double dpbx = w * (A[0] * gx + A[2] * gy), dpby = w * (A[1] * gx + A[3] * gy);
                a[IX] += wb * st.gx[ox] * st.wy[oy] - dpbx;
                a[IY] += wb * st.wx[ox] * st.gy[oy] - dpby;
            }
        a[IVX] = vb[0];
        a[IVY] = vb[1];
        double Kb[4];
        for (int k = 0; k < 4; ++k) {
            a[IC + k] = pm * Ab[k];
            Kb[k] = ss * Ab[k];
        }
        corot_bwd(q + IF, lam[p], mu[p], Kb, a + IF, &glam[p], &gmu[p]);
    }
    return ORC_OK;
}

int orc_diff_backward(const orc_diff_config* c, int32_t n, const double* state0, const double* lambda,
                      const double* mu, const double* bar_final, double* bar_state0, double* grad_lambda,
                      double* grad_mu, int64_t* confined) {
    int e = check_config(c, n);
    if (e) return e;
    const int32_t T = c->steps;
    const size_t per = (size_t)n * SW;
    grid_t g;
    if (grid_alloc(c, &g)) return ORC_EINVALID;
    const size_t nn = (size_t)(g.W * g.H);
    double* states = (double*)malloc(sizeof(double) * (per * (size_t)(T + 1) + 1));
    double* ab = (double*)malloc(sizeof(double) * (2 * per + 1));
    double* node_b = (double*)malloc(sizeof(double) * nn * 5);
    if (!states || !ab || !node_b) {
        free(states);
        free(ab);
        free(node_b);
        free(g.m);
        return ORC_EINVALID;
    }
    memcpy(states, state0, sizeof(double) * per);
    for (int32_t t = 0; t < T && !e; ++t) {
        e = grid_pass(c, n, states + per * t, lambda, mu, &g);
        if (!e) e = g2p_pass(c, n, states + per * t, &g, states + per * (t + 1));
    }
    double* ab_next = ab;
    double* ab_cur = ab + per;
    memcpy(ab_next, bar_final, sizeof(double) * per);
    memset(grad_lambda, 0, sizeof(double) * (size_t)n);
    memset(grad_mu, 0, sizeof(double) * (size_t)n);
    if (confined) *confined = 0;
    for (int32_t t = T - 1; t >= 0 && !e; --t) {
        e = grid_pass(c, n, states + per * t, lambda, mu, &g);
        if (!e)
            e = step_bwd(c, n, states + per * t, lambda, mu, &g, ab_next, ab_cur, grad_lambda, grad_mu, node_b,
                         confined);
        double* sw = ab_next;
        ab_next = ab_cur;
        ab_cur = sw;
    }
    if (!e) memcpy(bar_state0, ab_next, sizeof(double) * per);
    free(states);
    free(ab);
    free(node_b);
    free(g.m);
    return e;
}
