/*
 * CPU oracle for the 2D MLS-MPM step path — TEST INFRASTRUCTURE ONLY.
 * See mpm_oracle.h for scope and for what pins each function.
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off; the integer path relies
 * on every f32 product and sum being rounded separately).
 */
#include "mpm_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* reference/solver.ts:266-283 */
#define WALL_CUSHION_CELLS 8.0
#define MIN_LIQUID_DETERMINANT 0.5
#define MAX_LIQUID_DETERMINANT 2.0
#define MIN_POWDER_DETERMINANT 0.5
#define MIN_NEO_HOOKEAN_STRETCH 0.05
#define MAX_NEO_HOOKEAN_STRETCH 20.0
#define MAX_DEFORMATION_INCREMENT 0.25
#define MAX_DEFORMATION_SUBSTEPS 64
/* materials.ts:40-53 */
#define LIQUID_CONTACT_SEPARATION_FRACTION 0.25
#define LIQUID_CONTACT_APPROACH_SPEED_MULTIPLIER 1.0
#define LIQUID_CONTACT_RELAXATION_SECONDS 0.05
#define LIQUID_STABLE_ORDER_EPSILON 1e-3
#define SNOW_COMPRESSION_YIELD 0.025
#define SNOW_STRETCH_YIELD 0.0075
#define SNOW_HARDENING 10.0
#define SNOW_MIN_PLASTIC_VOLUME 0.6
#define SNOW_MAX_PLASTIC_VOLUME 20.0
#define SNOW_MAX_HARDENING_EXPONENT 20.0

typedef struct { double m00, m01, m10, m11; } mat2;

static int finite(double v) { return isfinite(v); }
static int mat_finite(mat2 m) { return finite(m.m00) && finite(m.m01) && finite(m.m10) && finite(m.m11); }
static double det2(mat2 m) { return m.m00 * m.m11 - m.m01 * m.m10; }  /* solver.ts:2501-2503 */
static mat2 mul2(mat2 a, mat2 b) {                                      /* solver.ts:2474-2481 */
    mat2 r = {a.m00 * b.m00 + a.m01 * b.m10, a.m00 * b.m01 + a.m01 * b.m11,
              a.m10 * b.m00 + a.m11 * b.m10, a.m10 * b.m01 + a.m11 * b.m11};
    return r;
}
static int is_liquid(int model) { return model == ORC_MODEL_WATER || model == ORC_MODEL_OIL; }

/* decompose2x2 (solver.ts:2425-2463): closed-form SVD through the eigen-angle
 * of F^T F. */
typedef struct { double s1, s2, u1x, u1y, u2x, u2y, v1x, v1y, v2x, v2y; mat2 rot; } decomp2;
static int decompose2x2(mat2 F, decomp2* d) {
    double c00 = F.m00 * F.m00 + F.m10 * F.m10;
    double c01 = F.m00 * F.m01 + F.m10 * F.m11;
    double c11 = F.m01 * F.m01 + F.m11 * F.m11;
    double radius = hypot(c00 - c11, 2 * c01);
    double s1 = sqrt(fmax(0.0, (c00 + c11 + radius) * 0.5));
    double s2 = det2(F) / s1;
    if (!finite(s1) || !finite(s2) || s1 <= 0 || s2 <= 0) return ORC_EDETERMINANT;
    double angle = 0.5 * atan2(2 * c01, c00 - c11);
    double c = cos(angle), s = sin(angle);
    d->s1 = s1; d->s2 = s2;
    d->v1x = c; d->v1y = s; d->v2x = -s; d->v2y = c;
    d->u1x = (F.m00 * d->v1x + F.m01 * d->v1y) / s1;
    d->u1y = (F.m10 * d->v1x + F.m11 * d->v1y) / s1;
    d->u2x = -d->u1y; d->u2y = d->u1x;
    d->rot.m00 = d->u1x * d->v1x + d->u2x * d->v2x;
    d->rot.m01 = d->u1x * d->v1y + d->u2x * d->v2y;
    d->rot.m10 = d->u1y * d->v1x + d->u2y * d->v2x;
    d->rot.m11 = d->u1y * d->v1y + d->u2y * d->v2y;
    if (!mat_finite(d->rot)) return ORC_ENUMERIC;
    return ORC_OK;
}
static mat2 reconstruct(const decomp2* d, double s1, double s2) {
    mat2 r = {s1 * d->u1x * d->v1x + s2 * d->u2x * d->v2x, s1 * d->u1x * d->v1y + s2 * d->u2x * d->v2y,
              s1 * d->u1y * d->v1x + s2 * d->u2y * d->v2x, s1 * d->u1y * d->v1y + s2 * d->u2y * d->v2y};
    return r;
}

/* constitutiveStressTimesFt (solver.ts:2148-2213) with neoHookeanFirstPiola
 * (:2036-2061) and granularStressTimesFt (:2222-2265). */
static int stress_times_ft(const orc_particle* p, mat2 F, mat2* out) {
    if (is_liquid(p->model)) {
        double J = det2(F);
        double s = p->param > 0 ? -J * p->lambda * (pow(1.0 / J, p->param) - 1.0)
                                : p->lambda * fmin(J - 1.0, 0.0);
        mat2 r = {s, 0, 0, s};
        if (!mat_finite(r)) return ORC_ENUMERIC;
        *out = r;
        return ORC_OK;
    }
    if (p->model == ORC_MODEL_SNOW) {
        decomp2 d;
        int e = decompose2x2(F, &d);
        if (e) return e;
        double hexp = fmax(-SNOW_MAX_HARDENING_EXPONENT,
                           fmin(SNOW_MAX_HARDENING_EXPONENT, SNOW_HARDENING * (1 - p->plastic_volume)));
        double h = exp(hexp);
        double lambda = p->lambda * h, mu = p->mu * h;
        double J = det2(F);
        double shear = 2 * mu, vol = lambda * J * (J - 1);
        double d00 = F.m00 - d.rot.m00, d01 = F.m01 - d.rot.m01, d10 = F.m10 - d.rot.m10, d11 = F.m11 - d.rot.m11;
        mat2 r = {shear * (d00 * F.m00 + d01 * F.m01) + vol, shear * (d00 * F.m10 + d01 * F.m11),
                  shear * (d10 * F.m00 + d11 * F.m01), shear * (d10 * F.m10 + d11 * F.m11) + vol};
        if (!mat_finite(r)) return ORC_ENUMERIC;
        *out = r;
        return ORC_OK;
    }
    if (p->model == ORC_MODEL_POWDER) {
        decomp2 d;
        int e = decompose2x2(F, &d);
        if (e) return e;
        double l1 = log(d.s1), l2 = log(d.s2), lv = l1 + l2;
        double blv = fmax(log(MIN_POWDER_DETERMINANT), lv);
        double pressure = fmax(0.0, -(p->lambda + p->mu) * blv);
        double dev = 0.5 * (l1 - l2);
        double maxdev = (p->param * pressure) / (2 * p->mu);
        double pdev = fmin(dev, maxdev);
        double mean = 0.5 * lv;
        double ps1 = 2 * p->mu * (mean + pdev) + p->lambda * lv;
        double ps2 = 2 * p->mu * (mean - pdev) + p->lambda * lv;
        mat2 r = {ps1 * d.u1x * d.u1x + ps2 * d.u2x * d.u2x, ps1 * d.u1x * d.u1y + ps2 * d.u2x * d.u2y,
                  ps1 * d.u1y * d.u1x + ps2 * d.u2y * d.u2x, ps1 * d.u1y * d.u1y + ps2 * d.u2y * d.u2y};
        if (!mat_finite(r)) return ORC_ENUMERIC;
        *out = r;
        return ORC_OK;
    }
    /* Neo-Hookean */
    double J = det2(F);
    mat2 it = {F.m11 / J, -F.m10 / J, -F.m01 / J, F.m00 / J};
    double pressure = p->lambda * log(J);
    mat2 P = {p->mu * (F.m00 - it.m00) + pressure * it.m00, p->mu * (F.m01 - it.m01) + pressure * it.m01,
              p->mu * (F.m10 - it.m10) + pressure * it.m10, p->mu * (F.m11 - it.m11) + pressure * it.m11};
    if (!mat_finite(P)) return ORC_ENUMERIC;
    mat2 Ft = {F.m00, F.m10, F.m01, F.m11};
    mat2 r = mul2(P, Ft);
    double scale = fmax(0.0, 1.0 - p->param);
    r.m00 *= scale; r.m01 *= scale; r.m10 *= scale; r.m11 *= scale;
    *out = r;
    return ORC_OK;
}

/* projectSnowDeformation (solver.ts:2267-2303) */
static int project_snow(mat2 F, double pv, mat2* out, double* pv_out) {
    decomp2 d;
    int e = decompose2x2(F, &d);
    if (e) return e;
    double lo = 1 - SNOW_COMPRESSION_YIELD, hi = 1 + SNOW_STRETCH_YIELD;
    double s1 = fmax(lo, fmin(hi, d.s1)), s2 = fmax(lo, fmin(hi, d.s2));
    double upv = fmax(SNOW_MIN_PLASTIC_VOLUME, fmin(SNOW_MAX_PLASTIC_VOLUME, (pv * det2(F)) / (s1 * s2)));
    mat2 r = reconstruct(&d, s1, s2);
    if (!mat_finite(r) || !finite(upv)) return ORC_ENUMERIC;
    if (det2(r) <= 0) return ORC_EDETERMINANT;
    *out = r;
    *pv_out = upv;
    return ORC_OK;
}

/* projectGranularDeformation (solver.ts:2305-2357) */
static int project_powder(mat2 F, const orc_particle* p, mat2* out) {
    decomp2 d;
    int e = decompose2x2(F, &d);
    if (e) return e;
    double l1 = log(d.s1), l2 = log(d.s2), lv = l1 + l2;
    double minlv = log(MIN_POWDER_DETERMINANT);
    double blv = fmax(minlv, lv);
    double pressure = fmax(0.0, -(p->lambda + p->mu) * blv);
    double dev = 0.5 * (l1 - l2);
    double maxdev = (p->param * pressure) / (2 * p->mu);
    if (dev <= maxdev && lv >= minlv) { *out = F; return ORC_OK; }
    double pdev = fmin(dev, maxdev);
    double mean = 0.5 * blv;
    mat2 r = reconstruct(&d, exp(mean + pdev), exp(mean - pdev));
    if (!mat_finite(r)) return ORC_ENUMERIC;
    if (det2(r) <= 0) return ORC_EDETERMINANT;
    *out = r;
    return ORC_OK;
}

/* projectLiquidDeformation (solver.ts:2359-2392) */
static int project_liquid(mat2 F, mat2* out) {
    double J = det2(F);
    double cn = F.m00 + F.m11, sn = F.m10 - F.m01;
    double mag = hypot(cn, sn);
    if (!finite(mag) || mag <= 0 || !finite(J) || J <= 0) return ORC_EDETERMINANT;
    double bd = fmin(MAX_LIQUID_DETERMINANT, fmax(MIN_LIQUID_DETERMINANT, J));
    double sc = sqrt(bd), c = cn / mag, s = sn / mag;
    mat2 r = {sc * c, -sc * s, sc * s, sc * c};
    *out = r;
    return ORC_OK;
}

/* projectNeoHookeanDeformation (solver.ts:2394-2423) */
static int project_nh(mat2 F, mat2* out) {
    decomp2 d;
    int e = decompose2x2(F, &d);
    if (e) return e;
    double s1 = fmax(MIN_NEO_HOOKEAN_STRETCH, fmin(MAX_NEO_HOOKEAN_STRETCH, d.s1));
    double s2 = fmax(MIN_NEO_HOOKEAN_STRETCH, fmin(MAX_NEO_HOOKEAN_STRETCH, d.s2));
    mat2 r = reconstruct(&d, s1, s2);
    if (!mat_finite(r)) return ORC_ENUMERIC;
    if (det2(r) <= 0) return ORC_EDETERMINANT;
    *out = r;
    return ORC_OK;
}

/* updateDeformationGradient (solver.ts:2561-2592) */
static int update_deformation(mat2 F, mat2 C, double dt, mat2* out) {
    double row = fmax(fabs(C.m00) + fabs(C.m01), fabs(C.m10) + fabs(C.m11));
    double inc = dt * row;
    if (!finite(inc)) return ORC_ENUMERIC;
    double s = fmax(1.0, ceil(inc / MAX_DEFORMATION_INCREMENT));
    if (s > MAX_DEFORMATION_SUBSTEPS) return ORC_ESUBSTEP;
    int steps = (int)s;
    double h = dt / steps;
    mat2 u = {1 + h * C.m00, h * C.m01, h * C.m10, 1 + h * C.m11};
    mat2 r = F;
    for (int k = 0; k < steps; ++k) r = mul2(u, r);
    if (!mat_finite(r)) return ORC_ENUMERIC;
    *out = r;
    return ORC_OK;
}

/* confineAxis (solver.ts:2528-2540) */
static void confine_axis(double x, double v, double maximum, double* xo, double* vo) {
    if (x < 0 || (x == 0 && v < 0)) { *xo = 0; *vo = 0; return; }
    if (x > maximum || (x == maximum && v > 0)) { *xo = maximum; *vo = 0; return; }
    *xo = x; *vo = v;
}
/* applyWallCushion / applyWallTangentialFriction (solver.ts:2542-2559) */
static double wall_cushion(double v, double coord, double extent) {
    double dist = v < 0 ? coord : extent - coord;
    double cushion = fmin(WALL_CUSHION_CELLS, extent / 4);
    return v * fmax(0.0, fmin(1.0, dist / cushion));
}
static double wall_friction(double v, double coord, double extent, double friction) {
    if (friction == 0) return v;
    double dist = fmax(0.0, fmin(coord, extent - coord));
    double cushion = fmin(WALL_CUSHION_CELLS, extent / 4);
    double infl = 1 - fmax(0.0, fmin(1.0, dist / cushion));
    return v * (1 - friction * infl);
}

/* updateParticleStencil (solver.ts:2017-2034) */
typedef struct { long bx, by; double fx, fy, wx[3], wy[3]; } stencil_t;
static void make_stencil(double x, double y, double dx, stencil_t* s) {
    double gx = x / dx, gy = y / dx;
    s->bx = (long)floor(gx - 0.5);
    s->by = (long)floor(gy - 0.5);
    s->fx = gx - (double)s->bx;
    s->fy = gy - (double)s->by;
    s->wx[0] = 0.5 * (1.5 - s->fx) * (1.5 - s->fx);
    s->wx[1] = 0.75 - (s->fx - 1) * (s->fx - 1);
    s->wx[2] = 0.5 * (s->fx - 0.5) * (s->fx - 0.5);
    s->wy[0] = 0.5 * (1.5 - s->fy) * (1.5 - s->fy);
    s->wy[1] = 0.75 - (s->fy - 1) * (s->fy - 1);
    s->wy[2] = 0.5 * (s->fy - 0.5) * (s->fy - 0.5);
}

enum { N_M, N_QX, N_QY, N_VX, N_VY, N_WM, N_WQX, N_WQY, N_WVX, N_WVY, N_WV, N_OM, N_OQX, N_OQY, N_OVX, N_OVY, N_OV, N_FIELDS };

static double particle_mass(const orc_particle* p, const orc_config* c) {
    return is_liquid(p->model) ? p->density * p->volume : c->particle_mass;  /* solver.ts:2130-2146 */
}

static int validate(const orc_particle* in, int32_t n, const orc_config* c) {
    if (!(c->dt > 0) || c->cells_x < 1 || c->cells_y < 1 || !(c->cell_size > 0) ||
        !(c->particle_mass > 0) || !(c->particle_volume > 0) || c->wall_friction < 0 || c->wall_friction > 1)
        return ORC_EINVALID;
    double W = c->cells_x * c->cell_size, H = c->cells_y * c->cell_size;
    for (int32_t i = 0; i < n; ++i) {
        const orc_particle* p = &in[i];
        if (p->model == ORC_MODEL_REMOVED) continue;
        if (p->bonded_vertex != 0) return ORC_EUNSUPPORTED;
        if (p->model < 0 || p->model > 4) return ORC_EINVALID;
        if (!(p->volume > 0)) return ORC_EINVALID;
        if (p->x < 0 || p->x > W || p->y < 0 || p->y > H) return ORC_EINVALID;
        mat2 F = {p->f00, p->f01, p->f10, p->f11};
        double J = det2(F);
        if (!finite(J) || J <= 0) return ORC_EINVALID;
        if (p->plastic_volume < SNOW_MIN_PLASTIC_VOLUME || p->plastic_volume > SNOW_MAX_PLASTIC_VOLUME)
            return ORC_EINVALID;
    }
    return ORC_OK;
}

/* P2G (solver.ts:418-479) + grid update (:494-608). */
static int p2g_and_update(const orc_particle* in, int32_t n, const orc_config* c, double* G,
                          orc_diagnostics* diag) {
    const long W = c->cells_x + 3, H = c->cells_y + 3;
    const double dx = c->cell_size, gscale = 4.0 / (dx * dx);
    memset(G, 0, sizeof(double) * (size_t)(W * H) * N_FIELDS);
    for (int32_t i = 0; i < n; ++i) {
        const orc_particle* p = &in[i];
        if (p->model == ORC_MODEL_REMOVED) continue;
        stencil_t s;
        make_stencil(p->x, p->y, dx, &s);
        if (s.bx < -1 || s.by < -1 || s.bx + 2 > c->cells_x + 1 || s.by + 2 > c->cells_y + 1) return ORC_EINVALID;
        mat2 F = {p->f00, p->f01, p->f10, p->f11};
        mat2 PFt;
        int e = stress_times_ft(p, F, &PFt);
        if (e) return e;
        double m = particle_mass(p, c);
        double ss = -c->dt * p->volume * gscale;
        double liqV = is_liquid(p->model) ? p->volume * det2(F) : 0;
        mat2 A = {p->c00 * m + PFt.m00 * ss, p->c01 * m + PFt.m01 * ss, p->c10 * m + PFt.m10 * ss,
                  p->c11 * m + PFt.m11 * ss};
        if (!mat_finite(A)) return ORC_ENUMERIC;
        for (int oy = 0; oy < 3; ++oy)
            for (int ox = 0; ox < 3; ++ox) {
                double w = s.wx[ox] * s.wy[oy];
                if (w == 0) continue;
                long key = (s.by + oy + 1) * W + (s.bx + ox + 1);
                double* nd = G + key * N_FIELDS;
                double dpx = (ox - s.fx) * dx, dpy = (oy - s.fy) * dx;
                double qx = m * p->vx + A.m00 * dpx + A.m01 * dpy;
                double qy = m * p->vy + A.m10 * dpx + A.m11 * dpy;
                if (p->model == ORC_MODEL_WATER) {
                    nd[N_WM] += w * m; nd[N_WQX] += w * qx; nd[N_WQY] += w * qy; nd[N_WV] += w * liqV;
                } else if (p->model == ORC_MODEL_OIL) {
                    nd[N_OM] += w * m; nd[N_OQX] += w * qx; nd[N_OQY] += w * qy; nd[N_OV] += w * liqV;
                } else {
                    nd[N_M] += w * m; nd[N_QX] += w * qx; nd[N_QY] += w * qy;
                }
            }
    }
    double gm = 0, gpx = 0, gpy = 0;
    for (long k = 0; k < W * H; ++k) {
        const double* nd = G + k * N_FIELDS;
        gm += nd[N_M] + nd[N_WM] + nd[N_OM];
        gpx += nd[N_QX] + nd[N_WQX] + nd[N_OQX];
        gpy += nd[N_QY] + nd[N_WQY] + nd[N_OQY];
    }
    if (diag) { diag->grid_mass = gm; diag->grid_px = gpx; diag->grid_py = gpy; }

    const double gmag = hypot(c->gravity_x, c->gravity_y);
    const double vscale = 1.0 / (2 * dx * dx);
    for (long sy = 0; sy < H; ++sy)
        for (long sx = 0; sx < W; ++sx) {
            double* nd = G + (sy * W + sx) * N_FIELDS;
            long xi = sx - 1, yi = sy - 1;
            /* updateGridVelocity (solver.ts:1507-1524) */
            double vel[3][2];
            const int mi[3] = {N_M, N_WM, N_OM};
            for (int ph = 0; ph < 3; ++ph) {
                double mass = nd[mi[ph]];
                if (mass <= 0) { vel[ph][0] = vel[ph][1] = 0; continue; }
                vel[ph][0] = nd[mi[ph] + 1] / mass + c->dt * c->gravity_x;
                vel[ph][1] = nd[mi[ph] + 2] / mass + c->dt * c->gravity_y;
            }
            /* solveLiquidGridContact + volumeGradient (solver.ts:1293-1409) */
            if (nd[N_WM] > 0 && nd[N_OM] > 0 && nd[N_WV] > 0 && nd[N_OV] > 0 && xi >= 0 && xi <= c->cells_x &&
                yi >= 0 && yi <= c->cells_y) {
#define VOL(X, Y, F) (((X) < 0 || (X) >= W || (Y) < 0 || (Y) >= H) ? 0.0 : G[((Y) * W + (X)) * N_FIELDS + (F)])
                double wgx = (VOL(sx + 1, sy, N_WV) - VOL(sx - 1, sy, N_WV)) * vscale;
                double wgy = (VOL(sx, sy + 1, N_WV) - VOL(sx, sy - 1, N_WV)) * vscale;
                double ogx = (VOL(sx + 1, sy, N_OV) - VOL(sx - 1, sy, N_OV)) * vscale;
                double ogy = (VOL(sx, sy + 1, N_OV) - VOL(sx, sy - 1, N_OV)) * vscale;
#undef VOL
                double gx = ogx - wgx, gy = ogy - wgy;
                double gl = hypot(gx, gy);
                if (finite(gl) && gl > 1e-12) {
                    double nx = gx / gl, ny = gy / gl;
                    double wd = nd[N_WM] / nd[N_WV], od = nd[N_OM] / nd[N_OV];
                    double dscale = fmax(wd, od);
                    double contrast = (od - wd) / dscale;
                    double align = gmag > 0 ? (nx * c->gravity_x + ny * c->gravity_y) / gmag : 0;
                    if (!finite(wd) || !finite(od) || !finite(contrast)) return ORC_ENUMERIC;
                    if (contrast * align < -LIQUID_STABLE_ORDER_EPSILON) { nx = -nx; ny = -ny; }
                    double rel = (vel[2][0] - vel[1][0]) * nx + (vel[2][1] - vel[1][1]) * ny;
                    double overlap = fmin(1.0, (2 * fmin(nd[N_WV], nd[N_OV])) / (nd[N_WV] + nd[N_OV]));
                    double minsep = fmax(LIQUID_CONTACT_SEPARATION_FRACTION * overlap * dx /
                                             LIQUID_CONTACT_RELAXATION_SECONDS,
                                         fmax(-rel, 0.0) * LIQUID_CONTACT_APPROACH_SPEED_MULTIPLIER);
                    if (rel < minsep) {
                        double imp = (minsep - rel) / (1 / nd[N_WM] + 1 / nd[N_OM]);
                        vel[1][0] -= (imp / nd[N_WM]) * nx;
                        vel[1][1] -= (imp / nd[N_WM]) * ny;
                        vel[2][0] += (imp / nd[N_OM]) * nx;
                        vel[2][1] += (imp / nd[N_OM]) * ny;
                    }
                }
            }
            for (int ph = 0; ph < 3; ++ph) {
                double vx = wall_cushion(vel[ph][0], (double)xi, (double)c->cells_x);
                double vy = wall_cushion(vel[ph][1], (double)yi, (double)c->cells_y);
                vx = wall_friction(vx, (double)yi, (double)c->cells_y, c->wall_friction);
                vy = wall_friction(vy, (double)xi, (double)c->cells_x, c->wall_friction);
                if (!finite(vx) || !finite(vy)) return ORC_ENUMERIC;
                nd[mi[ph] + 3] = vx;
                nd[mi[ph] + 4] = vy;
            }
        }
    return ORC_OK;
}

int orc_grid_f64(const orc_particle* in, int32_t n, const orc_config* c, double* node_out) {
    int e = validate(in, n, c);
    if (e) return e;
    return p2g_and_update(in, n, c, node_out, NULL);
}

int orc_step_f64(const orc_particle* in, int32_t n, const orc_config* c, orc_particle* out,
                 orc_diagnostics* diag) {
    int e = validate(in, n, c);
    if (e) return e;
    const long W = c->cells_x + 3, H = c->cells_y + 3;
    const double dx = c->cell_size, gscale = 4.0 / (dx * dx);
    double* G = (double*)malloc(sizeof(double) * (size_t)(W * H) * N_FIELDS);
    if (!G) return ORC_EINVALID;
    orc_diagnostics dg;
    memset(&dg, 0, sizeof dg);
    for (int32_t i = 0; i < n; ++i) {
        if (in[i].model == ORC_MODEL_REMOVED) continue;
        double m = particle_mass(&in[i], c);
        dg.input_mass += m;
        dg.input_px += m * in[i].vx;
        dg.input_py += m * in[i].vy;
    }
    e = p2g_and_update(in, n, c, G, &dg);
    if (e) { free(G); return e; }

    const double Wd = c->cells_x * dx, Hd = c->cells_y * dx;
    for (int32_t i = 0; i < n && !e; ++i) {
        const orc_particle* p = &in[i];
        orc_particle* q = &out[i];
        *q = *p;
        if (p->model == ORC_MODEL_REMOVED) continue;
        /* G2P (solver.ts:610-773) */
        stencil_t s;
        make_stencil(p->x, p->y, dx, &s);
        const int water = p->model == ORC_MODEL_WATER, oil = p->model == ORC_MODEL_OIL, liquid = water || oil;
        double vx = 0, vy = 0, c00 = 0, c01 = 0, c10 = 0, c11 = 0, lm = 0, lv = 0;
        for (int oy = 0; oy < 3; ++oy)
            for (int ox = 0; ox < 3; ++ox) {
                double w = s.wx[ox] * s.wy[oy];
                if (w == 0) continue;
                long key = (s.by + oy + 1) * W + (s.bx + ox + 1);
                const double* nd = G + key * N_FIELDS;
                double dpx = (ox - s.fx) * dx, dpy = (oy - s.fy) * dx;
                double nvx = water ? nd[N_WVX] : oil ? nd[N_OVX] : nd[N_VX];
                double nvy = water ? nd[N_WVY] : oil ? nd[N_OVY] : nd[N_VY];
                vx += w * nvx;
                vy += w * nvy;
                if (liquid) {
                    lm += w * (nd[N_WM] + nd[N_OM]);
                    lv += w * (nd[N_WV] + nd[N_OV]);
                }
                c00 += w * nvx * dpx; c01 += w * nvx * dpy;
                c10 += w * nvy * dpx; c11 += w * nvy * dpy;
            }
        mat2 F = {p->f00, p->f01, p->f10, p->f11};
        if (liquid) {
            if (lm <= 0 || lv <= 0) { e = ORC_ENUMERIC; break; }
            double m = particle_mass(p, c);
            double pd = m / (p->volume * det2(F));
            double ld = lm / lv;
            if (!finite(pd) || !finite(ld) || pd <= 0 || ld <= 0) { e = ORC_ENUMERIC; break; }
            double rel = (pd - ld) / pd;
            vx += c->dt * c->gravity_x * rel;
            vy += c->dt * c->gravity_y * rel;
        }
        mat2 C = {c00 * gscale, c01 * gscale, c10 * gscale, c11 * gscale};
        if (!finite(vx) || !finite(vy) || !mat_finite(C)) { e = ORC_ENUMERIC; break; }
        mat2 trial;
        e = update_deformation(F, C, c->dt, &trial);
        if (e) break;
        double J = det2(trial);
        if (!finite(J) || J <= 0) { e = ORC_EDETERMINANT; break; }
        mat2 Fn = trial;
        double pv = p->plastic_volume;
        if (p->model == ORC_MODEL_SNOW) e = project_snow(trial, pv, &Fn, &pv);
        else if (p->model == ORC_MODEL_POWDER) e = project_powder(trial, p, &Fn);
        else if (liquid) e = project_liquid(trial, &Fn);
        else {
            double tr = trial.m00 * trial.m00 + trial.m01 * trial.m01 + trial.m10 * trial.m10 + trial.m11 * trial.m11;
            if (!finite(tr)) { e = ORC_ENUMERIC; break; }
            if (J < MIN_NEO_HOOKEAN_STRETCH * sqrt(tr) || tr > MAX_NEO_HOOKEAN_STRETCH * MAX_NEO_HOOKEAN_STRETCH)
                e = project_nh(trial, &Fn);
        }
        if (e) break;
        double nx = p->x + c->dt * vx, ny = p->y + c->dt * vy;
        if (!finite(nx) || !finite(ny)) { e = ORC_ENUMERIC; break; }
        double bx, bvx, by, bvy, fx, fvx, fy, fvy;
        confine_axis(nx, vx, Wd, &bx, &bvx);
        confine_axis(ny, vy, Hd, &by, &bvy);
        /* projectParticleOutsideTriangles is the identity without bonded bodies */
        confine_axis(bx, bvx, Wd, &fx, &fvx);
        confine_axis(by, bvy, Hd, &fy, &fvy);
        q->x = fx; q->y = fy; q->vx = fvx; q->vy = fvy;
        q->c00 = C.m00; q->c01 = C.m01; q->c10 = C.m10; q->c11 = C.m11;
        q->f00 = Fn.m00; q->f01 = Fn.m01; q->f10 = Fn.m10; q->f11 = Fn.m11;
        q->plastic_volume = pv;
    }
    free(G);
    if (e) return e;
    for (int32_t i = 0; i < n; ++i) {
        if (out[i].model == ORC_MODEL_REMOVED) continue;
        double m = particle_mass(&out[i], c);
        dg.output_mass += m;
        dg.output_px += m * out[i].vx;
        dg.output_py += m * out[i].vy;
    }
    if (diag) *diag = dg;
    return ORC_OK;
}

int orc_run_f64(orc_particle* p, int32_t n, const orc_config* c, int32_t steps) {
    orc_particle* tmp = (orc_particle*)malloc(sizeof(orc_particle) * (size_t)(n > 0 ? n : 1));
    if (!tmp) return ORC_EINVALID;
    for (int32_t s = 0; s < steps; ++s) {
        int e = orc_step_f64(p, n, c, tmp, NULL);
        if (e) { free(tmp); return e; }
        memcpy(p, tmp, sizeof(orc_particle) * (size_t)n);
    }
    free(tmp);
    return ORC_OK;
}

/* ======================= float32 integer path ======================== */

#define FIXED_POINT_SCALE 16777216.0f      /* shaders.ts:173 */
#define FIXED_POINT_LIMIT 2147483520.0f    /* shaders.ts:974 */
#define ERR_STENCIL 1u
#define ERR_DEFORMATION 2u
#define ERR_P2G 4u
#define ERR_INVALID_MATERIAL 128u
#define ERR_NONFINITE_PARTICLE 256u
#define ERR_FIXED_POINT 65536u

static float wf(const uint32_t* r, int k) { float f; memcpy(&f, &r[k], 4); return f; }
static int ff(float v) { return isfinite(v); }

/* determinant2x2 (shaders.ts:261-266): compensated product difference. */
static float det_comp(float a, float b, float c, float d) {
    float sp = b * c;
    float spe = fmaf(b, c, -sp);
    return fmaf(a, d, -sp) - spe;
}

/* Solid stress P*F^T evaluated in FP64 from the f32 deformation and rounded to
 * f32 once per component (the GPU kernel does the same). */
static int solid_stress_f64(uint32_t model, const float* F, float lambda, float mu, float param, float pv,
                            float out[4]) {
    orc_particle p;
    memset(&p, 0, sizeof p);
    p.model = (int32_t)model;
    p.lambda = lambda; p.mu = mu; p.param = param; p.plastic_volume = pv;
    if (model == ORC_MODEL_NEO_HOOKEAN) p.param = param < 0 ? 0 : (param > 1 ? 1 : param);
    mat2 Fm = {F[0], F[1], F[2], F[3]};
    mat2 r;
    if (stress_times_ft(&p, Fm, &r)) return 0;
    out[0] = (float)r.m00; out[1] = (float)r.m01; out[2] = (float)r.m10; out[3] = (float)r.m11;
    return 1;
}

/* prepareParticle (shaders.ts:808-963). Returns flags; writes A. */
static uint32_t prepare_one(const uint32_t* r, const orc_params_f32* prm, float A[4]) {
    uint32_t model = r[13];
    float F[4] = {wf(r, 8), wf(r, 9), wf(r, 10), wf(r, 11)};
    float C[4] = {wf(r, 4), wf(r, 5), wf(r, 6), wf(r, 7)};
    float pv = wf(r, 12), lambda = wf(r, 14), mu = wf(r, 15), density = wf(r, 16), volume = wf(r, 17),
          param = wf(r, 18);
    A[0] = A[1] = A[2] = A[3] = 0.0f;
    if (model == 0xFFFFFFFFu) return 0;
    float det = det_comp(F[0], F[1], F[2], F[3]);
    if (!ff(det) || det <= 0.0f || !ff(volume) || volume <= 0.0f) return ERR_DEFORMATION;
    float S[4] = {0, 0, 0, 0};
    if (model == 2u || model == 3u) {
        if (!ff(density) || density <= 0.0f || !ff(lambda) || lambda <= 0.0f || !ff(param) || param < 0.0f)
            return ERR_P2G | ERR_INVALID_MATERIAL;
        if (param > 0.0f) {
            float inv = 1.0f / det;
            float pw = (float)pow((double)inv, (double)param);
            float pressure = lambda * (pw - 1.0f);
            float stress = -det * pressure;
            S[0] = stress; S[3] = stress;
        } else {
            float jm1 = det - 1.0f;
            float pressure = lambda * (jm1 < 0.0f ? jm1 : 0.0f);
            S[0] = pressure; S[3] = pressure;
        }
    } else if (model == 0u) {
        if (!ff(param) || param < 0.0f || param > 1.0f) return ERR_P2G | ERR_INVALID_MATERIAL;
        if (!solid_stress_f64(model, F, lambda, mu, param, pv, S)) return ERR_P2G | ERR_NONFINITE_PARTICLE;
    } else if (model == 1u) {
        if (!ff(pv) || pv < 0.6f || pv > 20.0f) return ERR_P2G | ERR_INVALID_MATERIAL;
        if (!solid_stress_f64(model, F, lambda, mu, param, pv, S)) return ERR_P2G | ERR_INVALID_MATERIAL;
    } else if (model == 4u) {
        if (!ff(lambda) || !ff(mu) || mu <= 0.0f || !ff(param) || param < 0.0f)
            return ERR_P2G | ERR_INVALID_MATERIAL;
        if (!solid_stress_f64(model, F, lambda, mu, param, pv, S)) return ERR_P2G | ERR_NONFINITE_PARTICLE;
    } else {
        return ERR_P2G | ERR_INVALID_MATERIAL;
    }
    float mass = (model == 2u || model == 3u) ? density * volume : prm->particle_mass;
    float ss = (-prm->dt * volume) * prm->gradient_scale;
    for (int k = 0; k < 4; ++k) {
        float a = mass * C[k];
        float b = ss * S[k];
        A[k] = a + b;
    }
    for (int k = 0; k < 4; ++k)
        if (!ff(S[k]) || !ff(A[k])) return ERR_P2G | ERR_NONFINITE_PARTICLE;
    return 0;
}

uint32_t orc_prepare_f32(const uint32_t* packed20, uint32_t n, const orc_params_f32* prm, float* affine_out) {
    uint32_t flags = 0;
    for (uint32_t i = 0; i < n; ++i) flags |= prepare_one(packed20 + 20 * (size_t)i, prm, affine_out + 4 * (size_t)i);
    return flags;
}

/* quadraticWeights (shaders.ts:500-509) */
static void weights_f32(float f, float w[3]) {
    float l = 1.5f - f, c = f - 1.0f, r = f - 0.5f;
    w[0] = (0.5f * l) * l;
    w[1] = 0.75f - c * c;
    w[2] = (0.5f * r) * r;
}

static int stencil_f32(const uint32_t* r, const orc_params_f32* prm, int* bx, int* by, float* fx, float* fy) {
    float gx = wf(r, 0) / prm->cell_size, gy = wf(r, 1) / prm->cell_size;
    if (!ff(gx) || !ff(gy)) return 0;
    float b0 = floorf(gx - 0.5f), b1 = floorf(gy - 0.5f);
    *bx = (int)b0; *by = (int)b1;
    if (*bx < -1 || *by < -1 || *bx + 2 > (int)prm->cells_x + 1 || *by + 2 > (int)prm->cells_y + 1) return 0;
    *fx = gx - (float)*bx;
    *fy = gy - (float)*by;
    return 1;
}

static uint32_t fixed_add(int64_t* acc, float value) {
    float scaled = rintf(value * FIXED_POINT_SCALE);
    if (!ff(scaled) || fabsf(scaled) > FIXED_POINT_LIMIT) return ERR_FIXED_POINT;
    *acc += (int64_t)(int32_t)scaled;
    return 0;
}

uint32_t orc_fixed_p2g_f32(const uint32_t* packed20, uint32_t n, const orc_params_f32* prm, int32_t* grid_out) {
    const size_t words = (size_t)prm->node_count * 11;
    int64_t* acc = (int64_t*)calloc(words, sizeof(int64_t));
    uint32_t flags = 0;
    for (uint32_t i = 0; i < n && acc; ++i) {
        const uint32_t* r = packed20 + 20 * (size_t)i;
        uint32_t model = r[13];
        if (model == 0xFFFFFFFFu) continue;
        float A[4];
        uint32_t f = prepare_one(r, prm, A);
        if (f) { flags |= f; continue; }
        int bx, by;
        float fx, fy;
        if (!stencil_f32(r, prm, &bx, &by, &fx, &fy)) { flags |= ERR_STENCIL; continue; }
        float wx[3], wy[3];
        weights_f32(fx, wx);
        weights_f32(fy, wy);
        int liquid = model == 2u || model == 3u;
        float mass = liquid ? wf(r, 16) * wf(r, 17) : prm->particle_mass;
        float volume = wf(r, 17) * det_comp(wf(r, 8), wf(r, 9), wf(r, 10), wf(r, 11));
        float vx = wf(r, 2), vy = wf(r, 3);
        for (int oy = 0; oy < 3; ++oy)
            for (int ox = 0; ox < 3; ++ox) {
                float w = wx[ox] * wy[oy];
                float dpx = ((float)ox - fx) * prm->cell_size;
                float dpy = ((float)oy - fy) * prm->cell_size;
                float mx = mass * vx, ax = A[0] * dpx, ay = A[1] * dpy;
                float qx = (mx + ax) + ay;
                float my = mass * vy, bxv = A[2] * dpx, byv = A[3] * dpy;
                float qy = (my + bxv) + byv;
                if (!ff(w) || w < 0.0f || !ff(qx) || !ff(qy)) { flags |= ERR_P2G | ERR_NONFINITE_PARTICLE; goto next; }
                size_t node = (size_t)(by + oy + 1) * prm->node_width + (size_t)(bx + ox + 1);
                size_t field = node * 11 + (model == 2u ? 3 : model == 3u ? 7 : 0);
                flags |= fixed_add(&acc[field], (w * mass) / prm->particle_mass);
                flags |= fixed_add(&acc[field + 1], (w * qx) / prm->particle_mass);
                flags |= fixed_add(&acc[field + 2], (w * qy) / prm->particle_mass);
                if (liquid) flags |= fixed_add(&acc[field + 3], (w * volume) / prm->particle_volume);
            }
    next:;
    }
    if (!acc) return ERR_FIXED_POINT;
    for (size_t k = 0; k < words; ++k) {
        if (acc[k] > INT32_MAX || acc[k] < INT32_MIN) flags |= ERR_FIXED_POINT;
        grid_out[k] = (int32_t)(uint32_t)(uint64_t)acc[k];
    }
    free(acc);
    return flags;
}

void orc_decode_fixed_f32(const int32_t* grid, uint32_t node_count, const orc_params_f32* prm, float* node_out) {
    for (size_t k = 0; k < (size_t)node_count * 11; ++k) {
        size_t field = k % 11;
        float v = (float)grid[k] / FIXED_POINT_SCALE;
        int is_volume = field == 6 || field == 10;
        node_out[k] = v * (is_volume ? prm->particle_volume : prm->particle_mass);
    }
}

uint32_t orc_bin(const uint32_t* packed20, uint32_t n, const orc_params_f32* prm, uint32_t* key_out,
                 uint32_t* start_out, uint32_t* perm_out) {
    uint32_t flags = 0;
    memset(start_out, 0, sizeof(uint32_t) * ((size_t)prm->node_count + 1));
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t* r = packed20 + 20 * (size_t)i;
        key_out[i] = 0xFFFFFFFFu;
        if (r[13] == 0xFFFFFFFFu) continue;
        int bx, by;
        float fx, fy;
        if (!stencil_f32(r, prm, &bx, &by, &fx, &fy)) { flags |= ERR_STENCIL; continue; }
        key_out[i] = (uint32_t)(by + 1) * prm->node_width + (uint32_t)(bx + 1);
        start_out[key_out[i] + 1]++;
    }
    for (uint32_t k = 0; k < prm->node_count; ++k) start_out[k + 1] += start_out[k];
    uint32_t* cursor = (uint32_t*)malloc(sizeof(uint32_t) * ((size_t)prm->node_count + 1));
    if (!cursor) return flags | ERR_STENCIL;
    memcpy(cursor, start_out, sizeof(uint32_t) * ((size_t)prm->node_count + 1));
    for (uint32_t i = 0; i < n; ++i)
        if (key_out[i] != 0xFFFFFFFFu) perm_out[cursor[key_out[i]]++] = i;
    free(cursor);
    return flags;
}
