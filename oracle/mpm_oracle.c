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
/* reference/solver.ts:277-282, materials.ts:44 */
#define DEFAULT_BOND_STIFFNESS 0.35
#define BOND_SOLVER_ITERATIONS 2
#define CONTACT_SKIN_CELLS 0.1
#define CONTACT_IMPULSE_RELAXATION 0.25
#define FIELD_BAND_CELLS 2.0
#define GRID_CONTACT_REACH_CELLS 0.75
#define FLUID_TANGENTIAL_DRAG_RATE 4.0
#define ORC_PI 3.14159265358979323846

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

/* ---- bonded triangle bodies (config-2 contact) ---- */

/* Nearest triangle of each grid node within the band (buildTriangleField,
 * solver.ts:1064-1165).  Arrays are indexed by node key; body = -1 where no
 * triangle lies within the band. */
typedef struct {
    long W, H;
    int32_t* body;
    double *dist, *nx, *ny, *b0, *b1, *b2, *ix, *iy;
    const int32_t* starts;
    int32_t nbodies;
} tri_field;

/* triangleDistanceSample (solver.ts:1190-1291) */
static int tri_sample(double px, double py, const double* t, double* sd, double* nx, double* ny, double* bary) {
    const double ax = t[0], ay = t[1], bx = t[2], by = t[3], cx = t[4], cy = t[5];
    const double abx = bx - ax, aby = by - ay, acx = cx - ax, acy = cy - ay;
    const double area = abx * acy - aby * acx;
    if (!finite(area) || fabs(area) <= 1e-10) return ORC_ENUMERIC;
    const double bcx = cx - bx, bcy = cy - by, cax = ax - cx, cay = ay - cy;
    const double c_ab = abx * (py - ay) - aby * (px - ax);
    const double c_bc = bcx * (py - by) - bcy * (px - bx);
    const double c_ca = cax * (py - cy) - cay * (px - cx);
    const int inside = area > 0 ? (c_ab >= -1e-10 && c_bc >= -1e-10 && c_ca >= -1e-10)
                                : (c_ab <= 1e-10 && c_bc <= 1e-10 && c_ca <= 1e-10);
    double best = INFINITY, bnx = 0, bny = 0;
    for (int e = 0; e < 3; ++e) {
        const double sx = t[2 * e], sy = t[2 * e + 1], ex = t[(2 * e + 2) % 6], ey = t[(2 * e + 3) % 6];
        const double dx = ex - sx, dy = ey - sy, len2 = dx * dx + dy * dy;
        const double tt = fmax(0.0, fmin(1.0, ((px - sx) * dx + (py - sy) * dy) / len2));
        const double qx = sx + dx * tt, qy = sy + dy * tt;
        const double ddx = px - qx, ddy = py - qy, d2 = ddx * ddx + ddy * ddy;
        if (d2 >= best) continue;
        best = d2;
        if (e == 0) { bary[0] = 1 - tt; bary[1] = tt; bary[2] = 0; }
        else if (e == 1) { bary[0] = 0; bary[1] = 1 - tt; bary[2] = tt; }
        else { bary[0] = tt; bary[1] = 0; bary[2] = 1 - tt; }
        if (d2 > 1e-20 && !inside) {
            const double d = sqrt(d2);
            bnx = ddx * (1 / d);
            bny = ddy * (1 / d);
        } else {
            const double o = area > 0 ? 1 : -1, el = sqrt(len2);
            bnx = (o * dy) * (1 / el);
            bny = (-o * dx) * (1 / el);
        }
    }
    const double d = sqrt(best);
    if (!finite(d) || !finite(bnx) || !finite(bny)) return ORC_ENUMERIC;
    *sd = inside ? -d : d;
    *nx = bnx;
    *ny = bny;
    return ORC_OK;
}

static void tri_vertices(const orc_particle* p, int32_t start, double* t) {
    for (int v = 0; v < 3; ++v) {
        t[2 * v] = p[start + v].x;
        t[2 * v + 1] = p[start + v].y;
    }
}

static int tri_field_build(const orc_particle* in, const orc_config* c, tri_field* f) {
    const long nn = f->W * f->H;
    for (long k = 0; k < nn; ++k) f->body[k] = -1;
    const double dx = c->cell_size, band = dx * FIELD_BAND_CELLS;
    for (int32_t b = 0; b < f->nbodies; ++b) {
        const int32_t start = f->starts[b];
        double t[6];
        tri_vertices(in, start, t);
        const double mnx = fmin(t[0], fmin(t[2], t[4])), mxx = fmax(t[0], fmax(t[2], t[4]));
        const double mny = fmin(t[1], fmin(t[3], t[5])), mxy = fmax(t[1], fmax(t[3], t[5]));
        const long x0 = (long)fmax(-1.0, floor((mnx - band) / dx) - 1), x1 = (long)fmin(c->cells_x + 1, ceil((mxx + band) / dx) + 1);
        const long y0 = (long)fmax(-1.0, floor((mny - band) / dx) - 1), y1 = (long)fmin(c->cells_y + 1, ceil((mxy + band) / dx) + 1);
        for (long yi = y0; yi <= y1; ++yi) {
            const double py = yi * dx;
            if (py < mny - band || py > mxy + band) continue;
            for (long xi = x0; xi <= x1; ++xi) {
                const double px = xi * dx;
                if (px < mnx - band || px > mxx + band) continue;
                double sd, nx, ny, bary[3] = {1, 0, 0};
                int e = tri_sample(px, py, t, &sd, &nx, &ny, bary);
                if (e) return e;
                if (sd > band) continue;
                const long key = (yi + 1) * f->W + (xi + 1);
                if (f->body[key] >= 0 && sd >= f->dist[key]) continue;
                f->body[key] = start;
                f->dist[key] = sd;
                f->nx[key] = nx;
                f->ny[key] = ny;
                f->b0[key] = bary[0];
                f->b1[key] = bary[1];
                f->b2[key] = bary[2];
                f->ix[key] = 0;
                f->iy[key] = 0;
            }
        }
    }
    return ORC_OK;
}

/* resolveFluidTriangleContact (solver.ts:1462-1505): updates the fluid
 * velocity and accumulates the impulse on the body. */
static int fluid_triangle(double fm, double* v, const double* sv, double nx, double ny, double inv_body,
                          double dt, double* ix, double* iy) {
    const double inv_eff = 1 / fm + inv_body;
    if (!finite(inv_eff) || inv_eff <= 0) return ORC_ENUMERIC;
    const double rx = v[0] - sv[0], ry = v[1] - sv[1];
    const double rn = rx * nx + ry * ny;
    const double nimp = rn < 0 ? -rn / inv_eff : 0;
    const double tx = rx - nx * rn, ty = ry - ny * rn;
    const double damping = 1 - exp(-FLUID_TANGENTIAL_DRAG_RATE * dt);
    const double fx = nx * nimp + tx * (-damping / inv_eff), fy = ny * nimp + ty * (-damping / inv_eff);
    v[0] = v[0] + fx * (1 / fm);
    v[1] = v[1] + fy * (1 / fm);
    *ix += -fx;
    *iy += -fy;
    if (!finite(v[0]) || !finite(v[1]) || !finite(*ix) || !finite(*iy)) return ORC_ENUMERIC;
    return ORC_OK;
}

/* P2G (solver.ts:418-479) + grid update (:494-608). */
/* P2G contributions of one particle (solver.ts:418-479) into G. */
static int p2g_particle(const orc_particle* p, const orc_config* c, double* G) {
    const long W = c->cells_x + 3;
    const double dx = c->cell_size, gscale = 4.0 / (dx * dx);
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
    return ORC_OK;
}

/* Grid update of node (sx, sy) (solver.ts:1293-1524): velocities from the
 * summed momenta, liquid and fluid-triangle contact, walls.  Reads the
 * volumes of the neighbours, writes only this node's velocities. */
static int grid_node(const orc_particle* in, const orc_config* c, double* G, long sx, long sy,
                     tri_field* tf) {
    const long W = c->cells_x + 3, H = c->cells_y + 3;
    const double dx = c->cell_size;
    const double gmag = hypot(c->gravity_x, c->gravity_y);
    const double vscale = 1.0 / (2 * dx * dx);
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
        /* drag spring-damper of finishGridVelocity (kernels/shaders.ts:1748-1777):
         * nodes inside the domain within the radius of the target */
        if (c->drag_radius > 0 && xi >= 0 && yi >= 0 && xi <= c->cells_x && yi <= c->cells_y) {
            const double ddx = c->drag_x - (double)xi * dx, ddy = c->drag_y - (double)yi * dx;
            const double dist = hypot(ddx, ddy);
            if (dist < c->drag_radius) {
                const double falloff = 1 - dist / c->drag_radius;
                const double ax = ddx * c->drag_stiffness - vel[ph][0] * c->drag_damping;
                const double ay = ddy * c->drag_stiffness - vel[ph][1] * c->drag_damping;
                vel[ph][0] += c->dt * falloff * ax;
                vel[ph][1] += c->dt * falloff * ay;
            }
        }
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
    /* solveFluidTriangleGridContact (solver.ts:1411-1460) */
    const long key = sy * W + sx;
    if (tf && tf->body[key] >= 0) {
        const double radius = sqrt(c->particle_volume / ORC_PI);
        if (tf->dist[key] <= radius + GRID_CONTACT_REACH_CELLS * dx) {
            const int32_t b = tf->body[key];
            const double bw[3] = {tf->b0[key], tf->b1[key], tf->b2[key]};
            double sv[2] = {0, 0}, inv_body = 0;
            for (int v = 0; v < 3; ++v) {
                sv[0] += bw[v] * in[b + v].vx;
                sv[1] += bw[v] * in[b + v].vy;
                inv_body += (bw[v] * bw[v]) / c->particle_mass;
            }
            if (inv_body > 0) {
                if (nd[N_WM] > 0) {
                    int e = fluid_triangle(nd[N_WM], vel[1], sv, tf->nx[key], tf->ny[key], inv_body, c->dt,
                                           &tf->ix[key], &tf->iy[key]);
                    if (e) return e;
                }
                if (nd[N_OM] > 0) {
                    int e = fluid_triangle(nd[N_OM], vel[2], sv, tf->nx[key], tf->ny[key], inv_body, c->dt,
                                           &tf->ix[key], &tf->iy[key]);
                    if (e) return e;
                }
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
    return ORC_OK;
}

static int p2g_and_update(const orc_particle* in, int32_t n, const orc_config* c, double* G,
                          orc_diagnostics* diag, tri_field* tf) {
    const long W = c->cells_x + 3, H = c->cells_y + 3;
    memset(G, 0, sizeof(double) * (size_t)(W * H) * N_FIELDS);
    for (int32_t i = 0; i < n; ++i) {
        if (in[i].model == ORC_MODEL_REMOVED) continue;
        int e = p2g_particle(&in[i], c, G);
        if (e) return e;
    }
    double gm = 0, gpx = 0, gpy = 0;
    for (long k = 0; k < W * H; ++k) {
        const double* nd = G + k * N_FIELDS;
        gm += nd[N_M] + nd[N_WM] + nd[N_OM];
        gpx += nd[N_QX] + nd[N_WQX] + nd[N_OQX];
        gpy += nd[N_QY] + nd[N_WQY] + nd[N_OQY];
    }
    if (diag) { diag->grid_mass = gm; diag->grid_px = gpx; diag->grid_py = gpy; }

    for (long sy = 0; sy < H; ++sy)
        for (long sx = 0; sx < W; ++sx) {
            int e = grid_node(in, c, G, sx, sy, tf);
            if (e) return e;
        }
    return ORC_OK;
}

int orc_grid_f64(const orc_particle* in, int32_t n, const orc_config* c, double* node_out) {
    int e = validate(in, n, c);
    if (e) return e;
    return p2g_and_update(in, n, c, node_out, NULL, NULL);
}

/* projectParticleOutsideTriangles (solver.ts:1526-1590) */
static int project_outside(const orc_config* c, const tri_field* f, double r, double* x, double* y) {
    stencil_t s;
    make_stencil(*x, *y, c->cell_size, &s);
    const double band = c->cell_size * FIELD_BAND_CELLS;
    double sd = 0, closest = INFINITY, snx = 0, sny = 0;
    int32_t cb = -1;
    long keys[9];
    for (int oy = 0; oy < 3; ++oy)
        for (int ox = 0; ox < 3; ++ox) {
            const double w = s.wx[ox] * s.wy[oy];
            const long key = (s.by + oy + 1) * f->W + (s.bx + ox + 1);
            const int has = key >= 0 && key < f->W * f->H && f->body[key] >= 0;
            keys[oy * 3 + ox] = has ? key : -1;
            sd += w * (has ? f->dist[key] : band);
            if (has && f->dist[key] < closest) {
                closest = f->dist[key];
                cb = f->body[key];
                snx = f->nx[key];
                sny = f->ny[key];
            }
        }
    if (cb < 0 || sd >= r) return ORC_OK;
    double nx = 0, ny = 0;
    for (int oy = 0; oy < 3; ++oy)
        for (int ox = 0; ox < 3; ++ox) {
            const long key = keys[oy * 3 + ox];
            if (key < 0 || f->body[key] != cb) continue;
            const double w = s.wx[ox] * s.wy[oy];
            nx += w * f->nx[key];
            ny += w * f->ny[key];
        }
    const double len = hypot(nx, ny);
    if (len > 1e-12 && finite(len)) {
        nx = nx / len;
        ny = ny / len;
    } else {
        nx = snx;
        ny = sny;
    }
    const double corr = r - sd;
    *x = *x + corr * nx;
    *y = *y + corr * ny;
    if (!finite(*x) || !finite(*y)) return ORC_ENUMERIC;
    return ORC_OK;
}

/* One step of particles already validated: P2G, grid update (with the
 * triangle field's fluid contact when tf is given), G2P with projection of
 * liquid particles out of the triangles.  Fills the input and grid sums of dg. */
/* G2P of one particle (solver.ts:610-773) from the updated grid G: q
 * receives the particle after the step. */
static int g2p_particle(const orc_particle* p, orc_particle* q, const orc_config* c, const double* G,
                        const tri_field* tf) {
    const long W = c->cells_x + 3;
    const double dx = c->cell_size, gscale = 4.0 / (dx * dx);
    const double Wd = c->cells_x * dx, Hd = c->cells_y * dx;
    int e = ORC_OK;
    *q = *p;
    if (p->model == ORC_MODEL_REMOVED) return ORC_OK;
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
        if (lm <= 0 || lv <= 0) return ORC_ENUMERIC;
        double m = particle_mass(p, c);
        double pd = m / (p->volume * det2(F));
        double ld = lm / lv;
        if (!finite(pd) || !finite(ld) || pd <= 0 || ld <= 0) return ORC_ENUMERIC;
        double rel = (pd - ld) / pd;
        vx += c->dt * c->gravity_x * rel;
        vy += c->dt * c->gravity_y * rel;
    }
    mat2 C = {c00 * gscale, c01 * gscale, c10 * gscale, c11 * gscale};
    if (!finite(vx) || !finite(vy) || !mat_finite(C)) return ORC_ENUMERIC;
    mat2 trial;
    e = update_deformation(F, C, c->dt, &trial);
    if (e) return e;
    double J = det2(trial);
    if (!finite(J) || J <= 0) return ORC_EDETERMINANT;
    mat2 Fn = trial;
    double pv = p->plastic_volume;
    if (p->model == ORC_MODEL_SNOW) e = project_snow(trial, pv, &Fn, &pv);
    else if (p->model == ORC_MODEL_POWDER) e = project_powder(trial, p, &Fn);
    else if (liquid) e = project_liquid(trial, &Fn);
    else {
        double tr = trial.m00 * trial.m00 + trial.m01 * trial.m01 + trial.m10 * trial.m10 + trial.m11 * trial.m11;
        if (!finite(tr)) return ORC_ENUMERIC;
        if (J < MIN_NEO_HOOKEAN_STRETCH * sqrt(tr) || tr > MAX_NEO_HOOKEAN_STRETCH * MAX_NEO_HOOKEAN_STRETCH)
            e = project_nh(trial, &Fn);
    }
    if (e) return e;
    double nx = p->x + c->dt * vx, ny = p->y + c->dt * vy;
    if (!finite(nx) || !finite(ny)) return ORC_ENUMERIC;
    double bx, bvx, by, bvy, fx, fvx, fy, fvy;
    confine_axis(nx, vx, Wd, &bx, &bvx);
    confine_axis(ny, vy, Hd, &by, &bvy);
    if (tf && tf->nbodies > 0 && liquid) {
        e = project_outside(c, tf, sqrt((p->volume * det2(Fn)) / ORC_PI), &bx, &by);
        if (e) return e;
    }
    confine_axis(bx, bvx, Wd, &fx, &fvx);
    confine_axis(by, bvy, Hd, &fy, &fvy);
    q->x = fx; q->y = fy; q->vx = fvx; q->vy = fvy;
    q->c00 = C.m00; q->c01 = C.m01; q->c10 = C.m10; q->c11 = C.m11;
    q->f00 = Fn.m00; q->f01 = Fn.m01; q->f10 = Fn.m10; q->f11 = Fn.m11;
    q->plastic_volume = pv;
    return ORC_OK;
}

static int step_core(const orc_particle* in, int32_t n, const orc_config* c, tri_field* tf, orc_particle* out,
                     orc_diagnostics* dgp) {
    int e = ORC_OK;
    const long W = c->cells_x + 3, H = c->cells_y + 3;
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
    e = p2g_and_update(in, n, c, G, &dg, tf);
    if (e) { free(G); return e; }

    for (int32_t i = 0; i < n && !e; ++i) e = g2p_particle(&in[i], &out[i], c, G, tf);
    free(G);
    *dgp = dg;
    return e;
}

static void output_sums(const orc_particle* out, int32_t n, const orc_config* c, orc_diagnostics* dg) {
    for (int32_t i = 0; i < n; ++i) {
        if (out[i].model == ORC_MODEL_REMOVED) continue;
        double m = particle_mass(&out[i], c);
        dg->output_mass += m;
        dg->output_px += m * out[i].vx;
        dg->output_py += m * out[i].vy;
    }
}

int orc_step_f64(const orc_particle* in, int32_t n, const orc_config* c, orc_particle* out,
                 orc_diagnostics* diag) {
    int e = validate(in, n, c);
    if (e) return e;
    orc_diagnostics dg;
    e = step_core(in, n, c, NULL, out, &dg);
    if (e) return e;
    output_sums(out, n, c, &dg);
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

/* All-core timing variant of orc_run_f64 (the CPU baseline of the
 * benchmark; the checker is always orc_step_f64).  Per step: validation in
 * chunks, then P2G scattered by bands of stencil-base rows with `threads`
 * OpenMP threads.  A particle whose base row is r writes node rows r..r+2, so
 * bands of at least three rows taken in two passes, even bands then odd
 * ones, never write a node at the same time.  The grid update is a parallel
 * loop over node rows (a node reads its neighbours' volumes and writes only
 * its own velocities) and G2P a parallel loop over particles into the other
 * of two buffers.  Node sums follow the band order rather than the particle
 * order, so results equal orc_run_f64's to rounding. */
int orc_run_f64_parallel(orc_particle* p, int32_t n, const orc_config* c, int32_t steps, int32_t threads) {
    if (threads < 1) threads = 1;
    const long W = c->cells_x + 3, H = c->cells_y + 3;
    const long rows = (long)c->cells_y + 2;  /* stencil-base rows by + 1 in [0, cells_y + 1] */
    long band = rows / (4L * threads);
    if (band < 3) band = 3;
    const long nbands = (rows + band - 1) / band;
    double* G = (double*)malloc(sizeof(double) * (size_t)(W * H) * N_FIELDS);
    orc_particle* tmp = (orc_particle*)malloc(sizeof(orc_particle) * (size_t)(n > 0 ? n : 1));
    int32_t* order = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    int32_t* band_of = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
    long* start = (long*)malloc(sizeof(long) * (size_t)(nbands + 1));
    int err = ORC_OK;
    if (!G || !tmp || !order || !band_of || !start) err = ORC_EINVALID;
    orc_particle* cur = p;  /* the state the step starts from; the two buffers swap each step */
    orc_particle* nxt = tmp;
    for (int32_t st = 0; st < steps && !err; ++st) {
        /* the first failing chunk's error, as the serial scan would report */
        const int32_t chunk = (n + threads - 1) / threads;
        int first = threads, code = validate(cur, 0, c);
        if (!code) {
#pragma omp parallel for num_threads(threads) schedule(static)
            for (int t = 0; t < threads; ++t) {
                const int32_t lo = t * chunk, hi = lo + chunk < n ? lo + chunk : n;
                const int e = lo < hi ? validate(cur + lo, hi - lo, c) : ORC_OK;
                if (e) {
#pragma omp critical
                    {
                        if (t < first) { first = t; code = e; }
                    }
                }
            }
        }
        err = code;
        if (err) break;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            if (cur[i].model == ORC_MODEL_REMOVED) { band_of[i] = -1; continue; }
            stencil_t s;
            make_stencil(cur[i].x, cur[i].y, c->cell_size, &s);
            long r = s.by + 1;
            if (r < 0) r = 0;
            if (r >= rows) r = rows - 1;
            band_of[i] = (int32_t)(r / band);
        }
        for (long b = 0; b <= nbands; ++b) start[b] = 0;
        for (int32_t i = 0; i < n; ++i)
            if (band_of[i] >= 0) ++start[band_of[i] + 1];
        for (long b = 0; b < nbands; ++b) start[b + 1] += start[b];
        for (int32_t i = 0; i < n; ++i)
            if (band_of[i] >= 0) order[start[band_of[i]]++] = i;
        for (long b = nbands; b > 0; --b) start[b] = start[b - 1];
        start[0] = 0;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long k = 0; k < W * H * N_FIELDS; ++k) G[k] = 0.0;
        for (int parity = 0; parity < 2; ++parity) {
#pragma omp parallel for num_threads(threads) schedule(dynamic, 1)
            for (long b = parity; b < nbands; b += 2)
                for (long k = start[b]; k < start[b + 1]; ++k) {
                    const int e = p2g_particle(&cur[order[k]], c, G);
                    if (e) {
#pragma omp atomic write
                        err = e;
                    }
                }
        }
        if (err) break;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (long sy = 0; sy < H; ++sy)
            for (long sx = 0; sx < W; ++sx) {
                const int e = grid_node(cur, c, G, sx, sy, NULL);
                if (e) {
#pragma omp atomic write
                    err = e;
                }
            }
        if (err) break;
#pragma omp parallel for num_threads(threads) schedule(static)
        for (int32_t i = 0; i < n; ++i) {
            const int e = g2p_particle(&cur[i], &nxt[i], c, G, NULL);
            if (e) {
#pragma omp atomic write
                err = e;
            }
        }
        if (err) break;
        orc_particle* t = cur;
        cur = nxt;
        nxt = t;
    }
    if (!err && cur != p) memcpy(p, cur, sizeof(orc_particle) * (size_t)n);
    free(G);
    free(tmp);
    free(order);
    free(band_of);
    free(start);
    return err;
}

/* ---- config-2 step: bonded triangle bodies and fracture bonds ---- */

/* collectBondedBodyStarts (solver.ts:1023-1062) and the particle checks of
 * validateParticles with bonded vertices allowed. */
static int validate_bodies(const orc_particle* in, int32_t n, const uint32_t* oid, const orc_config* c,
                           const orc_body_config* bc, int32_t* starts, int32_t* nbodies) {
    *nbodies = 0;
    for (int32_t i = 0; i < n; ++i) {
        if (in[i].bonded_vertex == 0) continue;
        if (in[i].bonded_vertex != 1) return ORC_EINVALID;          /* must start at vertex 0 */
        if (!(bc->bond_rest_length > 0)) return ORC_EINVALID;
        if (i + 2 >= n) return ORC_EINVALID;
        for (int v = 0; v < 3; ++v) {
            const orc_particle* m = &in[i + v];
            if (m->bonded_vertex != v + 1 || oid[i + v] != oid[i]) return ORC_EINVALID;
            if (is_liquid(m->model) || m->model == ORC_MODEL_REMOVED) return ORC_EINVALID;
        }
        starts[(*nbodies)++] = i;
        i += 2;
    }
    /* the plain checks, with bonded vertices masked */
    orc_particle* tmp = (orc_particle*)malloc(sizeof(orc_particle) * (size_t)(n > 0 ? n : 1));
    if (!tmp) return ORC_EINVALID;
    memcpy(tmp, in, sizeof(orc_particle) * (size_t)n);
    for (int32_t i = 0; i < n; ++i) tmp[i].bonded_vertex = 0;
    const int e = validate(tmp, n, c);
    free(tmp);
    return e;
}

/* validateMpmSolidBondGraph (reference/solid-bonds.ts:11-66) */
static int validate_bonds(const orc_particle* in, int32_t n, const uint32_t* oid, const orc_bond* b, int32_t nb) {
    for (int32_t k = 0; k < nb; ++k) {
        const int32_t f = b[k].first, s = b[k].second;
        if (f < 0 || f >= n || s < 0 || s >= n || f == s) return ORC_EINVALID;
        if (!(b[k].rest_length > 0) || !finite(b[k].rest_length)) return ORC_EINVALID;
        if (!(b[k].failure_strain > 0) || !finite(b[k].failure_strain)) return ORC_EINVALID;
        if (!(b[k].damage >= 0 && b[k].damage <= 1)) return ORC_EINVALID;
        if (oid[f] != oid[s]) return ORC_EINVALID;
        if (in[f].bonded_vertex != 0 || in[s].bonded_vertex != 0) return ORC_EINVALID;
        if (in[f].model != ORC_MODEL_NEO_HOOKEAN || in[s].model != ORC_MODEL_NEO_HOOKEAN) return ORC_EINVALID;
        for (int32_t j = 0; j < k; ++j)
            if ((b[j].first == f && b[j].second == s) || (b[j].first == s && b[j].second == f)) return ORC_EINVALID;
    }
    return ORC_OK;
}

/* trianglePairContact (solver.ts:1804-1874): 1 and the contact normal
 * (pointing from B to A) and penetration, or 0. */
static int pair_contact(const double* A, const double* B, double skin, double* nx, double* ny, double* pen) {
    int separated = 0, have = 0;
    double largest_gap = -INFINITY, smallest_overlap = INFINITY, axx = 0, axy = 0;
    for (int ti = 0; ti < 2; ++ti) {
        const double* T = ti == 0 ? A : B;
        for (int e = 0; e < 3; ++e) {
            const double ex = T[(2 * e + 2) % 6] - T[2 * e], ey = T[(2 * e + 3) % 6] - T[2 * e + 1];
            const double el = hypot(ex, ey);
            if (el <= 1e-10) return 0;
            const double ux = ey / el, uy = -ex / el;
            double mina = INFINITY, maxa = -INFINITY, minb = INFINITY, maxb = -INFINITY;
            for (int v = 0; v < 3; ++v) {
                const double a = A[2 * v] * ux + A[2 * v + 1] * uy, bb = B[2 * v] * ux + B[2 * v + 1] * uy;
                mina = fmin(mina, a);
                maxa = fmax(maxa, a);
                minb = fmin(minb, bb);
                maxb = fmax(maxb, bb);
            }
            const double gap = fmax(minb - maxa, mina - maxb);
            if (gap > 0) {
                separated = 1;
                if (gap > largest_gap) {
                    largest_gap = gap;
                    axx = ux;
                    axy = uy;
                    have = 1;
                }
            } else if (!separated && -gap < smallest_overlap) {
                smallest_overlap = -gap;
                axx = ux;
                axy = uy;
                have = 1;
            }
        }
    }
    if (separated && largest_gap > skin) return 0;
    if (!have) return 0;
    const double dx = (A[0] + A[2] + A[4]) / 3 - (B[0] + B[2] + B[4]) / 3;
    const double dy = (A[1] + A[3] + A[5]) / 3 - (B[1] + B[3] + B[5]) / 3;
    const double sign = dx * axx + dy * axy >= 0 ? 1 : -1;
    *nx = sign * axx;
    *ny = sign * axy;
    *pen = separated ? skin - largest_gap : skin + smallest_overlap;
    return *pen > 0;
}

static void set_confined(orc_particle* q, double x, double vx, double y, double vy, const orc_config* c) {
    double fx, fvx, fy, fvy;
    confine_axis(x, vx, c->cells_x * c->cell_size, &fx, &fvx);
    confine_axis(y, vy, c->cells_y * c->cell_size, &fy, &fvy);
    q->x = fx;
    q->vx = fvx;
    q->y = fy;
    q->vy = fvy;
}

/* applyTriangleContactImpulses, enforceBondedTriangleConstraints,
 * resolveBondedTriangleContacts (solver.ts:1592-1802). */
static int bodies_after_g2p(const orc_particle* in, orc_particle* out, const uint32_t* oid, const orc_config* c,
                            const orc_body_config* bc, const tri_field* f) {
    const int32_t nb = f->nbodies;
    const double dt = c->dt;
    /* field impulses, barycentric-weighted, in each body's node order */
    for (int32_t b = 0; b < nb; ++b) {
        const int32_t start = f->starts[b];
        double imp[3][2] = {{0, 0}, {0, 0}, {0, 0}};
        double t[6];
        tri_vertices(in, start, t);
        const double dx = c->cell_size, band = dx * FIELD_BAND_CELLS;
        const double mnx = fmin(t[0], fmin(t[2], t[4])), mxx = fmax(t[0], fmax(t[2], t[4]));
        const double mny = fmin(t[1], fmin(t[3], t[5])), mxy = fmax(t[1], fmax(t[3], t[5]));
        const long x0 = (long)fmax(-1.0, floor((mnx - band) / dx) - 1), x1 = (long)fmin(c->cells_x + 1, ceil((mxx + band) / dx) + 1);
        const long y0 = (long)fmax(-1.0, floor((mny - band) / dx) - 1), y1 = (long)fmin(c->cells_y + 1, ceil((mxy + band) / dx) + 1);
        for (long yi = y0; yi <= y1; ++yi)
            for (long xi = x0; xi <= x1; ++xi) {
                const long key = (yi + 1) * f->W + (xi + 1);
                if (f->body[key] != start) continue;
                const double bw[3] = {f->b0[key], f->b1[key], f->b2[key]};
                for (int v = 0; v < 3; ++v) {
                    imp[v][0] += bw[v] * f->ix[key];
                    imp[v][1] += bw[v] * f->iy[key];
                }
            }
        for (int v = 0; v < 3; ++v) {
            orc_particle* q = &out[start + v];
            const double m = particle_mass(q, c);
            q->vx = q->vx + imp[v][0] / m;
            q->vy = q->vy + imp[v][1] / m;
            if (!finite(q->vx) || !finite(q->vy)) return ORC_ENUMERIC;
        }
    }
    /* PBD edge constraints */
    const double ratio = dt / bc->bond_reference_dt;
    double k = fmin(bc->bond_stiffness * ratio * ratio, 1.0);
    if (!finite(k)) return ORC_ENUMERIC;
    k = fmax(0.0, fmin(1.0, k));
    static const int edges[3][2] = {{0, 1}, {1, 2}, {2, 0}};
    for (int32_t b = 0; b < nb; ++b) {
        const int32_t start = f->starts[b];
        double orig[3][2], pos[3][2];
        for (int v = 0; v < 3; ++v) {
            orig[v][0] = pos[v][0] = out[start + v].x;
            orig[v][1] = pos[v][1] = out[start + v].y;
        }
        for (int it = 0; it < BOND_SOLVER_ITERATIONS; ++it)
            for (int e = 0; e < 3; ++e) {
                const int a = edges[e][0], s2 = edges[e][1];
                const double dx = pos[s2][0] - pos[a][0], dy = pos[s2][1] - pos[a][1];
                const double d = hypot(dx, dy);
                if (!finite(d) || d <= 1e-10) return ORC_ENUMERIC;
                const double sc = ((d - bc->bond_rest_length) * k * 0.5) / d;
                pos[a][0] = pos[a][0] + dx * sc;
                pos[a][1] = pos[a][1] + dy * sc;
                pos[s2][0] = pos[s2][0] - dx * sc;
                pos[s2][1] = pos[s2][1] - dy * sc;
            }
        for (int v = 0; v < 3; ++v) {
            orc_particle* q = &out[start + v];
            const double vx = q->vx + (pos[v][0] - orig[v][0]) / dt, vy = q->vy + (pos[v][1] - orig[v][1]) / dt;
            set_confined(q, pos[v][0], vx, pos[v][1], vy, c);
        }
    }
    if (nb < 2) return ORC_OK;
    /* SAT pair contacts between bodies of different objects */
    const double skin = c->cell_size * CONTACT_SKIN_CELLS;
    double* tri = (double*)malloc(sizeof(double) * 6 * (size_t)nb);
    double* vel = (double*)malloc(sizeof(double) * 2 * (size_t)nb);
    int32_t* ncont = (int32_t*)calloc((size_t)nb, sizeof(int32_t));
    int32_t* other = (int32_t*)malloc(sizeof(int32_t) * (size_t)nb * (size_t)nb);
    double* cn = (double*)malloc(sizeof(double) * 2 * (size_t)nb * (size_t)nb);
    if (!tri || !vel || !ncont || !other || !cn) {
        free(tri); free(vel); free(ncont); free(other); free(cn);
        return ORC_EINVALID;
    }
    for (int32_t b = 0; b < nb; ++b) {
        const int32_t s0 = f->starts[b];
        tri_vertices(in, s0, tri + 6 * b);
        vel[2 * b] = (out[s0].vx + out[s0 + 1].vx + out[s0 + 2].vx) / 3;
        vel[2 * b + 1] = (out[s0].vy + out[s0 + 1].vy + out[s0 + 2].vy) / 3;
    }
    for (int32_t i = 0; i < nb; ++i)
        for (int32_t j = i + 1; j < nb; ++j) {
            if (oid[f->starts[i]] == oid[f->starts[j]]) continue;
            double nx, ny, pen;
            if (!pair_contact(tri + 6 * i, tri + 6 * j, skin, &nx, &ny, &pen)) continue;
            const int32_t ki = ncont[i]++, kj = ncont[j]++;
            other[(size_t)i * nb + ki] = j;
            cn[2 * ((size_t)i * nb + ki)] = nx;
            cn[2 * ((size_t)i * nb + ki) + 1] = ny;
            other[(size_t)j * nb + kj] = i;
            cn[2 * ((size_t)j * nb + kj)] = -nx;
            cn[2 * ((size_t)j * nb + kj) + 1] = -ny;
        }
    const double body_mass = 3 * c->particle_mass;
    for (int32_t b = 0; b < nb; ++b) {
        double ix = 0, iy = 0;
        for (int32_t q = 0; q < ncont[b]; ++q) {
            const int32_t o = other[(size_t)b * nb + q];
            const double nx = cn[2 * ((size_t)b * nb + q)], ny = cn[2 * ((size_t)b * nb + q) + 1];
            const double w = 1.0 / (double)(ncont[b] > ncont[o] ? ncont[b] : ncont[o]);
            const double rn = (vel[2 * b] - vel[2 * o]) * nx + (vel[2 * b + 1] - vel[2 * o + 1]) * ny;
            if (rn < 0) {
                const double imp = -rn / (2 / body_mass);
                const double iw = w * CONTACT_IMPULSE_RELAXATION;
                ix += (iw * imp * nx) / body_mass;
                iy += (iw * imp * ny) / body_mass;
            }
        }
        for (int v = 0; v < 3; ++v) {
            orc_particle* q = &out[f->starts[b] + v];
            set_confined(q, q->x, q->vx + ix, q->y, q->vy + iy, c);
        }
    }
    free(tri); free(vel); free(ncont); free(other); free(cn);
    return ORC_OK;
}

/* updateMpmSolidBondDamage (reference/solid-bonds.ts:68-108): bond damage
 * from the tensile strain of the output positions, then particle damage (the
 * neo-Hookean `param`) = max over incident bonds. */
static int fracture_after_step(orc_particle* out, int32_t n, orc_bond* b, int32_t nb) {
    double* pd = (double*)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (!pd) return ORC_EINVALID;
    for (int32_t i = 0; i < n; ++i) pd[i] = -1;
    for (int32_t k = 0; k < nb; ++k) {
        const orc_particle *p = &out[b[k].first], *q = &out[b[k].second];
        const double d = hypot(q->x - p->x, q->y - p->y);
        if (!finite(d)) { free(pd); return ORC_ENUMERIC; }
        const double strain = fmax(0.0, (d - b[k].rest_length) / b[k].rest_length);
        const int broken = b[k].broken || strain >= b[k].failure_strain;
        const double dmg = broken ? 1.0 : fmax(b[k].damage, fmin(1.0, strain / b[k].failure_strain));
        b[k].damage = dmg;
        b[k].broken = broken;
        pd[b[k].first] = fmax(pd[b[k].first], dmg);
        pd[b[k].second] = fmax(pd[b[k].second], dmg);
    }
    for (int32_t i = 0; i < n; ++i)
        if (pd[i] >= 0 && pd[i] > out[i].param) out[i].param = pd[i];
    free(pd);
    return ORC_OK;
}

int orc_triangle_sample(double px, double py, const double* tri, double* out6) {
    double b[3] = {1, 0, 0};
    const int e = tri_sample(px, py, tri, &out6[0], &out6[1], &out6[2], b);
    out6[3] = b[0];
    out6[4] = b[1];
    out6[5] = b[2];
    return e;
}

int orc_triangle_pair_contact(const double* a, const double* b, double skin, double* out3) {
    return pair_contact(a, b, skin, &out3[0], &out3[1], &out3[2]);
}

int orc_step_bodies_f64(const orc_particle* in, int32_t n, const uint32_t* object_ids, const orc_config* c,
                        const orc_body_config* bc, orc_bond* bonds, int32_t nbonds, orc_particle* out,
                        orc_diagnostics* diag) {
    int32_t* starts = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n / 3 + 1));
    if (!starts) return ORC_EINVALID;
    tri_field f;
    memset(&f, 0, sizeof f);
    int32_t nb = 0;
    int e = validate_bodies(in, n, object_ids, c, bc, starts, &nb);
    if (!e && nb > 0 && (!(bc->bond_stiffness > 0) || bc->bond_stiffness > 1 || !(bc->bond_reference_dt > 0)))
        e = ORC_EINVALID;
    if (!e) e = validate_bonds(in, n, object_ids, bonds, nbonds);
    double* fbuf = NULL;
    if (!e && nb > 0) {
        f.W = c->cells_x + 3;
        f.H = c->cells_y + 3;
        const size_t nn = (size_t)(f.W * f.H);
        fbuf = (double*)malloc(sizeof(double) * nn * 8);
        f.body = (int32_t*)malloc(sizeof(int32_t) * nn);
        if (!fbuf || !f.body) e = ORC_EINVALID;
        else {
            f.dist = fbuf; f.nx = fbuf + nn; f.ny = fbuf + 2 * nn; f.b0 = fbuf + 3 * nn;
            f.b1 = fbuf + 4 * nn; f.b2 = fbuf + 5 * nn; f.ix = fbuf + 6 * nn; f.iy = fbuf + 7 * nn;
            f.starts = starts;
            f.nbodies = nb;
            e = tri_field_build(in, c, &f);
        }
    }
    orc_diagnostics dg;
    if (!e) e = step_core(in, n, c, nb > 0 ? &f : NULL, out, &dg);
    if (!e && nb > 0) e = bodies_after_g2p(in, out, object_ids, c, bc, &f);
    if (!e && nbonds > 0) e = fracture_after_step(out, n, bonds, nbonds);
    free(fbuf);
    free(f.body);
    free(starts);
    if (e) return e;
    output_sums(out, n, c, &dg);
    if (diag) *diag = dg;
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

/* b^e as the GPU's stress cache evaluates it: square-and-multiply in FP64 for
 * a whole exponent in [1, 64], pow() otherwise (the same operation sequence,
 * so the float32 results agree bit for bit). */
static double tait_pow_f64(double b, float e) {
    if (e >= 1.0f && e <= 64.0f && rintf(e) == e) {
        uint32_t n = (uint32_t)e;
        double acc = 1.0, x = b;
        while (n != 0u) {
            if (n & 1u) acc *= x;
            x *= x;
            n >>= 1;
        }
        return acc;
    }
    return pow(b, (double)e);
}
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
            float pw = (float)tait_pow_f64((double)inv, param);
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
