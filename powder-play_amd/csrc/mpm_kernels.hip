// gfx950 kernels of the 2D MLS-MPM microstep.
//
// Particles live in buckets of 4x4 stencil-base cells (double-buffered slot
// arrays, SlotBuf).  One microstep is two launches on one stream:
//   p2g_grid<FIXED, GENERAL>  per 20x20-node tile (28 and 12 also compiled):
//                             stage the particles of the 6x6 buckets whose
//                             stencils reach the tile in LDS (ordered by cell,
//                             then particle index), gather the quadratic
//                             B-spline P2G for a 22x22 node patch, run the grid
//                             update on the 20x20 interior, write node velocities
//   g2p<GENERAL>              per 8x8-bucket tile: stage the 34x34 node patch in
//                             LDS, G2P, F update, projections, CFL guard, and
//                             write each particle into its next bucket
// framed per batch by batch_begin and commit_last + fix_rollback +
// measure_speed.  Numerical semantics follow the WGSL kernels of
// src/lib/mpm/kernels/shaders.ts (cited per function); the fixed-point path
// reproduces their f32 operation order exactly (the library is built with
// -ffp-contract=off) so its int32 accumulators are bit-exact.
#include <float.h>
#include <math.h>

#include <algorithm>
#include <type_traits>

#define MPM_GLOBAL_POINTERS  // scene-descriptor pointers are global in device code
#include "mpm_internal.h"
#include "mpm_timing.h"

namespace mpm {
namespace {

constexpr float FIXED_POINT_SCALE = 16777216.0f;   // shaders.ts:173
constexpr float FIXED_POINT_LIMIT = 2147483520.0f; // shaders.ts:974
constexpr float WALL_CUSHION_CELLS = 8.0f;         // shaders.ts:117
constexpr float LIQUID_CONTACT_SEPARATION_FRACTION = 0.25f;  // materials.ts:40-43
constexpr float LIQUID_CONTACT_APPROACH_SPEED_MULTIPLIER = 1.0f;
constexpr float LIQUID_CONTACT_RELAXATION_SECONDS = 0.05f;
constexpr float LIQUID_STABLE_ORDER_EPSILON = 1e-3f;
constexpr double SNOW_COMPRESSION_YIELD = 0.025;   // materials.ts:46-53
constexpr double SNOW_STRETCH_YIELD = 0.0075;
constexpr double SNOW_HARDENING = 10.0;
constexpr double SNOW_MIN_PLASTIC_VOLUME = 0.6;
constexpr double SNOW_MAX_PLASTIC_VOLUME = 20.0;
constexpr double SNOW_MAX_HARDENING_EXPONENT = 20.0;
constexpr float MIN_LIQUID_DETERMINANT = 0.5f;     // shaders.ts:202-210
constexpr float MAX_LIQUID_DETERMINANT = 2.0f;
constexpr double MIN_POWDER_DETERMINANT = 0.5;
constexpr double MIN_NEO_HOOKEAN_STRETCH = 0.05;
constexpr double MAX_NEO_HOOKEAN_STRETCH = 20.0;
constexpr float MAX_DEFORMATION_INCREMENT = 0.25f;  // a power of two (dividing by it is exact scaling)
static_assert(MAX_DEFORMATION_INCREMENT == 0.25f, "the substep count scales by the exact reciprocal");
constexpr uint32_t MAX_DEFORMATION_SUBSTEPS = 64u;
// bonded triangle bodies (shaders.ts:175-177, materials.ts:44, reference/solver.ts:277-282)
constexpr float TRIANGLE_FIELD_BAND_CELLS = 2.0f;
constexpr float TRIANGLE_CONTACT_SKIN_CELLS = 0.1f;
constexpr float TRIANGLE_CONTACT_IMPULSE_RELAXATION = 0.25f;
constexpr float TRIANGLE_FLUID_TANGENTIAL_DRAG_RATE = 4.0f;
constexpr int BONDED_TRIANGLE_SOLVER_ITERATIONS = 2;

// P2G tile of TI x TI interior nodes and a one-node ring for the
// liquid-contact volume gradient ((TI+2)^2 nodes computed).  A node receives
// from the particles whose stencil-base cell c (= floor(x/dx - 0.5) + 1, the
// cell of the bucket keys) lies in [node - 2, node], so the tile and its ring
// need TI + 4 cells a side.  Tile t starts one node before a bucket boundary
// (interior nodes [TI t - 1, TI t + TI - 2]), which makes those cells exactly
// the TI/4 + 1 buckets [TI/4 t - 1, TI/4 t + TI/4 - 1]: no particle of the
// region lies outside every stencil of the tile.  Three sides are compiled:
//   28: 1024 threads, two tiles a CU (LDS), 32 node threads to a row
//   20:  512 threads, three tiles a CU
//   12:  256 threads, six tiles a CU
// Smaller tiles re-read more region buckets (each bucket of a region is
// staged by one to four tiles) but keep more blocks in flight per CU.
template <int TI_>
struct P2GTile {
    static constexpr int TI = TI_;                     // interior nodes per tile side
    static constexpr int TC = TI + 2;                  // with the one-node ring
    static constexpr int OWN = TI / BUCKET_CELLS;      // buckets a tile owns per side
    static constexpr int RB = OWN + 1;                 // region buckets per side
    static constexpr int NB = RB * RB;
    static constexpr int RC = RB * BUCKET_CELLS;       // region cells per side
    static constexpr int CELLS = RC * RC;
    static constexpr int THREADS = TI == 28 ? 1024 : (TI == 20 ? 512 : 256);
    static constexpr int CAP = (int)p2g_chunk_for(TI);  // staged particles per gather chunk
    static constexpr int ROW = (TC <= 32 && 32 * TC <= THREADS) ? 32 : TC;  // node threads per tile row
    // waves a SIMD keeps of the gather: its LDS admits 2 x 16, 3 x 8 or 6 x 4
    // waves a CU, so the smaller tiles may use 80 registers instead of 64
    static constexpr int GATHER_WAVES_PER_EU = TI == 28 ? 8 : 6;
    static_assert(TI == 28 || TI == 20 || TI == 12, "compiled tile sides");
    static_assert(TI % BUCKET_CELLS == 0, "tiles hold whole buckets");
    static_assert(CAP <= 2 * THREADS, "the gather loads at most two particles a thread");
    static_assert(TC * TC <= CAP, "the ring volumes reuse two staging arrays");
    static_assert(ROW * TC <= THREADS, "one node thread per tile node");
    static_assert(CELLS <= 2 * THREADS, "the cell scan takes two cells a thread");
    static_assert(NB <= 128, "the bucket scan takes two buckets a lane of one wave");
    static_assert(OWN * OWN <= THREADS, "one thread clears each owned bucket count");
};
constexpr int LDS_MATS = 64;                    // material rows cached in LDS
// G2P tile: TB x TB buckets (4TB x 4TB cells) and their (4TB+2)^2 node patch.
template <int TB_>
struct G2PTile {
    static constexpr int TB = TB_;
    static constexpr int TC = TB * BUCKET_CELLS;
    static constexpr int TN = TC + 2;
    static constexpr int NBK = TB * TB;  // buckets of the tile
    static_assert(NBK <= 64, "one wave scans the bucket counts");
};

__device__ __forceinline__ bool finitef(float v) { return (__float_as_uint(v) & 0x7f800000u) != 0x7f800000u; }
__device__ __forceinline__ bool finite2(float2 v) { return finitef(v.x) && finitef(v.y); }
__device__ __forceinline__ bool finite4(float4 v) { return finitef(v.x) && finitef(v.y) && finitef(v.z) && finitef(v.w); }
__device__ __forceinline__ uint32_t load_status(const Control* c) {
    return __atomic_load_n(&c->status[0], __ATOMIC_RELAXED);
}
// Status read shared by the whole block, for kernels that synchronise: every
// thread takes the same early exit, so no barrier waits on exited waves.
__device__ __forceinline__ bool block_status_set(const Control* c) {
    __shared__ uint32_t s_status;
    if (threadIdx.x == 0) s_status = load_status(c);
    __syncthreads();
    return s_status != 0u;
}
// determinant2x2 (shaders.ts:261-266): keeps the residual of the second
// product when the two terms nearly cancel.
__device__ __forceinline__ float det_comp(float4 m) {
    float sp = m.y * m.z;
    float spe = fmaf(m.y, m.z, -sp);
    return fmaf(m.x, m.w, -sp) - spe;
}
// Exact-product FP64 determinant used for the deformation-failure tests.
__device__ __forceinline__ double det_f64(float4 m) {
    return (double)m.x * (double)m.w - (double)m.y * (double)m.z;
}

// recordDeformationFailure / recordProjectedDeformationFailure (shaders.ts:224-248)
__device__ void record_failure(Control* c, uint32_t kind, uint32_t index, float a, float b, float d, float e) {
    uint32_t prev = atomicOr(&c->status[6], kind);
    if (prev == 0u) {
        atomicExch(&c->status[1], index);
        atomicExch(&c->status[2], __float_as_uint(a));
        atomicExch(&c->status[3], __float_as_uint(b));
        atomicExch(&c->status[4], __float_as_uint(d));
        atomicExch(&c->status[5], __float_as_uint(e));
    }
}

// ---------------------------------------------------------------- FP64 2x2
struct M2 { double a, b, c, d; };  // row-major m00 m01 m10 m11
struct Decomp { double s1, s2, u1x, u1y, u2x, u2y, v1x, v1y, v2x, v2y; M2 rot; };

__device__ __forceinline__ double det2(M2 m) { return m.a * m.d - m.b * m.c; }

// Closed-form 2x2 SVD through the eigen-angle of F^T F (decompose2x2,
// reference/solver.ts:2425-2463), evaluated in FP64.
__device__ bool decompose(M2 F, Decomp& d) {
    double c00 = F.a * F.a + F.c * F.c;
    double c01 = F.a * F.b + F.c * F.d;
    double c11 = F.b * F.b + F.d * F.d;
    double radius = hypot(c00 - c11, 2 * c01);
    double s1 = sqrt(fmax(0.0, (c00 + c11 + radius) * 0.5));
    double s2 = det2(F) / s1;
    if (!isfinite(s1) || !isfinite(s2) || s1 <= 0 || s2 <= 0) return false;
    double angle = 0.5 * atan2(2 * c01, c00 - c11);
    double co = cos(angle), si = sin(angle);
    d.s1 = s1; d.s2 = s2;
    d.v1x = co; d.v1y = si; d.v2x = -si; d.v2y = co;
    d.u1x = (F.a * d.v1x + F.b * d.v1y) / s1;
    d.u1y = (F.c * d.v1x + F.d * d.v1y) / s1;
    d.u2x = -d.u1y; d.u2y = d.u1x;
    d.rot.a = d.u1x * d.v1x + d.u2x * d.v2x;
    d.rot.b = d.u1x * d.v1y + d.u2x * d.v2y;
    d.rot.c = d.u1y * d.v1x + d.u2y * d.v2x;
    d.rot.d = d.u1y * d.v1y + d.u2y * d.v2y;
    return isfinite(d.rot.a) && isfinite(d.rot.b) && isfinite(d.rot.c) && isfinite(d.rot.d);
}
__device__ __forceinline__ M2 reconstruct(const Decomp& d, double s1, double s2) {
    return {s1 * d.u1x * d.v1x + s2 * d.u2x * d.v2x, s1 * d.u1x * d.v1y + s2 * d.u2x * d.v2y,
            s1 * d.u1y * d.v1x + s2 * d.u2y * d.v2x, s1 * d.u1y * d.v1y + s2 * d.u2y * d.v2y};
}
__device__ __forceinline__ bool finite_m2(M2 m) {
    return isfinite(m.a) && isfinite(m.b) && isfinite(m.c) && isfinite(m.d);
}

// Solid P*F^T in FP64 from the f32 deformation, rounded once to f32
// (constitutiveStressTimesFt, reference/solver.ts:2148-2265; GPU form
// shaders.ts:825-936).  Returns false when the stress is not finite.
__device__ bool solid_stress(uint32_t model, float4 Ff, float lambda, float mu, float param, float pv, float4& S) {
    M2 F = {Ff.x, Ff.y, Ff.z, Ff.w};
    M2 r;
    if (model == MODEL_SNOW) {
        Decomp d;
        if (!decompose(F, d)) return false;
        double hexp = fmax(-SNOW_MAX_HARDENING_EXPONENT,
                           fmin(SNOW_MAX_HARDENING_EXPONENT, SNOW_HARDENING * (1 - (double)pv)));
        double h = exp(hexp);
        double l = (double)lambda * h, m = (double)mu * h;
        double J = det2(F);
        double shear = 2 * m, vol = l * J * (J - 1);
        double d00 = F.a - d.rot.a, d01 = F.b - d.rot.b, d10 = F.c - d.rot.c, d11 = F.d - d.rot.d;
        r = {shear * (d00 * F.a + d01 * F.b) + vol, shear * (d00 * F.c + d01 * F.d),
             shear * (d10 * F.a + d11 * F.b), shear * (d10 * F.c + d11 * F.d) + vol};
    } else if (model == MODEL_POWDER) {
        Decomp d;
        if (!decompose(F, d)) return false;
        double l1 = log(d.s1), l2 = log(d.s2), lv = l1 + l2;
        double blv = fmax(log(MIN_POWDER_DETERMINANT), lv);
        double pressure = fmax(0.0, -((double)lambda + (double)mu) * blv);
        double dev = 0.5 * (l1 - l2);
        double maxdev = ((double)param * pressure) / (2 * (double)mu);
        double pdev = fmin(dev, maxdev);
        double mean = 0.5 * lv;
        double ps1 = 2 * (double)mu * (mean + pdev) + (double)lambda * lv;
        double ps2 = 2 * (double)mu * (mean - pdev) + (double)lambda * lv;
        r = {ps1 * d.u1x * d.u1x + ps2 * d.u2x * d.u2x, ps1 * d.u1x * d.u1y + ps2 * d.u2x * d.u2y,
             ps1 * d.u1y * d.u1x + ps2 * d.u2y * d.u2x, ps1 * d.u1y * d.u1y + ps2 * d.u2y * d.u2y};
    } else {  // Neo-Hookean with damage in [0, 1]
        double J = det2(F);
        M2 it = {F.d / J, -F.c / J, -F.b / J, F.a / J};
        double pressure = (double)lambda * log(J);
        M2 Pk = {(double)mu * (F.a - it.a) + pressure * it.a, (double)mu * (F.b - it.b) + pressure * it.b,
                 (double)mu * (F.c - it.c) + pressure * it.c, (double)mu * (F.d - it.d) + pressure * it.d};
        if (!finite_m2(Pk)) return false;
        r = {Pk.a * F.a + Pk.b * F.b, Pk.a * F.c + Pk.b * F.d, Pk.c * F.a + Pk.d * F.b, Pk.c * F.c + Pk.d * F.d};
        double dmg = (double)param < 0 ? 0.0 : ((double)param > 1 ? 1.0 : (double)param);
        double scale = fmax(0.0, 1.0 - dmg);
        r.a *= scale; r.b *= scale; r.c *= scale; r.d *= scale;
    }
    if (!finite_m2(r)) return false;
    S = make_float4((float)r.a, (float)r.b, (float)r.c, (float)r.d);
    return true;
}

// Stress P*F^T of prepareParticle (shaders.ts:808-963) with its material
// checks.  Liquids in f32 (tension-free linear EOS; Tait via FP64 pow);
// solids in FP64, rounded once.  GENERAL = false admits linear liquids only.
// Returns the status bits to raise.
// b^e in FP64 for the Tait power J^-n.  A whole exponent up to 64 (the Tait
// exponent is 7 in practice) takes square-and-multiply: a handful of FP64
// products, each within an ulp, so the result is within a few FP64 ulps of
// the exact power, far below the float32 rounding that follows.  Other
// exponents take pow().  The oracle's float32 restatement of the stress
// cache evaluates the same sequence.
__device__ __forceinline__ double tait_pow(double b, float e) {
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

// SOLIDS = false compiles out the solid models (scenes whose flags list no
// solid): fewer registers for the Tait-liquid kernels.
template <bool GENERAL, bool SOLIDS = true>
__device__ uint32_t particle_stress(const Material& mt, float4 F, float pv, float det, float4& S) {
    S = make_float4(0.f, 0.f, 0.f, 0.f);
    if (!finitef(det) || det <= 0.0f || !finitef(mt.volume) || mt.volume <= 0.0f) return ERR_DEFORMATION;
    const uint32_t model = mt.model;
    if (model == MODEL_WATER || model == MODEL_OIL) {
        if (!finitef(mt.density) || mt.density <= 0.0f || !finitef(mt.lambda) || mt.lambda <= 0.0f ||
            !finitef(mt.param) || mt.param < 0.0f)
            return ERR_P2G | ERR_INVALID_MATERIAL;
        if (mt.param > 0.0f) {  // Tait: -J * B * (J^-n - 1)
            if constexpr (!GENERAL) {
                return ERR_P2G | ERR_INVALID_MATERIAL;
            } else {
                const float inv = 1.0f / det;
                const float pw = (float)tait_pow((double)inv, mt.param);
                const float pressure = mt.lambda * (pw - 1.0f);
                const float stress = -det * pressure;
                S.x = stress; S.w = stress;
            }
        } else {  // tension-free linear: lambda * min(J - 1, 0)
            const float jm1 = det - 1.0f;
            const float pressure = mt.lambda * (jm1 < 0.0f ? jm1 : 0.0f);
            S.x = pressure; S.w = pressure;
        }
    } else {
        if constexpr (!GENERAL || !SOLIDS) {
            return ERR_P2G | ERR_INVALID_MATERIAL;
        } else {
            if (model == MODEL_NEO_HOOKEAN) {
                if (!finitef(mt.param) || mt.param < 0.0f || mt.param > 1.0f) return ERR_P2G | ERR_INVALID_MATERIAL;
                if (!solid_stress(model, F, mt.lambda, mt.mu, mt.param, pv, S)) return ERR_P2G | ERR_NONFINITE_PARTICLE;
            } else if (model == MODEL_SNOW) {
                if (!finitef(pv) || pv < 0.6f || pv > 20.0f) return ERR_P2G | ERR_INVALID_MATERIAL;
                if (!solid_stress(model, F, mt.lambda, mt.mu, mt.param, pv, S)) return ERR_P2G | ERR_INVALID_MATERIAL;
            } else if (model == MODEL_POWDER) {
                if (!finitef(mt.lambda) || !finitef(mt.mu) || mt.mu <= 0.0f || !finitef(mt.param) || mt.param < 0.0f)
                    return ERR_P2G | ERR_INVALID_MATERIAL;
                if (!solid_stress(model, F, mt.lambda, mt.mu, mt.param, pv, S)) return ERR_P2G | ERR_NONFINITE_PARTICLE;
            } else {
                return ERR_P2G | ERR_INVALID_MATERIAL;
            }
        }
    }
    if (!finite4(S)) return ERR_P2G | ERR_NONFINITE_PARTICLE;
    return 0;
}

// A = m*C + (-dt*V*4/dx^2) * P*F^T (shaders.ts:954-958).
__device__ __forceinline__ float4 affine_of(const Params& P, const Material& mt, float4 C, float4 S) {
    const float ss = (-P.dt * mt.volume) * P.grad_scale;
    return make_float4(mt.mass * C.x + ss * S.x, mt.mass * C.y + ss * S.y, mt.mass * C.z + ss * S.z,
                       mt.mass * C.w + ss * S.w);
}

// Stencil base and fraction (shaders.ts:667-676, :1023-1028).  inv_dx is
// exact_reciprocal(dx): for a power-of-two cell size the scaling is one
// exact multiply instead of a division, with the same result.
struct Stencil { int bx, by; float fx, fy; };
__device__ __forceinline__ bool make_stencil(float dx, float inv_dx, uint32_t cells_x, uint32_t cells_y, float2 p,
                                             Stencil& s) {
    const float gx = inv_dx != 0.0f ? p.x * inv_dx : p.x / dx, gy = inv_dx != 0.0f ? p.y * inv_dx : p.y / dx;
    // the tests combined with bitwise operators, so that they form no
    // divergent branch; the stencil is meaningful only when this returns
    // true (a non-finite coordinate is replaced before the conversion)
    const bool finite_x = finitef(gx), finite_y = finitef(gy);
    const bool finite = finite_x & finite_y;
    const float cx = finite ? gx : 0.0f, cy = finite ? gy : 0.0f;
    s.bx = (int)floorf(cx - 0.5f);
    s.by = (int)floorf(cy - 0.5f);
    s.fx = cx - (float)s.bx;
    s.fy = cy - (float)s.by;
    // stencilFitsHalo (shaders.ts:511-516)
    return finite & (s.bx >= -1) & (s.by >= -1) & (s.bx + 2 <= (int)cells_x + 1) & (s.by + 2 <= (int)cells_y + 1);
}
// quadraticWeights (shaders.ts:500-509), one component.
__device__ __forceinline__ float qweight(float f, int o) {
    float l = 1.5f - f, c = f - 1.0f, r = f - 0.5f;
    float w0 = (0.5f * l) * l, w1 = 0.75f - c * c, w2 = (0.5f * r) * r;
    return o == 0 ? w0 : (o == 1 ? w1 : w2);
}

// The quadratic B-spline at d = o - f, o the stencil offset and f in
// [0.5, 1.5) the fraction: 0.75 - d^2 at the middle offset, else
// (1.5 - |d|)^2 / 2.  For o = 0 and 1 this is bit for bit quadraticWeights
// (shaders.ts:500-509); for o = 2, 1.5 - |2 - f| equals its f - 0.5 but for
// the rounding of 2 - f when f < 1 (at most 2^-24 in t, below 2^-25 in the
// weight).  d is the reference's dpos / dx, so the gather forms it once.
// Both branches are computed and one is kept by a bit mask, so lanes of a
// wave at different offsets do not diverge.
__device__ __forceinline__ float bspline_of_offset(float d, bool middle) {
    const uint32_t m = 0u - (uint32_t)middle;
    const float t = 1.5f - fabsf(d);
    const float mid = 0.75f - d * d, side = (0.5f * t) * t;
    return __uint_as_float((__float_as_uint(mid) & m) | (__float_as_uint(side) & ~m));
}

// Whether the momentum m v + (A dx) d of a staged particle is finite at all
// nine stencil offsets, d = offset - fraction per axis, as the gather forms
// it: fmaf(a01, dy, fmaf(a00, dx, m)).  |d| <= 1.5, so finite inputs whose
// magnitudes bound every sum below 1e37 pass at once.  Otherwise the four
// corners decide: with the fraction in [0.5, 1.5) no d of the outer offsets
// 0 and 2 is zero, so a non-finite input makes every corner sum non-finite,
// and with finite inputs each rounded fmaf is monotone in d and in its
// addend, so the middle sums lie between corner sums.
__device__ __forceinline__ bool contributions_finite(float fx, float fy, float mx, float my, float4 Ad) {
    const float bx = fabsf(mx) + 2.0f * (fabsf(Ad.x) + fabsf(Ad.y));
    const float by = fabsf(my) + 2.0f * (fabsf(Ad.z) + fabsf(Ad.w));
    if (bx < 1e37f && by < 1e37f) return true;  // false for a NaN bound
    const float dx0 = 0.0f - fx, dx2 = 2.0f - fx, dy0 = 0.0f - fy, dy2 = 2.0f - fy;
    auto fin2 = [&](float ddx, float ddy) {
        return finitef(fmaf(Ad.y, ddy, fmaf(Ad.x, ddx, mx))) && finitef(fmaf(Ad.w, ddy, fmaf(Ad.z, ddx, my)));
    };
    return fin2(dx0, dy0) && fin2(dx2, dy0) && fin2(dx0, dy2) && fin2(dx2, dy2);
}

// Raise status bits from microstep `step` of the batch; the batch commits
// the microsteps before the first one that raised (commitParticles,
// shaders.ts:3067-3080).
__device__ __forceinline__ void raise_step(Control* c, uint32_t bits, uint32_t step) {
    atomicMin(&c->fail_step, step);
    atomicOr(&c->status[0], bits);
}
__device__ __forceinline__ void raise_step_at(Control* c, uint32_t bits, uint32_t step, uint32_t index) {
    atomicMin(&c->status[1], index);
    raise_step(c, bits, step);
}

// Slot in bucket `b` for each active lane.  Lanes are grouped by bucket
// with ballots first; then every group's first lane reserves the group's
// slots with one atomic, all groups in the same instruction, so the wave
// waits for a single round trip.  Every lane of the wave must call it
// (converged control flow).
// The reservation in two halves: wave_alloc_issue groups the lanes and
// issues the atomics; wave_alloc_finish hands each lane its slot once the
// atomics have returned, so independent work can run between the two.
struct WaveAlloc {
    uint32_t base;  // the group leader's atomic result (valid in the leader)
    int leader;
    unsigned long long group;
};
__device__ __forceinline__ WaveAlloc wave_alloc_issue(uint32_t* cnt, uint32_t b, bool active) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(active);
    WaveAlloc a{0u, lane, 0ull};
    while (pending) {
        const int first = __ffsll((long long)pending) - 1;
        // the lane is uniform (from a ballot): a lane read, not a permute
        // through LDS (which __shfl is), on every particle's chain
        const uint32_t fb = (uint32_t)__builtin_amdgcn_readlane((int)b, first);
        const unsigned long long same = __ballot(active && b == fb);
        if (active && b == fb) {
            a.group = same;
            a.leader = first;
        }
        pending &= ~same;
    }
    if (active && lane == a.leader) a.base = atomicAdd(&cnt[b], (uint32_t)__popcll(a.group));
    return a;
}
__device__ __forceinline__ uint32_t wave_alloc_finish(const WaveAlloc& a, bool active) {
    const int lane = threadIdx.x & 63;
    const uint32_t base = __shfl(a.base, a.leader, 64);
    return active ? base + (uint32_t)__popcll(a.group & ((1ull << lane) - 1ull)) : 0xffffffffu;
}
__device__ __forceinline__ uint32_t wave_alloc(uint32_t* cnt, uint32_t b, bool active) {
    return wave_alloc_finish(wave_alloc_issue(cnt, b, active), active);
}

__device__ __forceinline__ uint32_t bucket_of(const Stencil& s, uint32_t bw) {
    return (uint32_t)((s.by + 1) >> 2) * bw + (uint32_t)((s.bx + 1) >> 2);
}

// Tile of this block.  Workgroups are dispatched round-robin over the eight
// XCDs, each with its own L2; handing every XCD a contiguous band of tiles
// (in tile order) keeps neighbouring tiles, which re-read each other's halo
// particles and nodes, on the same L2.
constexpr uint32_t XCDS = QUEUE_XCDS;
__device__ __forceinline__ uint32_t xcd_tile(uint32_t tiles) {
    const uint32_t b = blockIdx.x, x = b % XCDS, k = b / XCDS;
    const uint32_t per = tiles / XCDS, extra = tiles % XCDS;
    return x * per + min(x, extra) + k;
}

// ------------------------------------------------------------ batch frame
__global__ void k_batch_begin(const SceneDesc* __restrict__ scenes) {
    Control* ctl = scenes[blockIdx.x].ctl;  // one block per scene
    if (threadIdx.x == 0) {
        ctl->completed = 0u;
        ctl->max_speed_bits = 0u;
        // a flag already raised stops the batch before its first microstep
        ctl->fail_step = ctl->status[0] != 0u ? 0u : 0xffffffffu;
    }
}

__global__ void k_clear_counts(uint32_t* cnt, uint32_t nb) {
    for (uint32_t k = blockIdx.x * blockDim.x + threadIdx.x; k < nb; k += gridDim.x * blockDim.x) cnt[k] = 0u;
}

// Bucket particles [first, first+count) of the caller-order arrays
// (insertParticleBucket, shaders.ts:657-693, as slot allocation) and fill
// the stress cache of scenes that carry it.
template <bool GENERAL>
__global__ __launch_bounds__(256) void k_bucket(Control* ctl, Home home, const Material* __restrict__ mats,
                                                SlotBuf out, uint32_t first, uint32_t count, float dx,
                                                uint32_t cells_x, uint32_t cells_y, uint32_t bw, uint32_t cap,
                                                AuxData aux, int buf) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t i = first + t;
    bool active = t < count;
    const uint32_t model = active ? home.model[i] : MODEL_REMOVED;
    active = active && model != MODEL_REMOVED;
    Stencil s = {0, 0, 0.0f, 0.0f};
    float2 p = make_float2(0.f, 0.f);
    if (active) {
        p = home.pos[i];
        if (!make_stencil(dx, exact_reciprocal(dx), cells_x, cells_y, p, s)) {
            raise_step(ctl, ERR_STENCIL, 0u);
            active = false;
        }
    }
    const uint32_t b = active ? bucket_of(s, bw) : 0u;
    const uint32_t k = wave_alloc(out.cnt, b, active);
    if (!active) return;
    if (k >= cap) {
        raise_step(ctl, ERR_INTERNAL_REBUCKET, 0u);
        return;
    }
    const size_t g = (size_t)b * cap + k;
    const float4 F = home.F[i];
    const uint32_t m = home.mat[i];
    out.pos[g] = p;
    out.vel[g] = home.vel[i];
    out.C[g] = home.C[i];
    out.F[g] = F;
    out.pid[g] = i;
    out.mat[g] = m;
    Material mt = mats[m];
    if (aux.aux_of) {  // bodies / fracture: where the particle sits, and its current damage
        const uint32_t a = aux.aux_of[i];
        if (a != NO_AUX) {
            aux.slot[buf][a] = (uint32_t)g;
            if (a >= aux.frac_aux0) mt.param = aux.damage[buf][a];
        }
    }
    if constexpr (GENERAL) {
        const float pv = home.plastic[i];
        out.plastic[g] = pv;
        float4 S;
        const uint32_t pf = particle_stress<true>(mt, F, pv, det_comp(F), S);
        out.S[g] = S;
        if (pf) {
            if (pf & ERR_P2G) raise_step_at(ctl, pf, 0u, i); else raise_step(ctl, pf, 0u);
        }
    }
}

// Buffer -> caller-order arrays (readParticles, solver.ts:2594-2731).
__global__ __launch_bounds__(256) void k_unbucket(SlotBuf in, Home home, uint32_t nb, uint32_t cap,
                                                  uint32_t general) {
    const size_t total = (size_t)nb * cap;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const uint32_t b = (uint32_t)g / cap, k = (uint32_t)g - b * cap;  // slot counts fit 32 bits
        if (k >= in.cnt[b]) continue;
        const uint32_t i = in.pid[g];
        home.pos[i] = in.pos[g];
        home.vel[i] = in.vel[g];
        home.C[i] = in.C[g];
        home.F[i] = in.F[g];
        if (general) home.plastic[i] = in.plastic[g];
    }
}

__global__ __launch_bounds__(256) void k_keys(Home home, uint32_t* keys, uint32_t n, float dx, uint32_t cells_x,
                                              uint32_t cells_y, uint32_t node_w) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    Stencil s;
    if (home.model[i] == MODEL_REMOVED || !make_stencil(dx, exact_reciprocal(dx), cells_x, cells_y, home.pos[i], s)) {
        keys[i] = INVALID_KEY;
        return;
    }
    keys[i] = (uint32_t)(s.by + 1) * node_w + (uint32_t)(s.bx + 1);
}

// ------------------------------------------------------------ P2G + grid
// a / b for finite non-zero b in the grid update: the hardware reciprocal,
// then one correction of the quotient (Markstein), within an ulp of the
// IEEE quotient in four instructions instead of the ten of the correctly
// rounded sequence.  The WGSL quotient it stands for carries 2.5 ulp.  The
// binning and the fixed-point quotients keep their exact forms.
__device__ __forceinline__ float quot(float a, float b) {
    const float r = __builtin_amdgcn_rcpf(b);
    const float q = a * r;
    return fmaf(fmaf(-q, b, a), r, q);
}

// finishGridVelocity (shaders.ts:1748-1777)
__device__ __forceinline__ float2 finish_velocity(const Params& P, float m, float qx, float qy, int nx, int ny) {
    if (m <= 0.0f) return make_float2(0.f, 0.f);
    float2 v = make_float2(quot(qx, m) + P.dt * P.gx, quot(qy, m) + P.dt * P.gy);
    if (P.drag_radius > 0.0f && nx >= 0 && ny >= 0 && nx <= (int)P.cells_x && ny <= (int)P.cells_y) {
        float px = (float)nx * P.dx, py = (float)ny * P.dx;
        float dxv = P.drag_target_x - px, dyv = P.drag_target_y - py;
        float dist = sqrtf(dxv * dxv + dyv * dyv);
        if (dist < P.drag_radius) {
            float falloff = 1.0f - dist / P.drag_radius;
            float ax = dxv * P.drag_stiffness - v.x * P.drag_damping;
            float ay = dyv * P.drag_stiffness - v.y * P.drag_damping;
            float s = P.dt * falloff;
            v.x = v.x + s * ax;
            v.y = v.y + s * ay;
        }
    }
    return v;
}
// applyWallCushion / applyWallTangentialFriction (shaders.ts:1438-1457).
// At dist >= cushion the clamped ratio is exactly 1 (division is monotone
// and cushion / cushion = 1), so the factor is exact without the division.
__device__ __forceinline__ float wall_cushion(float v, int coord, int extent) {
    float cushion = fminf(WALL_CUSHION_CELLS, (float)extent / 4.0f);
    float dist = v < 0.0f ? (float)coord : (float)(extent - coord);
    if (dist >= cushion) return v;
    return v * fminf(fmaxf(dist / cushion, 0.0f), 1.0f);
}
__device__ __forceinline__ float wall_friction(float v, int coord, int extent, float friction) {
    if (friction <= 0.0f) return v;
    float dist = fmaxf(0.0f, fminf((float)coord, (float)(extent - coord)));
    float cushion = fminf(WALL_CUSHION_CELLS, (float)extent / 4.0f);
    if (dist >= cushion) return v;
    float infl = 1.0f - fminf(fmaxf(dist / cushion, 0.0f), 1.0f);
    return v * (1.0f - friction * infl);
}

// a / b for the fixed-point contributions (b = particleMass or
// particleVolume).  With r = RN(1/b) and b in [2^-20, 2^20] (the host passes
// r = 0 otherwise) Markstein's correction q + fma(-q, b, a) * r returns the
// IEEE quotient for |a| in [2^-60, 2^100]; outside that range the quotient
// rounds to the same fixed-point word or raises the same overflow flag.
__device__ __forceinline__ float div_ref(float a, float b, float r) {
    if (r == 0.0f) return a / b;
    const float q = a * r;
    const float e = fmaf(-q, b, a);
    return fmaf(e, r, q);
}

__device__ __forceinline__ uint32_t fixed_round(float value, int32_t& out) {
    const float scaled = rintf(value * FIXED_POINT_SCALE);  // addFixedPoint (shaders.ts:971-985)
    // as selects: a non-finite or out-of-range value gives 0 and the flag
    const bool ok = fabsf(scaled) <= FIXED_POINT_LIMIT;  // false for NaN and infinities
    out = (int32_t)(ok ? scaled : 0.0f);
    return ok ? 0u : ERR_FIXED_POINT;
}

// Copy the first LDS_MATS material rows into LDS (all threads of the block;
// a barrier must follow before material_at is used).
__device__ __forceinline__ void stage_materials(Material* s_mat, const Material* mats, uint32_t nmat) {
    const uint32_t words = min(nmat, (uint32_t)LDS_MATS) * (uint32_t)(sizeof(Material) / 4);
    for (uint32_t e = threadIdx.x; e < words; e += blockDim.x)
        reinterpret_cast<uint32_t*>(s_mat)[e] = reinterpret_cast<const uint32_t*>(mats)[e];
}
__device__ __forceinline__ Material material_at(const Material* s_mat, const Material* mats, uint32_t m) {
    uint4 lo, hi;  // the row as two 16-byte words, from LDS or from global memory
    if (m < (uint32_t)LDS_MATS) {
        lo = reinterpret_cast<const uint4*>(s_mat)[2 * m];
        hi = reinterpret_cast<const uint4*>(s_mat)[2 * m + 1];
    } else {  // a different load form, so the two branches are not merged into one flat load
        typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
        const u32x4 a = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(mats) + 2 * m);
        const u32x4 b = __builtin_nontemporal_load(reinterpret_cast<const u32x4*>(mats) + 2 * m + 1);
        lo = make_uint4(a.x, a.y, a.z, a.w);
        hi = make_uint4(b.x, b.y, b.z, b.w);
    }
    Material r;
    r.model = lo.x;
    r.lambda = __uint_as_float(lo.y);
    r.mu = __uint_as_float(lo.z);
    r.density = __uint_as_float(lo.w);
    r.volume = __uint_as_float(hi.x);
    r.param = __uint_as_float(hi.y);
    r.mass = __uint_as_float(hi.z);
    r.phase = hi.w;
    return r;
}

// Block sum of three doubles in a fixed order (wave shuffles, then the
// waves in index order); thread 0 writes them to out[0..2].  Every thread of
// the block calls it.
__device__ void block_sum3_f64(double a, double b, double c, double* out) {
    __shared__ double s_part[3][16];  // blocks of up to 1024 threads
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_down(a, off, 64);
        b += __shfl_down(b, off, 64);
        c += __shfl_down(c, off, 64);
    }
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_part[0][wave] = a;
        s_part[1][wave] = b;
        s_part[2][wave] = c;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double ta = 0.0, tb = 0.0, tc = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
            ta += s_part[0][w];
            tb += s_part[1][w];
            tc += s_part[2][w];
        }
        out[0] = ta;
        out[1] = tb;
        out[2] = tc;
    }
}

// updateGrid (shaders.ts:1485-1746) of one interior node from its summed
// channels f = {solid m, qx, qy | water m, qx, qy, V | oil m, qx, qy, V}.
// `nvol(k, ox, oy)` returns the summed water (k = 0) or oil (k = 1) volume of
// the neighbour at offset (ox, oy).  Returns the status bits to raise.
// No fluid-triangle contact (scenes without bonded bodies).
struct NoTriangles {
    __device__ uint32_t operator()(uint32_t, float, float, float2&, float2&) const { return 0u; }
};

template <typename NVol, typename Tri = NoTriangles>
__device__ __forceinline__ uint32_t update_node(const Params& P, const Grid& grid, uint32_t scene_flags,
                                                uint32_t node, int nx, int ny, const float (&f)[11],
                                                uint32_t nonfinite_bits, NVol nvol, Tri tri = Tri()) {
    const float f_sm = f[0], f_sqx = f[1], f_sqy = f[2], f_wm = f[3], f_wqx = f[4], f_wqy = f[5], f_wv = f[6];
    const float f_om = f[7], f_oqx = f[8], f_oqy = f[9], f_ov = f[10];
    uint32_t flags = 0;
    bool finite = true;
#pragma unroll
    for (int k = 0; k < 11; ++k) finite = finite && finitef(f[k]);
    if (!finite) return nonfinite_bits;
    if (f_sm < 0.0f || f_wm < 0.0f || f_wv < 0.0f || f_om < 0.0f || f_ov < 0.0f) return ERR_GRID;
    float2 vs = finish_velocity(P, f_sm, f_sqx, f_sqy, nx, ny);
    float2 vw = finish_velocity(P, f_wm, f_wqx, f_wqy, nx, ny);
    float2 vo = finish_velocity(P, f_om, f_oqx, f_oqy, nx, ny);
    if (nx >= 0 && ny >= 0 && nx <= (int)P.cells_x && ny <= (int)P.cells_y && f_wm > 0.0f && f_om > 0.0f &&
        f_wv > 0.0f && f_ov > 0.0f) {
        // solveLiquidGridContact with phaseVolumeGradient (shaders.ts:1527-1594, :1779-1799)
        const float wgx = nvol(0, 1, 0) - nvol(0, -1, 0), wgy = nvol(0, 0, 1) - nvol(0, 0, -1);
        const float ogx = nvol(1, 1, 0) - nvol(1, -1, 0), ogy = nvol(1, 0, 1) - nvol(1, 0, -1);
        const float sepx = ogx - wgx, sepy = ogy - wgy;
        // (the hardware square root, within an ulp, where WGSL allows 2.5)
        const float len = __builtin_amdgcn_sqrtf(sepx * sepx + sepy * sepy);
        if (finitef(len) && len > 1e-8f) {
            float nrx = quot(sepx, len), nry = quot(sepy, len);
            const float wd = quot(f_wm, f_wv), od = quot(f_om, f_ov);
            const float dscale = fmaxf(wd, od);
            const float contrast = quot(od - wd, dscale);
            const float gmag = __builtin_amdgcn_sqrtf(P.gx * P.gx + P.gy * P.gy);
            float align = 0.0f;
            if (!finitef(wd) || wd <= 0.0f || !finitef(od) || od <= 0.0f || !finitef(contrast) || !finitef(gmag)) {
                flags |= ERR_GRID;
            } else {
                if (gmag > 0.0f) align = quot(nrx * P.gx + nry * P.gy, gmag);
                if (!finitef(align)) flags |= ERR_GRID;
                if (contrast * align < -LIQUID_STABLE_ORDER_EPSILON) { nrx = -nrx; nry = -nry; }
                const float rel = (vo.x - vw.x) * nrx + (vo.y - vw.y) * nry;
                const float overlap = fminf(fmaxf(quot(2.0f * fminf(f_wv, f_ov), f_wv + f_ov), 0.0f), 1.0f);
                const float minsep =
                    fmaxf(quot(LIQUID_CONTACT_SEPARATION_FRACTION * overlap * P.dx, LIQUID_CONTACT_RELAXATION_SECONDS),
                          fmaxf(-rel, 0.0f) * LIQUID_CONTACT_APPROACH_SPEED_MULTIPLIER);
                if (rel < minsep) {
                    const float iwm = quot(1.0f, f_wm), iom = quot(1.0f, f_om);
                    const float imp = quot(minsep - rel, iwm + iom);
                    const float ix = imp * nrx, iy = imp * nry;
                    vw.x = vw.x - ix * iwm; vw.y = vw.y - iy * iwm;
                    vo.x = vo.x + ix * iom; vo.y = vo.y + iy * iom;
                }
            }
        }
    }
    // solveFluidTriangleGridContact (shaders.ts:1596-1665) in scenes with bodies
    flags |= tri(node, f_wm, f_om, vw, vo);
    if (!finite2(vs) || !finite2(vw) || !finite2(vo)) return flags | ERR_GRID | ERR_TRIANGLE_CONTACT;
    const int cxn = (int)P.cells_x, cyn = (int)P.cells_y;
    // A node at least a cushion away from every wall keeps its velocities:
    // each cushion and friction distance below is then >= its cushion.
    const float cush_x = fminf(WALL_CUSHION_CELLS, (float)cxn / 4.0f);
    const float cush_y = fminf(WALL_CUSHION_CELLS, (float)cyn / 4.0f);
    if (!((float)nx >= cush_x && (float)(cxn - nx) >= cush_x && (float)ny >= cush_y && (float)(cyn - ny) >= cush_y)) {
        vs.x = wall_cushion(vs.x, nx, cxn); vs.y = wall_cushion(vs.y, ny, cyn);
        vw.x = wall_cushion(vw.x, nx, cxn); vw.y = wall_cushion(vw.y, ny, cyn);
        vo.x = wall_cushion(vo.x, nx, cxn); vo.y = wall_cushion(vo.y, ny, cyn);
        const float fr = P.wall_friction;
        vs.x = wall_friction(vs.x, ny, cyn, fr); vs.y = wall_friction(vs.y, nx, cxn, fr);
        vw.x = wall_friction(vw.x, ny, cyn, fr); vw.y = wall_friction(vw.y, nx, cxn, fr);
        vo.x = wall_friction(vo.x, ny, cyn, fr); vo.y = wall_friction(vo.y, nx, cxn, fr);
    }
    if (scene_flags & SCENE_SOLID) grid.v_solid[node] = vs;
    if (scene_flags & SCENE_WATER) grid.v_water[node] = vw;
    if (scene_flags & SCENE_OIL) grid.v_oil[node] = vo;
    if (scene_flags & (SCENE_WATER | SCENE_OIL)) grid.liquid[node] = make_float2(f_wm + f_om, f_wv + f_ov);
    return flags;
}

// resolveFluidTriangleContact (shaders.ts:1624-1665): the fluid phase's
// velocity after the normal and tangential-drag impulse, and the reaction
// impulse on the body accumulated in (ix, iy).
__device__ __forceinline__ uint32_t fluid_triangle_contact(float fm, float2& v, float svx, float svy, float nx,
                                                           float ny, float inv_body, float dt, float& ix,
                                                           float& iy) {
    const float inv_eff = 1.0f / fm + inv_body;
    if (!finitef(inv_eff) || inv_eff <= 0.0f) return ERR_TRIANGLE_CONTACT;
    const float rx = v.x - svx, ry = v.y - svy;
    const float rn = rx * nx + ry * ny;
    const float nimp = rn < 0.0f ? -rn / inv_eff : 0.0f;
    const float tx = rx - nx * rn, ty = ry - ny * rn;
    const float damping = 1.0f - expf(-TRIANGLE_FLUID_TANGENTIAL_DRAG_RATE * dt);
    const float fx = nx * nimp + tx * (-damping / inv_eff), fy = ny * nimp + ty * (-damping / inv_eff);
    v.x = v.x + fx * (1.0f / fm);
    v.y = v.y + fy * (1.0f / fm);
    ix -= fx;
    iy -= fy;
    return (finite2(v) && finitef(ix) && finitef(iy)) ? 0u : ERR_TRIANGLE_CONTACT;
}

// Fluid contact of one node with the triangle of its field, whose vertex
// velocities come from the step's input buffer; the node's field keeps the
// impulse on the body.
__device__ __forceinline__ uint32_t fluid_triangle_node(const Params& P, const AuxData& A, const SlotBuf& in,
                                                        int in_buf, uint32_t node, float wm, float om, float2& vw,
                                                        float2& vo) {
    if (A.nbodies == 0u || (wm <= 0.0f && om <= 0.0f)) return 0u;
    const uint32_t bp1 = A.field_body[node];
    if (bp1 == 0u) return 0u;
    const FieldNode fn = A.field[node];
    if (fn.dist > P.reach) return 0u;
    const uint32_t a0 = 3u * (bp1 - 1u);
    const float bw[3] = {fn.b0, fn.b1, fn.b2};
    float svx = 0.0f, svy = 0.0f, inv_body = 0.0f;
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const float2 pv = in.vel[A.slot[in_buf][a0 + v]];
        svx += bw[v] * pv.x;
        svy += bw[v] * pv.y;
        inv_body += (bw[v] * bw[v]) / P.particle_mass;
    }
    if (!(inv_body > 0.0f)) return 0u;
    float ix = 0.0f, iy = 0.0f;
    uint32_t flags = 0u;
    if (wm > 0.0f) flags |= fluid_triangle_contact(wm, vw, svx, svy, fn.nx, fn.ny, inv_body, P.dt, ix, iy);
    if (om > 0.0f) flags |= fluid_triangle_contact(om, vo, svx, svy, fn.nx, fn.ny, inv_body, P.dt, ix, iy);
    A.field[node].ix = ix;
    A.field[node].iy = iy;
    return flags;
}

// The scene and scene-local tile of this block of a step launch: blocks
// are spread over the XCDs in bands of consecutive global tiles (xcd_tile),
// and the tile -> scene map names the scene of each global tile (a
// one-scene launch skips the lookup).
#define SCENE_AT(MAP, BASE, G)                                                       \
    const uint32_t global_tile_ = (G);                                               \
    const SceneDesc& S = scenes[nscenes == 1u ? 0u : (MAP)[global_tile_]];           \
    const uint32_t tile = global_tile_ - S.BASE

// The tiles of a step kernel's block: without a queue the one tile of the
// block (XCD-aware order); with one, the block's share of the active-tile
// list of its XCD: block k of the K blocks an XCD runs (the launch has 8K
// blocks, dispatched to the XCDs in turn) takes entries k, k + K, ... of that
// XCD's list, which holds the active tiles of the XCD's range of the launch,
// so the XCD keeps the tiles (and the halos they share) the one-block-a-tile
// launch would give it.  No atomics: every value is uniform over the block.
template <class Body>
__device__ __forceinline__ void run_tiles(const TileQueue& q, uint32_t kind, uint32_t step, Body body) {
    if (q.sets == nullptr) {
        body(xcd_tile(gridDim.x));
        return;
    }
    const uint32_t x = blockIdx.x % QUEUE_XCDS, k = blockIdx.x / QUEUE_XCDS, K = gridDim.x / QUEUE_XCDS;
    const uint32_t n = q.sets[(step & 1u) * QUEUE_SET_WORDS + kind * QUEUE_XCDS + x];
    const uint32_t* list = q.list[kind] + xcd_range_start(q.tiles[kind], x);
    for (uint32_t i = k; i < n; i += K) body(list[i]);
}

// Stress-cache failures of the state a microstep starts from, raised once a
// microstep by the scene's first P2G tile (in a batch every scene has one),
// and the deferred words of the microstep's output cleared.
__device__ __forceinline__ void scene_deferred_flags(Control* ctl, int in_buf, uint32_t step) {
    const int out_buf = in_buf ^ 1;
    const uint32_t d = ctl->deferred[in_buf];
    if (d) {
        if (d & ERR_P2G) atomicMin(&ctl->status[1], ctl->deferred_index[in_buf]);
        raise_step(ctl, d, step);
    }
    ctl->deferred[out_buf] = 0u;
    ctl->deferred_index[out_buf] = 0xffffffffu;
}

// Shared head of both P2G kernels: status gate, deferred stress-cache flags,
// clearing the next bucket counts this tile owns, and the staged offsets of
// the region buckets (s_boff, one wave, two buckets a lane).  Returns
// false when a flag is already raised (the whole block returns).
template <class T, bool OCC = false>
__device__ __forceinline__ bool p2g_head(Control* ctl, const SlotBuf& in, uint32_t* cnt_next, uint32_t tile,
                                         uint32_t tiles_x, uint32_t bw, uint32_t bh, uint32_t cap, uint32_t step,
                                         int in_buf, uint32_t* s_boff) {
    const int tid = threadIdx.x;
    // After a failed microstep the output buffer of this one holds the
    // committed state: leave its counts and deferred flags untouched.
    if (block_status_set(ctl)) return false;
    // the occupancy gather only reads the committed state
    if (!OCC && tile == 0u && tid == 0) scene_deferred_flags(ctl, in_buf, step);
    const int tx = (int)(tile % tiles_x), ty = (int)(tile / tiles_x);
    // this tile owns buckets [OWN tx, OWN tx + OWN) x [OWN ty, OWN ty + OWN): clear their next counts
    if (!OCC && tid < T::OWN * T::OWN) {
        const int obx = T::OWN * tx + tid % T::OWN, oby = T::OWN * ty + tid / T::OWN;
        if (obx < (int)bw && oby < (int)bh) cnt_next[(uint32_t)oby * bw + (uint32_t)obx] = 0u;
    }
    const int rbx0 = tx * T::OWN - 1, rby0 = ty * T::OWN - 1;
    if (tid < 64) {
        uint32_t c[2];
#pragma unroll
        for (int k = 0; k < 2; ++k) {
            const int b = tid + 64 * k;
            const int gbx = rbx0 + b % T::RB, gby = rby0 + b / T::RB;
            c[k] = 0u;
            if (b < T::NB && gbx >= 0 && gby >= 0 && gbx < (int)bw && gby < (int)bh)
                c[k] = min(in.cnt[(uint32_t)gby * bw + (uint32_t)gbx], cap);
        }
        uint32_t i0 = c[0], i1 = c[1];
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t0 = __shfl_up(i0, off, 64), t1 = __shfl_up(i1, off, 64);
            if (tid >= off) { i0 += t0; i1 += t1; }
        }
        const uint32_t tot0 = __shfl(i0, 63, 64), tot1 = __shfl(i1, 63, 64);
        if (tid < T::NB) s_boff[tid] = i0 - c[0];
        if (tid + 64 < T::NB) s_boff[tid + 64] = tot0 + i1 - c[1];
        if (tid == 0) s_boff[T::NB] = tot0 + tot1;
    }
    return true;
}

// A P2G tile whose region holds no particle: its nodes receive
// nothing, and no particle is within reach of them in this microstep's G2P,
// so their velocities are not rewritten.  The kept fixed-point words and
// the diagnostics partial still read zero.
template <int TI>
__device__ void p2g_empty_tile(const Params& P, const Grid& grid, uint32_t keep_fixed, double* diag_grid,
                               uint32_t tile, int X0, int Y0) {
    if (keep_fixed) {
        for (int e = threadIdx.x; e < TI * TI; e += blockDim.x) {
            const int sx = X0 + e % TI, sy = Y0 + e / TI;
            if (sx >= 0 && sy >= 0 && sx < (int)P.node_w && sy < (int)P.node_h) {
                const size_t node = (size_t)sy * P.node_w + (size_t)sx;
#pragma unroll
                for (int k = 0; k < 11; ++k) grid.fixed[node * 11 + k] = 0;
            }
        }
    }
    if (P.diag_on && threadIdx.x == 0) {
        diag_grid[3 * tile] = 0.0;
        diag_grid[3 * tile + 1] = 0.0;
        diag_grid[3 * tile + 2] = 0.0;
    }
}

// Region bucket holding staged particle j of [s_boff[b0], s_boff[b1]).
__device__ __forceinline__ int region_bucket(const uint32_t* s_boff, int b0, int b1, uint32_t j) {
    int lo = b0, hi = b1 - 1;
    while (lo < hi) {  // last region bucket with s_boff <= j
        const int mid = (lo + hi + 1) >> 1;
        if (s_boff[mid] <= j) lo = mid; else hi = mid - 1;
    }
    return lo;
}

// Exclusive scan of s[0, n) in place (n <= 2 * THREADS); s[n] receives the
// total.  Every thread of the block must call it; it ends with a barrier.
template <int THREADS>
__device__ void block_exclusive_scan(uint32_t* s, int n, uint32_t* s_wave) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int i0 = 2 * tid;
    const uint32_t a = i0 < n ? s[i0] : 0u, b = i0 + 1 < n ? s[i0 + 1] : 0u;
    const uint32_t pair = a + b;
    uint32_t incl = pair;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t t = __shfl_up(incl, off, 64);
        if (lane >= off) incl += t;
    }
    if (lane == 63) s_wave[wave] = incl;
    __syncthreads();
    if (wave == 0) {
        constexpr int WAVES = THREADS / 64;
        const uint32_t w = lane < WAVES ? s_wave[lane] : 0u;
        uint32_t wi = w;
#pragma unroll
        for (int off = 1; off < WAVES; off <<= 1) {
            const uint32_t t = __shfl_up(wi, off, 64);
            if (lane >= off) wi += t;
        }
        if (lane < WAVES) s_wave[lane] = wi - w;
        if (lane == WAVES - 1) s_wave[WAVES] = wi;
    }
    __syncthreads();
    const uint32_t base = s_wave[wave] + incl - pair;
    if (i0 < n) s[i0] = base;
    if (i0 + 1 < n) s[i0 + 1] = base + a;
    if (tid == 0) s[n] = s_wave[THREADS / 64];
    __syncthreads();
}

// Float P2G: LDS-staged gather (p2g, shaders.ts:1072-1211) fused with
// updateGrid (:1485-1799).  One block per TI x TI-node tile: it gathers the
// tile and a one-node ring of complete sums, then updates the interior (the
// ring supplies the volume gradient).
//
// The tile's particles are those of its region buckets (for TI = 20, 6x6
// buckets: a 24x24-cell region).  They are staged in LDS in
// region-row-major cell order, ascending particle index inside a cell, so
// each node sums its three stencil rows as three contiguous ranges in a fixed
// order (rows, then cells, then particle index).  That order does not depend
// on the tile: the float sums are the same run to run and for every tile
// side.  A region holding more than CAP particles is processed in chunks of
// whole bucket rows, which keeps the order; only a single bucket row above
// CAP is split into single buckets, and then a node whose stencil row spans
// two buckets of different chunks adds them bucket by bucket.  The sums carry
// across chunks.
//
// OCC = true is the surface occupancy gather (gatherSurfaceOccupancy,
// shaders.ts:710-806) of the committed state: the same staging and gather,
// no grid update, no bucket bookkeeping and no flags; every node receives
// (water, oil) = 2 (sum w V det F) / dx^2, the display area of the liquid
// phases over the cell area.  The reference scales each contribution,
// ((V det F) 2) w / dx^2; here the sum is scaled once.
template <int TI, bool GENERAL, bool AUX, bool OCC>
__device__ __forceinline__ void p2g_gather_tile(const SceneDesc* __restrict__ scenes,
                                                const uint32_t* __restrict__ tile_scene, uint32_t nscenes,
                                                uint32_t step, uint32_t global_tile) {
    using T = P2GTile<TI>;
    constexpr int G_CAP = T::CAP;
    SCENE_AT(tile_scene, p2g_base, global_tile);
    const int in_buf = (int)((S.start_buf + step) & 1u);
    const Params* __restrict__ Pp = &S.P;
    Control* ctl = S.ctl;
    const SlotBuf in = S.buf[in_buf];
    uint32_t* cnt_next = S.buf[in_buf ^ 1].cnt;
    const Material* __restrict__ mats = S.mats;
    const Grid grid = S.grid;
    const uint32_t nmat = S.nmat, tiles_x = S.ptx, bw = S.bw, bh = S.bh, cap = S.cap, scene_flags = S.scene_flags;
    double* diag_grid = S.diag_grid;
    __shared__ uint32_t s_boff[T::NB + 1];    // staged offset of each region bucket
    __shared__ uint32_t s_chunk[T::NB + 1];   // chunk boundaries (region bucket indices)
    __shared__ uint32_t s_nchunks;
    __shared__ uint32_t s_wave[T::THREADS / 64 + 1];
    __shared__ Material s_mat[LDS_MATS];
    __shared__ uint32_t s_cell[T::CELLS + 1]; // per region cell: count, then start
    // staged particle records in gather order, packed so that a contribution
    // reads them in three wide LDS loads (and consecutive lanes read
    // consecutive 16-byte records, which keeps the banks apart)
    __shared__ float4 s_r0[G_CAP];     // stencil fraction fx, fy; momentum m vx, m vy
    __shared__ float4 s_r1[G_CAP];     // affine A (row-major)
    // mass, volume x det F, the tag (region column | phase << 8) as bits;
    // the fourth word holds the particle
    // indices at their provisional places, which the ranking reads while the
    // first three words are written
    __shared__ float4 s_r2[G_CAP];
    uint32_t* const s_pid = reinterpret_cast<uint32_t*>(s_r2) + 3;  // stride 4 words
    // In scenes without the stress cache a region of one chunk maps each
    // staged place to (region bucket << 16 | place in the bucket) in s_r1,
    // which (D) overwrites only after (A); elsewhere (A) searches s_boff
    // (measured: the extra barrier costs more than the search in the sparse
    // tiles of batched Tait tanks).
    uint32_t* const s_map = reinterpret_cast<uint32_t*>(s_r1);
    constexpr bool MAPPED = !GENERAL;

    stage_materials(s_mat, mats, nmat);
    if (!p2g_head<T, OCC>(ctl, in, cnt_next, tile, tiles_x, bw, bh, cap, step, in_buf, s_boff)) return;
    const int tid = threadIdx.x;
    const Params& P = *Pp;
    const int tx = (int)(tile % tiles_x), ty = (int)(tile / tiles_x);
    const int lx = tid % T::ROW, ly = tid / T::ROW;
    const int X0 = tx * TI - 1, Y0 = ty * TI - 1;  // first interior node
    const int W = (int)P.node_w, H = (int)P.node_h;
    const int rbx0 = tx * T::OWN - 1, rby0 = ty * T::OWN - 1;
    const int sx = X0 - 1 + lx, sy = Y0 - 1 + ly;  // this thread's node (storage coords)
    const bool node_ok = lx < T::TC && ly < T::TC && sx >= 0 && sx < W && sy >= 0 && sy < H;
    const bool interior = node_ok && lx >= 1 && lx <= TI && ly >= 1 && ly <= TI;
    for (int e = tid; e <= T::CELLS; e += T::THREADS) s_cell[e] = 0u;  // the first chunk's cell counts
    __syncthreads();
    if (s_boff[T::NB] == 0u) {
        if constexpr (OCC) {
            if (interior) grid.occupancy[(uint32_t)sy * (uint32_t)W + (uint32_t)sx] = make_float2(0.f, 0.f);
        } else {
            p2g_empty_tile<TI>(P, grid, 0u, diag_grid, tile, X0, Y0);
        }
        return;
    }
    // one chunk unless the region holds more particles than the staging
    // arrays (the total is the same for every thread, so is the branch)
    const bool chunked = s_boff[T::NB] > (uint32_t)G_CAP;
    if (chunked && tid == 0) {
        uint32_t n = 0;
        s_chunk[0] = 0u;
        for (int r = 0; r < T::RB; ++r) {
            const int lo = r * T::RB, hi = lo + T::RB;
            if (s_boff[hi] - s_boff[s_chunk[n]] <= (uint32_t)G_CAP) continue;
            if ((int)s_chunk[n] < lo) s_chunk[++n] = (uint32_t)lo;
            if (s_boff[hi] - s_boff[lo] > (uint32_t)G_CAP) {
                for (int b = lo; b < hi; ++b)
                    if (s_boff[b + 1] - s_boff[s_chunk[n]] > (uint32_t)G_CAP && b > (int)s_chunk[n])
                        s_chunk[++n] = (uint32_t)b;
            }
        }
        s_chunk[++n] = T::NB;
        s_nchunks = n;
    }
    if (MAPPED && !chunked) {  // the slot map, several threads a region bucket
        constexpr int TPB = T::THREADS / T::NB;
        if (tid < TPB * T::NB) {
            const int b = tid / TPB;
            const uint32_t b0 = s_boff[b], n = s_boff[b + 1] - b0;
            for (uint32_t k = (uint32_t)(tid % TPB); k < n; k += TPB) s_map[b0 + k] = (uint32_t)b << 16 | k;
        }
    }
    if (MAPPED || chunked) __syncthreads();

    float acc[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) acc[k] = 0.0f;
    uint32_t flags = 0;
    const uint32_t nchunks = chunked ? s_nchunks : 1u;
    const int cx0 = rbx0 * BUCKET_CELLS, cy0 = rby0 * BUCKET_CELLS;  // first region cell (stencil-base coords)
    constexpr int PER = G_CAP / T::THREADS + (G_CAP % T::THREADS ? 1 : 0);  // staged particles a thread

    for (uint32_t ch = 0; ch < nchunks; ++ch) {
        const int b0 = chunked ? (int)s_chunk[ch] : 0, b1 = chunked ? (int)s_chunk[ch + 1] : T::NB;
        const uint32_t c0 = s_boff[b0], c1 = s_boff[b1];
        if (ch > 0u) {  // the first chunk's counts were cleared before the head's barrier
            for (int e = tid; e <= T::CELLS; e += T::THREADS) s_cell[e] = 0u;
            __syncthreads();
        }
        // (A) positions: region cell of each particle, counted per cell
        uint32_t cell[PER], pid[PER], gi[PER];  // cell: region cell << 16 | place in the cell
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t j = c0 + (uint32_t)tid + (uint32_t)(k * T::THREADS);
            cell[k] = 0xffffffffu;
            pid[k] = 0u; gi[k] = 0u;
            if (j >= c1) continue;
            int rb;
            uint32_t place;
            if (MAPPED && !chunked) {
                const uint32_t e = s_map[j];
                rb = (int)(e >> 16);
                place = e & 0xffffu;
            } else {
                rb = region_bucket(s_boff, b0, b1, j);
                place = j - s_boff[rb];
            }
            const int gbx = rbx0 + rb % T::RB, gby = rby0 + rb / T::RB;
            gi[k] = ((uint32_t)gby * bw + (uint32_t)gbx) * cap + place;
            const float2 p = in.pos[gi[k]];
            pid[k] = in.pid[gi[k]];
            Stencil sc;
            if (make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, p, sc) && ((sc.bx + 1) >> 2) == gbx &&
                ((sc.by + 1) >> 2) == gby) {
                const uint32_t rc = (uint32_t)((sc.by + 1 - cy0) * T::RC + (sc.bx + 1 - cx0));
                cell[k] = rc << 16 | atomicAdd(&s_cell[rc], 1u);
            } else {
                flags |= ERR_STENCIL;
            }
        }
        __syncthreads();
        // (B) cell starts
        block_exclusive_scan<T::THREADS>(s_cell, T::CELLS, s_wave);
        // (C) particle indices at their provisional places
#pragma unroll
        for (int k = 0; k < PER; ++k)
            if (cell[k] != 0xffffffffu) s_pid[4u * (s_cell[cell[k] >> 16] + (cell[k] & 0xffffu))] = pid[k];
        __syncthreads();
        // (D) rank by particle index inside the cell (the summation order),
        // then stage the prepared affine (prepareParticle, shaders.ts:808-963)
        // and the per-particle factors of the contributions
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (cell[k] == 0xffffffffu) continue;
            // the inputs load while the rank is counted
            const uint32_t g = gi[k];
            const float2 pos = in.pos[g];
            const uint32_t mat = in.mat[g];
            const float2 v = in.vel[g];
            const float4 C = in.C[g];
            const float4 F = in.F[g];
            float4 S = make_float4(0.f, 0.f, 0.f, 0.f);
            float pv = 1.0f;
            if constexpr (GENERAL) {
                S = in.S[g];
                pv = in.plastic[g];
            }
            const uint32_t rc = cell[k] >> 16;
            const uint32_t q0 = s_cell[rc], q1 = s_cell[rc + 1];  // q1 > q0: this particle is one of them
            uint32_t q = q0;
            // the cell's first four places in one batch of reads (an index
            // past the cell reads its last place and counts nothing), then
            // the rest of a fuller cell
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) {
                const uint32_t t = q0 + u;
                const uint32_t other = s_pid[4u * min(t, q1 - 1u)];
                q += ((t < q1) & (other < pid[k])) ? 1u : 0u;
            }
            for (uint32_t t = q0 + 4u; t < q1; ++t) q += s_pid[4u * t] < pid[k] ? 1u : 0u;
            Stencil sc;  // recomputed (it fitted in (A)) rather than held across the barriers
            make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, pos, sc);
            const Material mt = material_at(s_mat, mats, mat);
            const float det = det_comp(F);
            uint32_t pf = 0;
            if constexpr (GENERAL) {
                if (!finitef(det) || det <= 0.0f || !finitef(mt.volume) || mt.volume <= 0.0f) pf = ERR_DEFORMATION;
            } else {
                pf = particle_stress<false>(mt, F, pv, det, S);
            }
            const float4 A = affine_of(P, mt, C, S);
            if (!pf && !finite4(A)) pf = ERR_P2G | ERR_NONFINITE_PARTICLE;
            // The reference tests every stencil weight for finiteness and
            // sign (shaders.ts:1153-1157).  A stencil that make_stencil
            // accepted has |x/dx| < 2^17, so x/dx - 0.5, its floor and the
            // fraction are exact and the fraction lies in [0.5, 1.5): all
            // three quadratic weights are then finite and non-negative, and
            // the test cannot fail.
            if (pf && !OCC) {
                if (pf & ERR_P2G) atomicMin(&ctl->status[1], pid[k]);
                flags |= pf;
            }
            const float mvx = mt.mass * v.x, mvy = mt.mass * v.y;
            // the affine scaled by the cell size: a contribution's momentum
            // is then m v + (A dx) d with d = offset - fraction, two fused
            // multiply-adds per component
            const float4 Ad = make_float4(A.x * P.dx, A.y * P.dx, A.z * P.dx, A.w * P.dx);
            // a non-finite contribution fails the microstep (shaders.ts:1153-1172);
            // the nine momenta are tested here, once a staged particle, so the
            // gather loop does not test each pair
            bool drop = pf != 0u;
            if constexpr (!OCC) {
                if (!drop && !contributions_finite(sc.fx, sc.fy, mvx, mvy, Ad)) {
                    flags |= ERR_P2G | ERR_NONFINITE_PARTICLE;
                    drop = true;
                }
            }
            // a failed particle is staged as a solid of zero mass, volume and
            // momentum: its contributions add exact zeros
            const uint32_t col = rc % (uint32_t)T::RC;
            // word by word (a select of whole vectors compiles to a stack copy)
            auto kept = [drop](float v) { return drop ? 0.0f : v; };
            s_r0[q] = make_float4(sc.fx, sc.fy, kept(mvx), kept(mvy));
            s_r1[q] = make_float4(kept(Ad.x), kept(Ad.y), kept(Ad.z), kept(Ad.w));
            *reinterpret_cast<float3*>(&s_r2[q]) =
                make_float3(kept(mt.mass), kept(mt.volume * det), __uint_as_float(col | (drop ? 0u : mt.phase << 8)));
        }
        __syncthreads();
        // (E) gather: stencil rows ly..ly+2, columns lx..lx+2 of the region.
        // The three row ranges run as one loop, so a wave iterates over the
        // largest per-node pair count rather than the largest per cell.
        if (node_ok) {
            uint32_t q0[3], n[3];
#pragma unroll
            for (int dy = 0; dy < 3; ++dy) {
                const int row = (ly + dy) * T::RC + lx;
                q0[dy] = s_cell[row];
                n[dy] = s_cell[row + 3] - q0[dy];
            }
            const uint32_t n01 = n[0] + n[1], total = n01 + n[2];
            // pair k lies in row range dy = (k >= n[0]) + (k >= n01), at
            // q = k + base[dy]; the choices below are selects, so lanes at
            // different rows or columns do not diverge, and the three LDS reads
            // of a pair are issued together (one index scales to all three)
            const uint32_t base0 = q0[0], base1 = q0[1] - n[0], base2 = q0[2] - n01;
#pragma unroll 1
            for (uint32_t k = 0; k < total; ++k) {
                const bool row2 = k < n[0], row1 = !row2 && k < n01;  // oy = 2, oy = 1
                const uint32_t q = k + (row2 ? base0 : (row1 ? base1 : base2));
                const float4 r0 = s_r0[q], r1 = s_r1[q], r2 = s_r2[q];
                const uint32_t tag = __float_as_uint(r2.z);
                // the empty asm takes one word of each read, so the three reads
                // are issued back to back and waited for once (left alone, the
                // scheduler waits for the tag before issuing the others, and
                // defers the mass and volume into the phase branches)
                asm volatile("" ::"v"(r0.x), "v"(r1.x), "v"(r2.x));
                const uint32_t ph = tag >> 8;
                const int ox = lx + 2 - (int)(tag & 0xffu);
                // d = offset - fraction per axis, the reference's dpos / dx
                const float ddx = (float)ox - r0.x;
                const float ddy = (row2 ? 2.0f : (row1 ? 1.0f : 0.0f)) - r0.y;
                const float w = bspline_of_offset(ddx, ox == 1) * bspline_of_offset(ddy, row1);
                const float qx = fmaf(r1.y, ddy, fmaf(r1.x, ddx, r0.z));
                const float qy = fmaf(r1.w, ddy, fmaf(r1.z, ddx, r0.w));
                // each product is added with one rounding
                if constexpr (GENERAL) {
                    if (ph == 0u) {
                        acc[0] = fmaf(w, r2.x, acc[0]); acc[1] = fmaf(w, qx, acc[1]); acc[2] = fmaf(w, qy, acc[2]);
                    } else if (ph == 1u) {
                        acc[3] = fmaf(w, r2.x, acc[3]); acc[4] = fmaf(w, qx, acc[4]);
                        acc[5] = fmaf(w, qy, acc[5]); acc[6] = fmaf(w, r2.y, acc[6]);
                    } else {
                        acc[7] = fmaf(w, r2.x, acc[7]); acc[8] = fmaf(w, qx, acc[8]);
                        acc[9] = fmaf(w, qy, acc[9]); acc[10] = fmaf(w, r2.y, acc[10]);
                    }
                } else {
                    // Scenes without the stress cache hold water and oil only
                    // (and failed particles, staged with zero mass, volume and
                    // momentum).  Both phases' sums take every pair, the other
                    // phase's with weight exactly zero: fmaf(0, x, s) == s for
                    // the finite staged values and any sum that is not -0, which
                    // a sum starting at +0 never is.  No divergent branch, so
                    // none of its exec-mask bookkeeping either.
                    const float ww = ph == 1u ? w : 0.0f, wo = w - ww;
                    acc[3] = fmaf(ww, r2.x, acc[3]); acc[4] = fmaf(ww, qx, acc[4]);
                    acc[5] = fmaf(ww, qy, acc[5]); acc[6] = fmaf(ww, r2.y, acc[6]);
                    acc[7] = fmaf(wo, r2.x, acc[7]); acc[8] = fmaf(wo, qx, acc[8]);
                    acc[9] = fmaf(wo, qy, acc[9]); acc[10] = fmaf(wo, r2.y, acc[10]);
                }
            }
        }
        __syncthreads();
    }

    if constexpr (OCC) {
        if (interior) {
            const float cell_area = P.dx * P.dx;
            grid.occupancy[(uint32_t)sy * (uint32_t)W + (uint32_t)sx] =
                make_float2((acc[6] * 2.0f) / cell_area, (acc[10] * 2.0f) / cell_area);
        }
        return;
    }
    // ring volumes for the gradient (the staging arrays are free now)
    float* s_vw = reinterpret_cast<float*>(s_r0);
    float* s_vo = s_vw + G_CAP;
    const int ring = ly * T::TC + lx;
    if (lx < T::TC && ly < T::TC) {
        s_vw[ring] = node_ok ? acc[6] : 0.0f;
        s_vo[ring] = node_ok ? acc[10] : 0.0f;
    }
    __syncthreads();
    if (node_ok && lx >= 1 && lx <= TI && ly >= 1 && ly <= TI) {
        const uint32_t node = (uint32_t)sy * (uint32_t)W + (uint32_t)sx;
        auto nvol = [&](int k, int ox, int oy) {
            const float* a = k == 0 ? s_vw : s_vo;
            return a[ring + oy * T::TC + ox];
        };
        if constexpr (AUX) {
            const AuxData& A = S.aux;
            flags |= update_node(P, grid, scene_flags, node, sx - 1, sy - 1, acc, ERR_P2G | ERR_NONFINITE_ACCUMULATION,
                                 nvol, [&](uint32_t nd, float wm, float om, float2& vw, float2& vo) {
                                     return fluid_triangle_node(P, A, in, in_buf, nd, wm, om, vw, vo);
                                 });
        } else {
            flags |= update_node(P, grid, scene_flags, node, sx - 1, sy - 1, acc, ERR_P2G | ERR_NONFINITE_ACCUMULATION,
                                 nvol);
        }
    }
    if (P.diag_on) {  // grid mass and momentum of this microstep (reference/solver.ts:481-492)
        const bool mine = node_ok && lx >= 1 && lx <= TI && ly >= 1 && ly <= TI;
        block_sum3_f64(mine ? (double)acc[0] + (double)acc[3] + (double)acc[7] : 0.0,
                       mine ? (double)acc[1] + (double)acc[4] + (double)acc[8] : 0.0,
                       mine ? (double)acc[2] + (double)acc[5] + (double)acc[9] : 0.0, diag_grid + 3 * tile);
    }
    if (flags) raise_step(ctl, flags, step);
}


template <int TI, bool GENERAL, bool AUX = false, bool OCC = false>
__global__ __launch_bounds__(P2GTile<TI>::THREADS)
__attribute__((amdgpu_waves_per_eu(P2GTile<TI>::GATHER_WAVES_PER_EU, 8)))
void k_p2g_gather(const SceneDesc* __restrict__ scenes, const uint32_t* __restrict__ tile_scene, uint32_t nscenes,
                  uint32_t step, TileQueue q) {
    run_tiles(q, 0u, step, [&](uint32_t t) { p2g_gather_tile<TI, GENERAL, AUX, OCC>(scenes, tile_scene, nscenes, step, t); });
}

// Fixed-point P2G (prepareAndScatterParticles + gatherAtomicGrid,
// shaders.ts:971-1070) fused with updateGrid.  Same tiles as the float
// gather, but every particle of the region adds its rounded int32
// contributions that land on the tile or its ring into 64-bit LDS words with
// integer atomics: integer sums do not depend on the order, and an exact sum
// outside int32 is the accumulator overflow of 65536.  Scenes without solids
// keep only the eight liquid channels, so two tiles fit in one CU's LDS.
template <int TI, bool GENERAL, bool SOLIDS, bool AUX>
__device__ __forceinline__ void p2g_fixed_tile(const SceneDesc* __restrict__ scenes,
                                               const uint32_t* __restrict__ tile_scene, uint32_t nscenes,
                                               uint32_t step, uint32_t global_tile) {
    using T = P2GTile<TI>;
    SCENE_AT(tile_scene, p2g_base, global_tile);
    const int in_buf = (int)((S.start_buf + step) & 1u);
    const Params* __restrict__ Pp = &S.P;
    Control* ctl = S.ctl;
    const SlotBuf in = S.buf[in_buf];
    uint32_t* cnt_next = S.buf[in_buf ^ 1].cnt;
    const Material* __restrict__ mats = S.mats;
    const Grid grid = S.grid;
    const uint32_t tiles_x = S.ptx, bw = S.bw, bh = S.bh, cap = S.cap, scene_flags = S.scene_flags;
    const uint32_t keep_fixed = S.keep_fixed;
    double* diag_grid = S.diag_grid;
    // words [solid m, qx, qy |] water m, qx, qy, V | oil m, qx, qy, V
    constexpr int NCH = SOLIDS ? 11 : 8;
    constexpr int SKIP = SOLIDS ? 0 : 3;  // channel index = word index - SKIP
    __shared__ unsigned long long s_acc[NCH][T::TC * T::TC];
    __shared__ uint32_t s_boff[T::NB + 1];

    for (int e = threadIdx.x; e < NCH * T::TC * T::TC; e += T::THREADS) (&s_acc[0][0])[e] = 0ull;
    if (!p2g_head<T>(ctl, in, cnt_next, tile, tiles_x, bw, bh, cap, step, in_buf, s_boff)) return;
    __syncthreads();
    const int tid = threadIdx.x;
    const Params& P = *Pp;
    const int tx = (int)(tile % tiles_x), ty = (int)(tile / tiles_x);
    const int X0 = tx * TI - 1, Y0 = ty * TI - 1;  // first interior node
    const int W = (int)P.node_w, H = (int)P.node_h;
    const int rbx0 = tx * T::OWN - 1, rby0 = ty * T::OWN - 1;
    if (s_boff[T::NB] == 0u) {
        p2g_empty_tile<TI>(P, grid, keep_fixed, diag_grid, tile, X0, Y0);
        return;
    }

    uint32_t flags = 0;
    const float mref = P.particle_mass, vref = P.particle_volume, dx = P.dx;
    const float rm = P.inv_mass_ref, rv = P.inv_volume_ref;
    const uint32_t total = s_boff[T::NB];
    // one particle a thread a pass: with two (both loads in flight) the
    // second one's inputs spill to scratch across the first one's scatter
    constexpr int PER = 1;
    for (uint32_t base = 0; base < total; base += (uint32_t)PER * T::THREADS) {
        uint32_t pid[PER], mat[PER];
        int gbx[PER], gby[PER];
        float2 p[PER], v[PER];
        float4 C[PER], F[PER], S[PER];
        float pv[PER];
        bool have[PER];
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t j = base + (uint32_t)tid + (uint32_t)(k * T::THREADS);
            have[k] = j < total;
            pid[k] = 0u; mat[k] = 0u; gbx[k] = 0; gby[k] = 0;
            p[k] = make_float2(0.f, 0.f); v[k] = p[k];
            C[k] = make_float4(0.f, 0.f, 0.f, 0.f); F[k] = C[k]; S[k] = C[k];
            pv[k] = 1.0f;
            if (!have[k]) continue;
            const int rb = region_bucket(s_boff, 0, T::NB, j);
            gbx[k] = rbx0 + rb % T::RB;
            gby[k] = rby0 + rb / T::RB;
            const size_t g = ((size_t)gby[k] * bw + (size_t)gbx[k]) * cap + (j - s_boff[rb]);
            p[k] = in.pos[g];
            pid[k] = in.pid[g];
            mat[k] = in.mat[g];
            v[k] = in.vel[g];
            C[k] = in.C[g];
            F[k] = in.F[g];
            if constexpr (GENERAL) {
                S[k] = in.S[g];
                pv[k] = in.plastic[g];
            }
        }
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            if (!have[k]) continue;
            Stencil sc;
            if (!make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, p[k], sc) || ((sc.bx + 1) >> 2) != gbx[k] ||
                ((sc.by + 1) >> 2) != gby[k]) {
                flags |= ERR_STENCIL;
                continue;
            }
            // tile-local node of stencil offset (0, 0); the tile's nodes are
            // storage nodes [X0-1, X0+TI] x [Y0-1, Y0+TI]
            const int nx0 = sc.bx + 2 - X0, ny0 = sc.by + 2 - Y0;
            if (nx0 + 2 < 0 || ny0 + 2 < 0 || nx0 >= T::TC || ny0 >= T::TC) continue;
            const Material mt = mats[mat[k]];
            const float det = det_comp(F[k]);
            uint32_t pf = 0;
            float4 Sk = S[k];
            if constexpr (GENERAL) {
                if (!finitef(det) || det <= 0.0f || !finitef(mt.volume) || mt.volume <= 0.0f) pf = ERR_DEFORMATION;
            } else {
                pf = particle_stress<false>(mt, F[k], pv[k], det, Sk);
            }
            const float4 A = affine_of(P, mt, C[k], Sk);
            if (!pf && !finite4(A)) pf = ERR_P2G | ERR_NONFINITE_PARTICLE;
            if (pf) {
                if (pf & ERR_P2G) atomicMin(&ctl->status[1], pid[k]);
                flags |= pf;
                continue;
            }
            const uint32_t ph = mt.phase;
            if (!SOLIDS && ph == 0u) {  // the scene flags list no solid: never taken
                flags |= ERR_P2G | ERR_INVALID_MATERIAL;
                continue;
            }
            const int c0 = (ph == 0u ? 0 : (ph == 1u ? 3 : 7)) - SKIP;
            const float m = mt.mass;
            const float mvx = m * v[k].x, mvy = m * v[k].y;
            const float vd = mt.volume * det;
            float wx[3], wy[3];
#pragma unroll
            for (int o = 0; o < 3; ++o) {
                wx[o] = qweight(sc.fx, o);
                wy[o] = qweight(sc.fy, o);
            }
            // The quotients by particleMass and particleVolume: the exact
            // reciprocal form when the host supplied both reciprocals (one
            // uniform test a particle instead of one a quotient), else
            // div_ref, which picks per quotient.
            auto scatter = [&](auto exact) {
                constexpr bool EXACT = decltype(exact)::value;
                auto quot_ref = [](float a, float b, float r) {
                    if constexpr (EXACT) {
                        const float q = a * r;
                        return fmaf(fmaf(-q, b, a), r, q);
                    } else {
                        return div_ref(a, b, r);
                    }
                };
    #pragma unroll
                for (int oy = 0; oy < 3; ++oy) {
                    const int ly = ny0 + oy;
                    const bool inside_y = (ly >= 0) & (ly < T::TC);
                    const float dpy = ((float)oy - sc.fy) * dx;
    #pragma unroll
                    for (int ox = 0; ox < 3; ++ox) {
                        const int lx = nx0 + ox;
                        const float dpx = ((float)ox - sc.fx) * dx;
                        const float w = wx[ox] * wy[oy];
                        const float qx = (mvx + A.x * dpx) + A.y * dpy;
                        const float qy = (mvy + A.z * dpx) + A.w * dpy;
                        // one predicate a pair: the node lies in the tile, and
                        // (tested only then) the contribution is finite
                        const bool inside = inside_y & (lx >= 0) & (lx < T::TC);
                        const bool fw = finitef(w), fqx = finitef(qx), fqy = finitef(qy);
                        const bool bad = !fw | (w < 0.0f) | !fqx | !fqy;
                        flags |= (inside & bad) ? (ERR_P2G | ERR_NONFINITE_PARTICLE) : 0u;
                        if (!(inside & !bad)) continue;
                        const int node = ly * T::TC + lx;
                        int32_t im, iqx, iqy, iv = 0;
                        flags |= fixed_round(quot_ref(w * m, mref, rm), im);
                        flags |= fixed_round(quot_ref(w * qx, mref, rm), iqx);
                        flags |= fixed_round(quot_ref(w * qy, mref, rm), iqy);
                        // (scenes without solids: every particle here is liquid)
                        if (!SOLIDS || ph != 0u) flags |= fixed_round(quot_ref(w * vd, vref, rv), iv);
                        atomicAdd(&s_acc[c0][node], (unsigned long long)(long long)im);
                        atomicAdd(&s_acc[c0 + 1][node], (unsigned long long)(long long)iqx);
                        atomicAdd(&s_acc[c0 + 2][node], (unsigned long long)(long long)iqy);
                        if (!SOLIDS || ph != 0u)
                            atomicAdd(&s_acc[c0 + 3][node], (unsigned long long)(long long)iv);
                    }
                }
            };
            if (rm != 0.0f && rv != 0.0f) scatter(std::true_type{});
            else scatter(std::false_type{});
        }
    }
    __syncthreads();

    // gatherAtomicGrid decode (shaders.ts:1045-1070) of the wrapped words
    auto sum_at = [&](int word, int node) -> long long {
        if (word < SKIP) return 0;
        return (long long)s_acc[word - SKIP][node];
    };
    auto decoded = [&](int word, int node) {
        const int32_t w32 = (int32_t)(uint32_t)(unsigned long long)sum_at(word, node);
        return ((float)w32 / FIXED_POINT_SCALE) * (word == 6 || word == 10 ? vref : mref);
    };
    const int lx = tid % T::ROW, ly = tid / T::ROW;
    const int sx = X0 - 1 + lx, sy = Y0 - 1 + ly;  // this thread's node (storage coords)
    if (lx >= 1 && lx <= TI && ly >= 1 && ly <= TI && sx >= 0 && sy >= 0 && sx < W && sy < H) {
        const int ring = ly * T::TC + lx;
        const uint32_t node = (uint32_t)sy * (uint32_t)W + (uint32_t)sx;
        float f[11];
#pragma unroll
        for (int k = 0; k < 11; ++k) {
            const long long sum = sum_at(k, ring);
            if (sum > 2147483647LL || sum < -2147483648LL) flags |= ERR_FIXED_POINT;
            if (keep_fixed) grid.fixed[(size_t)node * 11 + k] = (int32_t)(uint32_t)(unsigned long long)sum;
            f[k] = decoded(k, ring);
        }
        auto nvol = [&](int k, int ox, int oy) { return decoded(k == 0 ? 6 : 10, ring + oy * T::TC + ox); };
        if constexpr (AUX) {
            const AuxData& A = S.aux;
            flags |= update_node(P, grid, scene_flags, node, sx - 1, sy - 1, f, ERR_GRID, nvol,
                                 [&](uint32_t nd, float wm, float om, float2& vw, float2& vo) {
                                     return fluid_triangle_node(P, A, in, in_buf, nd, wm, om, vw, vo);
                                 });
        } else {
            flags |= update_node(P, grid, scene_flags, node, sx - 1, sy - 1, f, ERR_GRID, nvol);
        }
    }
    if (P.diag_on) {  // grid mass and momentum of this microstep (reference/solver.ts:481-492)
        const bool mine = lx >= 1 && lx <= TI && ly >= 1 && ly <= TI && sx >= 0 && sy >= 0 && sx < W && sy < H;
        const int ring = ly * T::TC + lx;
        block_sum3_f64(mine ? (double)decoded(0, ring) + (double)decoded(3, ring) + (double)decoded(7, ring) : 0.0,
                       mine ? (double)decoded(1, ring) + (double)decoded(4, ring) + (double)decoded(8, ring) : 0.0,
                       mine ? (double)decoded(2, ring) + (double)decoded(5, ring) + (double)decoded(9, ring) : 0.0,
                       diag_grid + 3 * tile);
    }
    if (flags) raise_step(ctl, flags, step);
}


template <int TI, bool GENERAL, bool SOLIDS, bool AUX = false>
__global__ __launch_bounds__(P2GTile<TI>::THREADS)
// LDS: 3 x 8 waves a CU for 20 with solids; with the stress cache and no
// solids, 80 registers at 3 blocks a CU (fewer scatter spills) beat 64 at 4
__attribute__((amdgpu_waves_per_eu((TI == 20 && (SOLIDS || GENERAL)) ? 6 : 8, 8)))
void k_p2g_fixed(const SceneDesc* __restrict__ scenes, const uint32_t* __restrict__ tile_scene, uint32_t nscenes,
                 uint32_t step, TileQueue q) {
    run_tiles(q, 0u, step, [&](uint32_t t) { p2g_fixed_tile<TI, GENERAL, SOLIDS, AUX>(scenes, tile_scene, nscenes, step, t); });
}

// -------------------------------------------------------------------- G2P
// integrateDeformation (shaders.ts:1459-1483)
__device__ __forceinline__ float4 integrate_deformation(float4 F, float4 C, float dt, uint32_t substeps) {
    const float h = substeps == 1u ? dt : dt / (float)substeps;  // dt / 1 is dt
    const float4 u = make_float4(1.0f + h * C.x, h * C.y, h * C.z, 1.0f + h * C.w);
    float4 r = F;
    for (uint32_t k = 0; k < substeps; ++k) {
        r = make_float4(fmaf(u.x, r.x, u.y * r.z), fmaf(u.x, r.y, u.y * r.w), fmaf(u.z, r.x, u.w * r.z),
                        fmaf(u.z, r.y, u.w * r.w));
    }
    return r;
}

// projectLiquidDeformation (shaders.ts:474-488).  The quotients and square
// roots are the hardware ones with one quotient correction (quot), within an
// ulp, where the WGSL allows 2.5 ulp; a non-finite or zero scale still gives
// a non-finite result, which the caller's finiteness test reports.
__device__ __forceinline__ float4 project_liquid(float4 m) {
    float nx = m.x + m.w, ny = m.z - m.y;
    float sc = fmaxf(fabsf(nx), fabsf(ny));
    float ax = quot(nx, sc), ay = quot(ny, sc);
    float len = __builtin_amdgcn_sqrtf(ax * ax + ay * ay);
    float rx = quot(ax, len), ry = quot(ay, len);
    float vol = det_comp(m);
    float bv = fminf(fmaxf(vol, MIN_LIQUID_DETERMINANT), MAX_LIQUID_DETERMINANT);
    float s = __builtin_amdgcn_sqrtf(bv);
    return make_float4(s * rx, s * -ry, s * ry, s * rx);
}

// particleWaveSpeedBound (shaders.ts:268-353)
template <bool GENERAL, bool SOLIDS = true>
__device__ float wave_speed_bound(const Params& P, uint32_t model, float4 F, float lambda, float mu, float density,
                                  float volume, float param) {
    if (model == MODEL_WATER || model == MODEL_OIL) {
        if (!finitef(lambda) || !finitef(density) || lambda <= 0.0f || density <= 0.0f) return FLT_MAX;
        float ws2 = quot(lambda, density);  // positive finite operands: within an ulp
        if (param > 0.0f) {
            if constexpr (!GENERAL) return FLT_MAX;
            float det = det_comp(F);
            if (!finitef(param) || param <= 0.0f || !finitef(det) || det <= 0.0f) return FLT_MAX;
            ws2 = param * ws2 * (float)tait_pow((double)(1.0f / det), param - 1.0f);
        }
        if (!finitef(ws2) || ws2 < 0.0f) return FLT_MAX;
        return __builtin_amdgcn_sqrtf(ws2);
    }
    if constexpr (!GENERAL || !SOLIDS) return FLT_MAX;
    if (model != MODEL_NEO_HOOKEAN) return FLT_MAX;
    if (!finitef(param)) return FLT_MAX;
    float remaining = 1.0f - fminf(fmaxf(param, 0.0f), 1.0f);
    if (remaining <= 0.0f) return 0.0f;
    float l = lambda * remaining, m = mu * remaining;
    float det = det_comp(F);
    float tr = F.x * F.x + F.y * F.y + F.z * F.z + F.w * F.w;
    if (!finitef(det) || det <= 0.0f || !finitef(tr) || !finitef(l) || !finitef(m) || m < 0.0f) return FLT_MAX;
    float disc = sqrtf(fmaxf(0.0f, tr * tr - 4.0f * det * det));
    float smax = 0.5f * (tr + disc);
    float vt = fabsf(l + m - l * (float)log((double)det));
    float rho = P.particle_mass / volume;
    float ws2 = (m * smax + vt) / rho;
    if (!finitef(rho) || rho <= 0.0f || !finitef(ws2) || ws2 < 0.0f) return FLT_MAX;
    return sqrtf(ws2);
}

// confineAxis (shaders.ts:522-534) as selects: the tests are combined with
// bitwise operators, so the compiler forms no divergent branch for them.
__device__ __forceinline__ float2 confine_axis(float x, float v, float maximum) {
    const bool low = (x < 0.0f) | ((x == 0.0f) & (v < 0.0f));
    const bool high = !low & ((x > maximum) | ((x == maximum) & (v > 0.0f)));
    return make_float2(low ? 0.0f : (high ? maximum : x), (low | high) ? 0.0f : v);
}

// g2p (shaders.ts:2114-2421) for one particle with the grid patch in LDS,
// plus validateCflParticle (:2423-2454).  Returns false when the particle
// raised a flag (the microstep then rolls back).
struct G2POut { float2 pos, vel; float4 C, F; float plastic; };

// projectParticleOutsideTriangles (shaders.ts:1801-1854): a liquid particle
// of radius r whose interpolated signed distance to the nearest triangle is
// below r moves out along the interpolated normal of that triangle.  In FP64:
// the stencil fraction of a float32 coordinate near 2^12 cells keeps only
// ~14 bits, which would move the interpolated distance by 1e-4 of a cell.
__device__ float2 project_out_of_triangles(const Params& P, const AuxData& A, float2 p, float r) {
    const double dx = P.dx, gx = (double)p.x / dx, gy = (double)p.y / dx;
    const double bxf = floor(gx - 0.5), byf = floor(gy - 0.5);
    if (!(bxf >= -1.0 && byf >= -1.0 && bxf + 2.0 <= (double)P.cells_x + 1.0 && byf + 2.0 <= (double)P.cells_y + 1.0))
        return p;
    const int bx = (int)bxf, by = (int)byf;
    const double fx = gx - bxf, fy = gy - byf;
    const double wx[3] = {0.5 * (1.5 - fx) * (1.5 - fx), 0.75 - (fx - 1.0) * (fx - 1.0), 0.5 * (fx - 0.5) * (fx - 0.5)};
    const double wy[3] = {0.5 * (1.5 - fy) * (1.5 - fy), 0.75 - (fy - 1.0) * (fy - 1.0), 0.5 * (fy - 0.5) * (fy - 0.5)};
    const double band = dx * TRIANGLE_FIELD_BAND_CELLS;
    double sd = 0.0, closest = INFINITY, snx = 0.0, sny = 0.0;
    uint32_t cb = 0u;
    const uint32_t W = P.node_w;
    for (int oy = 0; oy < 3; ++oy)
        for (int ox = 0; ox < 3; ++ox) {
            const uint32_t node = (uint32_t)(by + oy + 1) * W + (uint32_t)(bx + ox + 1);
            const uint32_t b = A.field_body[node];
            double d = band;
            if (b != 0u) {
                const FieldNode fn = A.field[node];
                d = fn.dist;
                if (d < closest) {
                    closest = d;
                    cb = b;
                    snx = fn.nx;
                    sny = fn.ny;
                }
            }
            sd += wx[ox] * wy[oy] * d;
        }
    if (cb == 0u || sd >= (double)r) return p;
    double nx = 0.0, ny = 0.0;
    for (int oy = 0; oy < 3; ++oy)
        for (int ox = 0; ox < 3; ++ox) {
            const uint32_t node = (uint32_t)(by + oy + 1) * W + (uint32_t)(bx + ox + 1);
            if (A.field_body[node] != cb) continue;
            const FieldNode fn = A.field[node];
            nx += wx[ox] * wy[oy] * fn.nx;
            ny += wx[ox] * wy[oy] * fn.ny;
        }
    const double len = sqrt(nx * nx + ny * ny);
    if (len > 1e-12 && isfinite(len)) {
        nx = nx / len;
        ny = ny / len;
    } else {
        nx = snx;
        ny = sny;
    }
    const double corr = (double)r - sd;
    return make_float2((float)((double)p.x + corr * nx), (float)((double)p.y + corr * ny));
}

// The G2P of one particle runs in three parts:
//   g2p_motion:  velocity and affine from the node patch, the liquid
//                density correction, and the unconfined new position;
//   g2p_deform:  deformation update with its doubling retry and the
//                constitutive projections;
//   g2p_cfl_ok:  the CFL guard of the new state.
// The new position, once confined, fixes the particle's next bucket, so a
// caller without bonded bodies reserves the output slot after g2p_motion and
// runs the other parts while that atomic is in flight.  Each part raises its
// own flags and returns false when the particle failed.
// Failure codes of the G2P parts, in the order the reference tests them; a
// particle reports its first failure (raise_g2p).
enum : uint32_t {
    G2P_OK = 0u,
    G2P_FAIL_LIQUID,       // liquid mass, volume or density not finite and positive: ERR_G2P, with the particle
    G2P_FAIL_G2P,          // velocity, affine, position or F not finite, or a model this variant lacks: ERR_G2P
    G2P_FAIL_SUBSTEPS,     // the deformation increment needs more than 64 substeps: ERR_SUBSTEP_LIMIT, with the particle
    G2P_FAIL_DEFORMATION,  // the updated F loses orientation: deformation record 1, ERR_DEFORMATION
    G2P_FAIL_PROJECTED,    // the projected F loses orientation: deformation record 2, ERR_PROJECTED_DEFORMATION
    G2P_FAIL_CFL,          // the CFL guard: deformation record 3, ERR_CFL_BOUND
    G2P_FAIL_STENCIL,      // a stencil outside the grid or in another bucket: ERR_STENCIL
};
// The first failure of a sequence of tests: `code` keeps the earliest.
__device__ __forceinline__ uint32_t first_failure(uint32_t code, bool failed, uint32_t kind) {
    return code != G2P_OK ? code : (failed ? kind : G2P_OK);
}
// Raise the status bits (and the deformation record) of a G2P failure code.
__device__ void raise_g2p(Control* ctl, uint32_t code, uint32_t pid, uint32_t step, const float (&rec)[4]) {
    switch (code) {
    case G2P_FAIL_LIQUID: raise_step_at(ctl, ERR_G2P, step, pid); break;
    case G2P_FAIL_G2P: raise_step(ctl, ERR_G2P, step); break;
    case G2P_FAIL_SUBSTEPS: raise_step_at(ctl, ERR_SUBSTEP_LIMIT, step, pid); break;
    case G2P_FAIL_DEFORMATION:
        record_failure(ctl, 1u, pid, rec[0], rec[1], rec[2], rec[3]);
        raise_step(ctl, ERR_DEFORMATION, step);
        break;
    case G2P_FAIL_PROJECTED:
        record_failure(ctl, 2u, pid, rec[0], rec[1], rec[2], rec[3]);
        raise_step(ctl, ERR_PROJECTED_DEFORMATION, step);
        break;
    case G2P_FAIL_CFL:
        record_failure(ctl, 3u, pid, rec[0], rec[1], rec[2], rec[3]);
        raise_step(ctl, ERR_CFL_BOUND, step);
        break;
    case G2P_FAIL_STENCIL: raise_step(ctl, ERR_STENCIL, step); break;
    default: break;
    }
}

// The node values come from `lv` (the particle's phase velocity) and `llq`
// (liquid mass and volume), node (x, y) at (x - x0) + (y - y0) * stride.
// Returns G2P_OK or the first failure code; on a failure the outputs are
// not meaningful (every test is evaluated, none branches).
// LIQUID = true: every particle of the scene is water or oil (the liquid
// terms then need no test of the model).
template <class NodePtr, bool LIQUID = false>
__device__ __forceinline__ uint32_t g2p_motion_code(const Params& P, const Material& mt, float2 p, float4 F,
                                                    const Stencil& s, NodePtr lv, NodePtr llq, int stride, int x0,
                                                    int y0, float2& vel, float4& Cout, float2& np_out) {
    const uint32_t model = mt.model;
    const bool liquid = LIQUID || model == MODEL_WATER || model == MODEL_OIL;
    float vx = 0.f, vy = 0.f, a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, pm = 0.f, pvol = 0.f;
    const float dx = P.dx;
    // The reference tests every weight, offset and node value for
    // finiteness here (shaders.ts:2165-2192).  When this kernel runs, the
    // same step's P2G has accepted every weight of the particle and the grid
    // update has accepted every node value it wrote (finite velocities,
    // finite non-negative liquid mass and volume); nodes outside the grid are
    // staged as zero.  The tests cannot fail, so they are not repeated.
    float wx[3], wy[3], dpx[3], dpy[3];
#pragma unroll
    for (int o = 0; o < 3; ++o) {
        wx[o] = qweight(s.fx, o);
        wy[o] = qweight(s.fy, o);
        dpx[o] = ((float)o - s.fx) * dx;
        dpy[o] = ((float)o - s.fy) * dx;
    }
    const int row0 = (s.by + 1 - y0) * stride + (s.bx + 1 - x0);  // stencil offset (0, 0)
#pragma unroll
    for (int oy = 0; oy < 3; ++oy) {
#pragma unroll
        for (int ox = 0; ox < 3; ++ox) {
            const float w = wx[ox] * wy[oy];
            const float2 g = lv[row0 + oy * stride + ox];
            const float wgx = w * g.x, wgy = w * g.y;
            vx = vx + wgx;
            vy = vy + wgy;
            // fused: one rounding per term where the WGSL sum may take two
            a0 = fmaf(wgx, dpx[ox], a0);
            a1 = fmaf(wgx, dpy[oy], a1);
            a2 = fmaf(wgy, dpx[ox], a2);
            a3 = fmaf(wgy, dpy[oy], a3);
        }
    }
    uint32_t code = G2P_OK;
    if (liquid) {
#pragma unroll
        for (int oy = 0; oy < 3; ++oy) {
#pragma unroll
            for (int ox = 0; ox < 3; ++ox) {
                const float w = wx[ox] * wy[oy];
                const float2 lq = llq[row0 + oy * stride + ox];
                pm = fmaf(w, lq.x, pm);
                pvol = fmaf(w, lq.y, pvol);
            }
        }
        const float pdet = det_comp(F);
        code = first_failure(code,
                             !finitef(pm) || pm <= 0.0f || !finitef(pvol) || pvol <= 0.0f || !finitef(pdet) ||
                                 pdet <= 0.0f,
                             G2P_FAIL_LIQUID);
        // (quotients of finite positive values: quot, within an ulp)
        const float pd = quot(mt.density, pdet);
        const float ld = quot(pm, pvol);
        const float rel = quot(pd - ld, pd);
        vx = vx + P.dt * (rel * P.gx);
        vy = vy + P.dt * (rel * P.gy);
        code = first_failure(code,
                             !finitef(pd) || pd <= 0.0f || !finitef(ld) || ld <= 0.0f || !finitef(rel) ||
                                 !finitef(vx) || !finitef(vy),
                             G2P_FAIL_LIQUID);
    }
    const float gs = P.grad_scale, dt = P.dt;
    const float4 C = make_float4(a0 * gs, a1 * gs, a2 * gs, a3 * gs);
    const float2 np = make_float2(p.x + dt * vx, p.y + dt * vy);
    code = first_failure(code, !finitef(vx) || !finitef(vy) || !finite4(C) || !finite2(np), G2P_FAIL_G2P);
    vel = make_float2(vx, vy);
    Cout = C;
    np_out = np;
    return code;
}

template <class NodePtr>
__device__ __forceinline__ bool g2p_motion(const Params& P, Control* ctl, const Material& mt, uint32_t pid, float2 p,
                                           float4 F, const Stencil& s, NodePtr lv, NodePtr llq, int stride, int x0,
                                           int y0, uint32_t step, float2& vel, float4& Cout, float2& np_out) {
    const uint32_t code = g2p_motion_code(P, mt, p, F, s, lv, llq, stride, x0, y0, vel, Cout, np_out);
    if (code != G2P_OK) {
        const float rec[4] = {0.f, 0.f, 0.f, 0.f};
        raise_g2p(ctl, code, pid, step, rec);
        return false;
    }
    return true;
}

template <bool GENERAL, bool SOLIDS = true>
__device__ bool g2p_deform(const Params& P, Control* ctl, const Material& mt, uint32_t pid, float4 F, float4 C,
                           float pv, uint32_t step, float4& Fout, float& pv_out, double& ndet_out) {
    const uint32_t model = mt.model;
    const bool liquid = model == MODEL_WATER || model == MODEL_OIL;
    const float dt = P.dt;
    const float rowsum = fmaxf(fabsf(C.x) + fabsf(C.y), fabsf(C.z) + fabsf(C.w));
    const float increment = dt * rowsum;
    const float required = ceilf(increment * (1.0f / MAX_DEFORMATION_INCREMENT));  // exact: a power of two
    if (!finitef(required) || required > (float)MAX_DEFORMATION_SUBSTEPS) {
        raise_step_at(ctl, ERR_SUBSTEP_LIMIT, step, pid);
        return false;
    }
    uint32_t substeps = required < 1.0f ? 1u : (uint32_t)required;
    float4 trial = integrate_deformation(F, C, dt, substeps);
    if (!finite4(trial)) {
        raise_step(ctl, ERR_G2P, step);
        return false;
    }
    double tdet = det_f64(trial);
    // A near-singular F can lose orientation to f32 rounding; retry that update
    // with doubled substeps (shaders.ts:2274-2286).  The sign test is FP64.
    while (isfinite(tdet) && tdet <= 0.0 && substeps < MAX_DEFORMATION_SUBSTEPS) {
        substeps = min(substeps * 2u, MAX_DEFORMATION_SUBSTEPS);
        trial = integrate_deformation(F, C, dt, substeps);
        if (!finite4(trial)) {
            raise_step(ctl, ERR_G2P, step);
            return false;
        }
        tdet = det_f64(trial);
    }
    if (!isfinite(tdet) || tdet <= 0.0) {
        record_failure(ctl, 1u, pid, (float)tdet, increment, (float)det_f64(F), 0.0f);
        raise_step(ctl, ERR_DEFORMATION, step);
        return false;
    }
    float4 Fn = trial;
    float pvn = pv;
    if (liquid) {
        Fn = project_liquid(trial);
    } else if constexpr (!GENERAL || !SOLIDS) {
        raise_step(ctl, ERR_G2P, step);
        return false;
    } else if (model == MODEL_SNOW) {
        Decomp d;
        if (!decompose({trial.x, trial.y, trial.z, trial.w}, d)) {
            raise_step(ctl, ERR_G2P, step);
            return false;
        }
        const double lo = 1 - SNOW_COMPRESSION_YIELD, hi = 1 + SNOW_STRETCH_YIELD;
        const double s1 = fmax(lo, fmin(hi, d.s1)), s2 = fmax(lo, fmin(hi, d.s2));
        const double upv = fmax(SNOW_MIN_PLASTIC_VOLUME, fmin(SNOW_MAX_PLASTIC_VOLUME, ((double)pv * tdet) / (s1 * s2)));
        const M2 r = reconstruct(d, s1, s2);
        Fn = make_float4((float)r.a, (float)r.b, (float)r.c, (float)r.d);
        pvn = (float)upv;
    } else if (model == MODEL_POWDER) {
        if (!finitef(mt.lambda) || !finitef(mt.mu) || mt.mu <= 0.0f || !finitef(mt.param) || mt.param < 0.0f) {
            raise_step(ctl, ERR_G2P, step);
            return false;
        }
        Decomp d;
        if (!decompose({trial.x, trial.y, trial.z, trial.w}, d)) {
            raise_step(ctl, ERR_G2P, step);
            return false;
        }
        const double l1 = log(d.s1), l2 = log(d.s2), lv2 = l1 + l2;
        const double minlv = log(MIN_POWDER_DETERMINANT);
        const double blv = fmax(minlv, lv2);
        const double pressure = fmax(0.0, -((double)mt.lambda + (double)mt.mu) * blv);
        const double dev = 0.5 * (l1 - l2);
        const double maxdev = ((double)mt.param * pressure) / (2 * (double)mt.mu);
        if (!(dev <= maxdev && lv2 >= minlv)) {
            const double pdev = fmin(dev, maxdev), mean = 0.5 * blv;
            const M2 r = reconstruct(d, exp(mean + pdev), exp(mean - pdev));
            Fn = make_float4((float)r.a, (float)r.b, (float)r.c, (float)r.d);
        }
    } else if (model == MODEL_NEO_HOOKEAN) {
        const float tr = trial.x * trial.x + trial.y * trial.y + trial.z * trial.z + trial.w * trial.w;
        if (!finitef(tr)) {
            raise_step(ctl, ERR_G2P, step);
            return false;
        }
        if (tdet < MIN_NEO_HOOKEAN_STRETCH * sqrt((double)tr) ||
            (double)tr > MAX_NEO_HOOKEAN_STRETCH * MAX_NEO_HOOKEAN_STRETCH) {
            Decomp d;
            if (!decompose({trial.x, trial.y, trial.z, trial.w}, d)) {
                raise_step(ctl, ERR_G2P, step);
                return false;
            }
            const double s1 = fmax(MIN_NEO_HOOKEAN_STRETCH, fmin(MAX_NEO_HOOKEAN_STRETCH, d.s1));
            const double s2 = fmax(MIN_NEO_HOOKEAN_STRETCH, fmin(MAX_NEO_HOOKEAN_STRETCH, d.s2));
            const M2 r = reconstruct(d, s1, s2);
            Fn = make_float4((float)r.a, (float)r.b, (float)r.c, (float)r.d);
        }
    } else {
        raise_step(ctl, ERR_G2P, step);
        return false;
    }
    const double ndet = det_f64(Fn);
    if (!finite4(Fn) || !finitef(pvn) || pvn < 0.6f || pvn > 20.0f) {
        raise_step(ctl, ERR_G2P, step);
        return false;
    }
    if (!isfinite(ndet) || ndet <= 0.0) {
        record_failure(ctl, 2u, pid, Fn.x, Fn.y, Fn.z, Fn.w);
        raise_step(ctl, ERR_PROJECTED_DEFORMATION, step);
        return false;
    }
    Fout = Fn;
    pv_out = pvn;
    ndet_out = ndet;
    return true;
}

// validateCflParticle (shaders.ts:2423-2454) of the new velocity and F, as a
// failure code; `rec` receives the deformation record of a failure.
template <bool GENERAL, bool SOLIDS = true>
__device__ __forceinline__ uint32_t g2p_cfl_code(const Params& P, const Material& mt, float2 vel, float4 Fn,
                                                 float (&rec)[4]) {
    const float speed = __builtin_amdgcn_sqrtf(vel.x * vel.x + vel.y * vel.y);
    const float ws = wave_speed_bound<GENERAL, SOLIDS>(P, mt.model, Fn, mt.lambda, mt.mu, mt.density, mt.volume,
                                                       mt.param);
    rec[0] = speed;
    rec[1] = ws;
    rec[2] = P.vel_bound;
    rec[3] = P.wave_bound;
    const bool bad = !finite2(vel) || !finitef(speed) || !finitef(ws) || speed > P.vel_bound || ws > P.wave_bound;
    return bad ? G2P_FAIL_CFL : G2P_OK;
}
template <bool GENERAL, bool SOLIDS = true>
__device__ bool g2p_cfl_ok(const Params& P, Control* ctl, const Material& mt, uint32_t pid, float2 vel, float4 Fn,
                           uint32_t step) {
    float rec[4];
    const uint32_t code = g2p_cfl_code<GENERAL, SOLIDS>(P, mt, vel, Fn, rec);
    if (code != G2P_OK) {
        raise_g2p(ctl, code, pid, step, rec);
        return false;
    }
    return true;
}

// g2p_deform for a liquid particle (integrateDeformation with the doubling
// retry, then projectLiquidDeformation) as a failure code: the tests are
// evaluated in the reference's order without branching (the retry loop
// aside, which only a lane that lost orientation enters), and `rec`
// receives the deformation record of a failure: (det of the update, the
// increment, det F) or the projected F.
__device__ __forceinline__ uint32_t g2p_deform_liquid_code(const Params& P, const Material& mt, float4 F, float4 C,
                                                           float pv, float4& Fout, float (&rec)[4]) {
    const float dt = P.dt;
    const float rowsum = fmaxf(fabsf(C.x) + fabsf(C.y), fabsf(C.z) + fabsf(C.w));
    const float increment = dt * rowsum;
    const float required = ceilf(increment * (1.0f / MAX_DEFORMATION_INCREMENT));  // exact: a power of two
    uint32_t code = first_failure(G2P_OK, !finitef(required) || required > (float)MAX_DEFORMATION_SUBSTEPS,
                                  G2P_FAIL_SUBSTEPS);
    uint32_t substeps = (code != G2P_OK || required < 1.0f) ? 1u : (uint32_t)required;
    float4 trial = integrate_deformation(F, C, dt, substeps);
    bool finite_trial = finite4(trial);
    double tdet = det_f64(trial);
    // A near-singular F can lose orientation to f32 rounding; retry that
    // update with doubled substeps (shaders.ts:2274-2286).  The sign test is FP64.
    while (code == G2P_OK && finite_trial && isfinite(tdet) && tdet <= 0.0 && substeps < MAX_DEFORMATION_SUBSTEPS) {
        substeps = min(substeps * 2u, MAX_DEFORMATION_SUBSTEPS);
        trial = integrate_deformation(F, C, dt, substeps);
        finite_trial = finite4(trial);
        tdet = det_f64(trial);
    }
    code = first_failure(code, !finite_trial, G2P_FAIL_G2P);
    const bool lost = code == G2P_OK && (!isfinite(tdet) || tdet <= 0.0);
    code = first_failure(code, lost, G2P_FAIL_DEFORMATION);
    code = first_failure(code, mt.model != MODEL_WATER && mt.model != MODEL_OIL, G2P_FAIL_G2P);
    const float4 Fn = project_liquid(trial);
    const double ndet = det_f64(Fn);
    code = first_failure(code, !finite4(Fn) || !finitef(pv) || pv < 0.6f || pv > 20.0f, G2P_FAIL_G2P);
    code = first_failure(code, !isfinite(ndet) || ndet <= 0.0, G2P_FAIL_PROJECTED);
    rec[0] = lost ? (float)tdet : Fn.x;
    rec[1] = lost ? increment : Fn.y;
    rec[2] = lost ? (float)det_f64(F) : Fn.z;
    rec[3] = lost ? 0.0f : Fn.w;
    Fout = Fn;
    return code;
}

// The whole G2P of one particle in the reference's order, for scenes with
// bonded bodies, whose liquid particles leave the triangles after the
// deformation update (the projection radius needs the new det F).
template <bool GENERAL, bool AUX, bool SOLIDS, class NodePtr>
__device__ bool g2p_particle(const Params& P, Control* ctl, const Material& mt, uint32_t pid, float2 p, float4 F,
                             float pv, const Stencil& s, NodePtr lv, NodePtr llq, int stride, int x0, int y0,
                             uint32_t step, G2POut& o, const AuxData* A = nullptr, bool cfl_here = false) {
    float2 vel, np;
    float4 C, Fn;
    float pvn;
    double ndet;
    if (!g2p_motion(P, ctl, mt, pid, p, F, s, lv, llq, stride, x0, y0, step, vel, C, np)) return false;
    if (!g2p_deform<GENERAL, SOLIDS>(P, ctl, mt, pid, F, C, pv, step, Fn, pvn, ndet)) return false;
    const bool liquid = mt.model == MODEL_WATER || mt.model == MODEL_OIL;
    // projectLiquidFromTriangles is the identity without bonded bodies; the
    // second confineAxis pass is then idempotent (shaders.ts:2376-2408).
    float2 bxv = confine_axis(np.x, vel.x, P.dom_w);
    float2 byv = confine_axis(np.y, vel.y, P.dom_h);
    if constexpr (AUX) {
        if (liquid && A->nbodies != 0u) {
            const float r = (float)sqrt(((double)mt.volume * ndet) / 3.14159265358979323846);
            const float2 q = project_out_of_triangles(P, *A, make_float2(bxv.x, byv.x), r);
            if (!finite2(q)) {
                raise_step(ctl, ERR_TRIANGLE_CONTACT, step);
                return false;
            }
            bxv = confine_axis(q.x, bxv.y, P.dom_w);
            byv = confine_axis(q.y, byv.y, P.dom_h);
        }
    }
    o.pos = make_float2(bxv.x, byv.x);
    o.vel = make_float2(bxv.y, byv.y);
    o.C = C;
    o.F = Fn;
    o.plastic = pvn;
    if (P.cfl_on >= 0.5f && (P.body_count == 0u || cfl_here))
        return g2p_cfl_ok<GENERAL, SOLIDS>(P, ctl, mt, pid, o.vel, Fn, step);
    return true;
}

// One block of G_THREADS threads per G_TB x G_TB-bucket tile: stage the
// tile's node patch, run G2P for its particles and write each into its next
// bucket.  The inputs of a thread's next particle are fetched before the
// current one is processed, so their latency overlaps the current G2P and
// slot reservation.  Compiled shapes: 8x8 buckets with 128, 256 or 512
// threads, 4x4 buckets with 128 or 256.
template <int G_TB, int G_THREADS, bool GENERAL, bool AUX, bool SOLIDS>
__device__ __forceinline__ void g2p_tile(const SceneDesc* __restrict__ scenes, const uint32_t* __restrict__ tile_scene,
                                         uint32_t nscenes, uint32_t step, uint32_t global_tile) {
    using T = G2PTile<G_TB>;
    constexpr int G_TN = T::TN, NBK = T::NBK;
    static_assert(G_THREADS >= 64, "the first wave scans the bucket counts");
    SCENE_AT(tile_scene, g2p_base, global_tile);
    const int out_buf = (int)((S.start_buf + step + 1u) & 1u);
    const Params* __restrict__ Pp = &S.P;
    Control* ctl = S.ctl;
    const SlotBuf in = S.buf[out_buf ^ 1];
    const SlotBuf out = S.buf[out_buf];
    const Material* __restrict__ mats = S.mats;
    const Grid grid = S.grid;
    const uint32_t nmat = S.nmat, tiles_x = S.gtx, bw = S.bw, bh = S.bh, cap = S.cap, scene_flags = S.scene_flags;
    __shared__ float2 s_vs[GENERAL ? G_TN * G_TN : 1];  // scenes without solids (not GENERAL) never read it
    __shared__ float2 s_vw[G_TN * G_TN];
    __shared__ float2 s_vo[G_TN * G_TN];
    __shared__ float2 s_lq[G_TN * G_TN];
    __shared__ uint32_t s_boff[NBK + 1];
    __shared__ Material s_mat[LDS_MATS];
    // tile place -> (bucket of the tile << 16 | place in the bucket), for
    // tiles of at most G2P_MAP particles (larger ones search s_boff), in
    // scenes without the stress cache (measured: the map slows the Tait
    // variant, whose registers already spill) and where the block stays
    // within 40 KB of LDS (four blocks a CU)
    constexpr int PATCH_BYTES = 3 * G_TN * G_TN * (int)sizeof(float2);
    constexpr uint32_t G2P_MAP =
        !GENERAL && PATCH_BYTES + 8192 + (int)sizeof(Material) * LDS_MATS <= 40960 ? 2048u : 0u;
    __shared__ uint32_t s_map[G2P_MAP > 0u ? G2P_MAP : 1u];
    if (block_status_set(ctl)) return;
    const Params& P = *Pp;
    const int tid = threadIdx.x;
    const int tbx = (int)(tile % tiles_x) * G_TB, tby = (int)(tile / tiles_x) * G_TB;
    const int nx0 = tbx * BUCKET_CELLS, ny0 = tby * BUCKET_CELLS;  // first node (= first cell) of the patch
    const int W = (int)P.node_w, H = (int)P.node_h;
    stage_materials(s_mat, mats, nmat);
    if (tid < 64) {
        const int gbx = tbx + tid % G_TB, gby = tby + tid / G_TB;
        uint32_t c = 0u;
        if (tid < NBK && gbx < (int)bw && gby < (int)bh) c = min(in.cnt[(uint32_t)gby * bw + (uint32_t)gbx], cap);
        // exclusive scan of the tile's counts in one wave
        uint32_t incl = c;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
            const uint32_t t = __shfl_up(incl, off, 64);
            if (tid >= off) incl += t;
        }
        if (tid < NBK) s_boff[tid] = incl - c;
        if (tid == NBK - 1) s_boff[NBK] = incl;
    }
    __syncthreads();
    if (s_boff[NBK] == 0u) return;  // no particle in this tile
    const uint32_t total = s_boff[NBK];

    // inputs of the thread's current (c_) and next (n_) particle
    float2 c_p = make_float2(0.f, 0.f), n_p = c_p;
    float4 c_F = make_float4(0.f, 0.f, 0.f, 0.f), n_F = c_F;
    uint32_t c_pid = 0u, c_mat = 0u, n_pid = 0u, n_mat = 0u;
    float c_pv = 1.0f, n_pv = 1.0f;
    int c_gbx = 0, c_gby = 0, n_gbx = 0, n_gby = 0;
    const bool mapped = total <= G2P_MAP;
#define G2P_FETCH(J, PFX, MAPPED)                                                      \
    do {                                                                               \
        const uint32_t j_ = (J);                                                       \
        if (j_ < total) {                                                              \
            int lo = 0;                                                                \
            uint32_t k_ = 0;                                                           \
            if (MAPPED) {                                                              \
                const uint32_t e_ = s_map[j_];                                         \
                lo = (int)(e_ >> 16);                                                  \
                k_ = e_ & 0xffffu;                                                     \
            } else {                                                                   \
                int hi = NBK - 1;                                                      \
                while (lo < hi) {                                                      \
                    const int mid = (lo + hi + 1) >> 1;                                \
                    if (s_boff[mid] <= j_) lo = mid; else hi = mid - 1;                \
                }                                                                      \
                k_ = j_ - s_boff[lo];                                                  \
            }                                                                          \
            PFX##gbx = tbx + lo % G_TB;                                                \
            PFX##gby = tby + lo / G_TB;                                                \
            const size_t g_ = ((size_t)PFX##gby * bw + (size_t)PFX##gbx) * cap + k_;  \
            PFX##p = in.pos[g_];                                                       \
            PFX##F = in.F[g_];                                                         \
            PFX##pid = in.pid[g_];                                                     \
            PFX##mat = in.mat[g_];                                                     \
            if constexpr (GENERAL) PFX##pv = in.plastic[g_];                           \
        }                                                                              \
    } while (0)
    // The first particle's inputs load while the node patch is staged.  The
    // patch loads of a thread are all issued before any is written to LDS
    // (in groups of up to five nodes), so the staging costs one memory
    // round trip per group instead of one per node.
    G2P_FETCH((uint32_t)tid, c_, false);
    if (mapped) {  // each bucket's places by G_THREADS / NBK threads (read after the barrier below)
        constexpr int TPB = G_THREADS / NBK;
        static_assert(TPB >= 1 && G_THREADS % NBK == 0, "whole threads a bucket");
        const int b = tid / TPB;
        const uint32_t b0 = s_boff[b], n = s_boff[b + 1] - b0;
        for (uint32_t k = (uint32_t)(tid % TPB); k < n; k += TPB) s_map[b0 + k] = (uint32_t)b << 16 | k;
    }
    {
        constexpr int NODES = G_TN * G_TN, PER = (NODES + G_THREADS - 1) / G_THREADS, GROUP = PER < 5 ? PER : 5;
        const bool has_s = GENERAL && (scene_flags & SCENE_SOLID), has_w = scene_flags & SCENE_WATER,
                   has_o = scene_flags & SCENE_OIL, has_l = scene_flags & (SCENE_WATER | SCENE_OIL);
#pragma unroll
        for (int r0 = 0; r0 < PER; r0 += GROUP) {
            float2 vs[GROUP], vw[GROUP], vo[GROUP], lq[GROUP];
#pragma unroll
            for (int r = 0; r < GROUP; ++r) {
                const int e = tid + (r0 + r) * G_THREADS;
                const int nx = nx0 + e % G_TN, ny = ny0 + e / G_TN;
                vs[r] = make_float2(0.f, 0.f);
                vw[r] = vs[r];
                vo[r] = vs[r];
                lq[r] = vs[r];
                if (r0 + r < PER && e < NODES && nx < W && ny < H) {
                    const uint32_t node = (uint32_t)ny * (uint32_t)W + (uint32_t)nx;
                    if (has_s) vs[r] = grid.v_solid[node];
                    if (has_w) vw[r] = grid.v_water[node];
                    if (has_o) vo[r] = grid.v_oil[node];
                    if (has_l) lq[r] = grid.liquid[node];
                }
            }
#pragma unroll
            for (int r = 0; r < GROUP; ++r) {
                const int e = tid + (r0 + r) * G_THREADS;
                if (r0 + r < PER && e < NODES) {
                    if (has_s) s_vs[e] = vs[r];
                    if (has_w) s_vw[e] = vw[r];
                    if (has_o) s_vo[e] = vo[r];
                    if (has_l) s_lq[e] = lq[r];
                }
            }
        }
    }
    __syncthreads();

    const AuxData* const Ax = &S.aux;
    uint32_t* const S_aux_slot_out = AUX ? S.aux.slot[out_buf] : nullptr;
    for (uint32_t base = 0; base < total; base += G_THREADS) {
        bool ok = base + (uint32_t)tid < total;
        G2POut o;
        uint32_t dest = 0;
        float4 S = make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t aux = NO_AUX;
        uint32_t k;
        if constexpr (AUX || (GENERAL && SOLIDS)) {
            // the reference's order: the slot is reserved once the particle
            // is complete (with bonded bodies the projection out of the
            // triangles moves the position after the deformation update; in
            // scenes with the stress cache the early reservation would hold
            // more registers than four waves a SIMD allow)
            G2P_FETCH(base + G_THREADS + (uint32_t)tid, n_, mapped);
            if (ok) {
                Material mt = material_at(s_mat, mats, c_mat);
                bool cfl_here = false;
                if constexpr (AUX) {
                    aux = Ax->aux_of[c_pid];
                    // fracture endpoints carry their own damage (the neo-Hookean parameter)
                    if (aux != NO_AUX && aux >= Ax->frac_aux0) mt.param = Ax->damage[out_buf ^ 1][aux];
                    // triangle vertices are checked after the body contacts (validateCflParticles)
                    cfl_here = aux == NO_AUX || aux >= Ax->frac_aux0;
                }
                Stencil s;
                if (!make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, c_p, s) || ((s.bx + 1) >> 2) != c_gbx ||
                    ((s.by + 1) >> 2) != c_gby) {
                    raise_step(ctl, ERR_STENCIL, step);
                    ok = false;
                } else {
                    const float2* lv = mt.phase == 1u ? s_vw : (mt.phase == 2u || !GENERAL ? s_vo : s_vs);
                    ok = g2p_particle<GENERAL, AUX, SOLIDS>(P, ctl, mt, c_pid, c_p, c_F, c_pv, s, lv,
                                                            (const float2*)s_lq, G_TN, nx0, ny0, step, o, Ax,
                                                            cfl_here);
                }
                if (ok) {
                    Stencil ns;
                    if (!make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, o.pos, ns)) {
                        raise_step(ctl, ERR_STENCIL, step);
                        ok = false;
                    } else {
                        dest = bucket_of(ns, bw);
                    }
                }
                if constexpr (GENERAL) {
                    if (ok) {
                        // stress cache of the new state; failures belong to the
                        // microstep that will start from it
                        const uint32_t pf = particle_stress<true, SOLIDS>(mt, o.F, o.plastic, det_comp(o.F), S);
                        if (pf) {
                            if (pf & ERR_P2G) atomicMin(&ctl->deferred_index[out_buf], c_pid);
                            atomicOr(&ctl->deferred[out_buf], pf);
                        }
                    }
                }
            }
            k = wave_alloc(out.cnt, dest, ok);
        } else if constexpr (GENERAL) {
            // Liquid scenes with the stress cache (Tait water or oil): the
            // reference's order with the slot reserved once the particle is
            // complete (an early reservation would hold more registers than
            // four waves a SIMD allow), the tests accumulated into a
            // first-failure code as in liquid-only scenes and raised once.
            G2P_FETCH(base + G_THREADS + (uint32_t)tid, n_, mapped);
            if (ok) {
                const Material mt = material_at(s_mat, mats, c_mat);
                uint32_t code = G2P_OK;
                float rec[4] = {0.f, 0.f, 0.f, 0.f};
                Stencil s;
                if (!make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, c_p, s) || ((s.bx + 1) >> 2) != c_gbx ||
                    ((s.by + 1) >> 2) != c_gby) {
                    code = G2P_FAIL_STENCIL;
                } else {
                    const float2* lv = mt.phase == 1u ? s_vw : s_vo;
                    float2 np;
                    code = g2p_motion_code<const float2*, true>(P, mt, c_p, c_F, s, lv, (const float2*)s_lq, G_TN,
                                                                 nx0, ny0, o.vel, o.C, np);
                    if (code == G2P_OK) code = g2p_deform_liquid_code(P, mt, c_F, o.C, c_pv, o.F, rec);
                    o.plastic = c_pv;
                    // the second confineAxis pass is idempotent without bodies
                    const float2 bxv = confine_axis(np.x, o.vel.x, P.dom_w);
                    const float2 byv = confine_axis(np.y, o.vel.y, P.dom_h);
                    o.pos = make_float2(bxv.x, byv.x);
                    o.vel = make_float2(bxv.y, byv.y);
                    if (P.cfl_on >= 0.5f && P.body_count == 0u && code == G2P_OK)
                        code = g2p_cfl_code<GENERAL, SOLIDS>(P, mt, o.vel, o.F, rec);
                    Stencil ns;
                    const bool placed =
                        code == G2P_OK && make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, o.pos, ns);
                    dest = placed ? bucket_of(ns, bw) : 0u;
                    code = first_failure(code, !placed, G2P_FAIL_STENCIL);
                }
                if (code != G2P_OK) {
                    raise_g2p(ctl, code, c_pid, step, rec);
                    ok = false;
                } else {
                    // stress cache of the new state; failures belong to the
                    // microstep that will start from it
                    const uint32_t pf = particle_stress<true, SOLIDS>(mt, o.F, o.plastic, det_comp(o.F), S);
                    if (pf) {
                        if (pf & ERR_P2G) atomicMin(&ctl->deferred_index[out_buf], c_pid);
                        atomicOr(&ctl->deferred[out_buf], pf);
                    }
                }
            }
            k = wave_alloc(out.cnt, dest, ok);
        } else {
            // The confined new position fixes the next bucket: its slot is
            // reserved here, and the atomic completes while the next
            // particle's inputs are fetched and this particle's deformation
            // update, projections and CFL guard run.  A particle that fails
            // after the reservation fails the microstep, whose output buffer
            // is then never read (rollback copies the committed buffer over
            // it, counts included).  The flags are those of the reference's
            // order: a new position outside the grid is reported only when
            // the rest of the particle's update succeeded.  The tests of the
            // particle's update accumulate a first-failure code without
            // branching; one branch raises it at the end.
            Material mt;
            bool placed = false;
            uint32_t code = G2P_OK;
            float rec[4] = {0.f, 0.f, 0.f, 0.f};
            if (ok) {
                mt = material_at(s_mat, mats, c_mat);
                Stencil s;
                if (!make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, c_p, s) || ((s.bx + 1) >> 2) != c_gbx ||
                    ((s.by + 1) >> 2) != c_gby) {
                    code = G2P_FAIL_STENCIL;
                } else {
                    const float2* lv = mt.phase == 1u ? s_vw : s_vo;
                    float2 np;
                    code = g2p_motion_code<const float2*, true>(P, mt, c_p, c_F, s, lv, (const float2*)s_lq, G_TN,
                                                                 nx0, ny0, o.vel, o.C, np);
                    // projectLiquidFromTriangles is the identity without
                    // bonded bodies; the second confineAxis pass is then
                    // idempotent (shaders.ts:2376-2408)
                    const float2 bxv = confine_axis(np.x, o.vel.x, P.dom_w);
                    const float2 byv = confine_axis(np.y, o.vel.y, P.dom_h);
                    o.pos = make_float2(bxv.x, byv.x);
                    o.vel = make_float2(bxv.y, byv.y);
                    Stencil ns;
                    placed = code == G2P_OK && make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, o.pos, ns);
                    dest = placed ? bucket_of(ns, bw) : 0u;
                }
            }
            const bool reserve = ok && code == G2P_OK && placed;
            const WaveAlloc wa = wave_alloc_issue(out.cnt, dest, reserve);
            G2P_FETCH(base + G_THREADS + (uint32_t)tid, n_, mapped);
            if (ok && code == G2P_OK) {
                code = g2p_deform_liquid_code(P, mt, c_F, o.C, c_pv, o.F, rec);
                o.plastic = c_pv;
                if (P.cfl_on >= 0.5f && P.body_count == 0u && code == G2P_OK)
                    code = g2p_cfl_code<GENERAL, SOLIDS>(P, mt, o.vel, o.F, rec);
                code = first_failure(code, !placed, G2P_FAIL_STENCIL);
            }
            k = wave_alloc_finish(wa, reserve);
            if (ok && code != G2P_OK) {
                raise_g2p(ctl, code, c_pid, step, rec);
                ok = false;
            }
        }
        if (ok) {
            if (k >= cap) {
                raise_step(ctl, ERR_INTERNAL_REBUCKET, step);
            } else {
                const size_t g = (size_t)dest * cap + k;
                out.pos[g] = o.pos;
                out.vel[g] = o.vel;
                out.C[g] = o.C;
                out.F[g] = o.F;
                out.pid[g] = c_pid;
                out.mat[g] = c_mat;
                if constexpr (GENERAL) {
                    out.plastic[g] = o.plastic;
                    out.S[g] = S;
                }
                if constexpr (AUX) {
                    if (aux != NO_AUX) S_aux_slot_out[aux] = (uint32_t)g;
                }
            }
        }
        c_p = n_p; c_F = n_F; c_pid = n_pid; c_mat = n_mat; c_pv = n_pv; c_gbx = n_gbx; c_gby = n_gby;
    }
#undef G2P_FETCH
}


template <int G_TB, int G_THREADS, bool GENERAL, bool AUX = false, bool SOLIDS = true>
__global__ __launch_bounds__(G_THREADS)
__attribute__((amdgpu_waves_per_eu(!GENERAL || !SOLIDS ? 4 : 1, 8)))  // liquid-only scenes: 4 waves a SIMD
void k_g2p(const SceneDesc* __restrict__ scenes, const uint32_t* __restrict__ tile_scene, uint32_t nscenes,
           uint32_t step, TileQueue q) {
    run_tiles(q, 1u, step, [&](uint32_t t) { g2p_tile<G_TB, G_THREADS, GENERAL, AUX, SOLIDS>(scenes, tile_scene, nscenes, step, t); });
}

// ------------------------------------------ config-2 bodies and fracture
// Scenes with bonded triangle bodies or fracture bonds run up to three more
// kernels per microstep around P2G and G2P: the triangle field before P2G,
// then, on the G2P output, the body contacts and the fracture update.  They
// cover one scene (such scenes are not batched).

// triangleDistanceSample (reference/solver.ts:1190-1291) in FP64: signed
// distance, outward normal and barycentric weights of the closest point.
__device__ bool tri_sample64(double px, double py, const double* t, double& sd, double& nx, double& ny,
                             double* bary) {
    const double ax = t[0], ay = t[1], bx = t[2], by = t[3], cx = t[4], cy = t[5];
    const double abx = bx - ax, aby = by - ay, acx = cx - ax, acy = cy - ay;
    const double area = abx * acy - aby * acx;
    if (!isfinite(area) || fabs(area) <= 1e-10) return false;
    const double bcx = cx - bx, bcy = cy - by, cax = ax - cx, cay = ay - cy;
    const double c_ab = abx * (py - ay) - aby * (px - ax);
    const double c_bc = bcx * (py - by) - bcy * (px - bx);
    const double c_ca = cax * (py - cy) - cay * (px - cx);
    const bool inside = area > 0 ? (c_ab >= -1e-10 && c_bc >= -1e-10 && c_ca >= -1e-10)
                                 : (c_ab <= 1e-10 && c_bc <= 1e-10 && c_ca <= 1e-10);
    double best = INFINITY, bnx = 0.0, bny = 0.0;
    bary[0] = 1.0;
    bary[1] = 0.0;
    bary[2] = 0.0;
#pragma unroll
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
            bnx = ddx * (1.0 / d);
            bny = ddy * (1.0 / d);
        } else {
            const double o = area > 0 ? 1.0 : -1.0, el = sqrt(len2);
            bnx = (o * dy) * (1.0 / el);
            bny = (-o * dx) * (1.0 / el);
        }
    }
    const double d = sqrt(best);
    if (!isfinite(d) || !isfinite(bnx) || !isfinite(bny)) return false;
    sd = inside ? -d : d;
    nx = bnx;
    ny = bny;
    return true;
}

__device__ __forceinline__ void load_triangle(const SlotBuf& b, const uint32_t* slot, uint32_t body, double* t) {
#pragma unroll
    for (int v = 0; v < 3; ++v) {
        const float2 p = b.pos[slot[3u * body + (uint32_t)v]];
        t[2 * v] = p.x;
        t[2 * v + 1] = p.y;
    }
}

constexpr int FIELD_THREADS = 256;
constexpr int FIELD_CHUNK = 256;  // triangles staged in LDS at a time

// buildTriangleField (reference/solver.ts:1064-1165; shaders.ts:1324-1436):
// for every node, the nearest triangle of the step's input state whose
// signed distance is within the band (ties keep the first body), and a
// cleared contact impulse.
__global__ __launch_bounds__(FIELD_THREADS) void k_tri_field(const SceneDesc* __restrict__ scenes, uint32_t step) {
    const SceneDesc& S = scenes[0];
    Control* ctl = S.ctl;
    if (block_status_set(ctl)) return;
    const Params& P = S.P;
    const AuxData& A = S.aux;
    const int in_buf = (int)((S.start_buf + step) & 1u);
    const SlotBuf in = S.buf[in_buf];
    __shared__ double s_tri[FIELD_CHUNK][6];
    __shared__ double s_box[FIELD_CHUNK][4];
    __shared__ uint32_t s_near[FIELD_CHUNK];  // bodies of the chunk whose band box meets this block's rows
    __shared__ uint32_t s_nnear;
    const uint32_t node = blockIdx.x * FIELD_THREADS + threadIdx.x;
    const bool live = node < P.node_count;
    const double dx = P.dx, band = (double)P.dx * TRIANGLE_FIELD_BAND_CELLS;
    const double px = (double)((int)(node % P.node_w) - 1) * dx, py = (double)((int)(node / P.node_w) - 1) * dx;
    double best = INFINITY, bnx = 0.0, bny = 0.0, bb[3] = {0.0, 0.0, 0.0};
    uint32_t best_body = 0u, err = 0u, bad_body = NO_AUX;
    for (uint32_t c0 = 0; c0 < A.nbodies; c0 += FIELD_CHUNK) {
        __syncthreads();
        for (uint32_t b = threadIdx.x; b < FIELD_CHUNK && c0 + b < A.nbodies; b += FIELD_THREADS) {
            double t[6];
            load_triangle(in, A.slot[in_buf], c0 + b, t);
            for (int k = 0; k < 6; ++k) s_tri[b][k] = t[k];
            s_box[b][0] = fmin(t[0], fmin(t[2], t[4])) - band;
            s_box[b][1] = fmax(t[0], fmax(t[2], t[4])) + band;
            s_box[b][2] = fmin(t[1], fmin(t[3], t[5])) - band;
            s_box[b][3] = fmax(t[1], fmax(t[3], t[5])) + band;
        }
        __syncthreads();
        const uint32_t nchunk = min((uint32_t)FIELD_CHUNK, A.nbodies - c0);
        // the block's nodes lie on storage rows [r0, r1]: one wave lists the
        // chunk's bodies whose box meets those rows, in ascending order
        // (ties keep the first body)
        if (threadIdx.x < 64u) {
            const uint32_t first = blockIdx.x * FIELD_THREADS;
            const uint32_t last = min(first + FIELD_THREADS, P.node_count) - 1u;
            const double y0 = (double)((int)(first / P.node_w) - 1) * dx, y1 = (double)((int)(last / P.node_w) - 1) * dx;
            uint32_t n = 0u;
            for (uint32_t b0 = 0; b0 < nchunk; b0 += 64u) {
                const uint32_t b = b0 + threadIdx.x;
                const bool meets = b < nchunk && !(y1 < s_box[b][2] || y0 > s_box[b][3]);
                const unsigned long long m = __ballot(meets);
                if (meets) s_near[n + (uint32_t)__popcll(m & ((1ull << threadIdx.x) - 1ull))] = b;
                n += (uint32_t)__popcll(m);
            }
            if (threadIdx.x == 0u) s_nnear = n;
        }
        __syncthreads();
        if (!live) continue;
        const uint32_t nnear = s_nnear;
        for (uint32_t k = 0; k < nnear; ++k) {
            const uint32_t b = s_near[k];
            if (px < s_box[b][0] || px > s_box[b][1] || py < s_box[b][2] || py > s_box[b][3]) continue;
            double sd, nx, ny, bary[3];
            if (!tri_sample64(px, py, s_tri[b], sd, nx, ny, bary)) {
                err |= ERR_TRIANGLE_CONTACT;
                bad_body = min(bad_body, c0 + b);
                continue;
            }
            if (sd <= band && sd < best) {
                best = sd;
                best_body = c0 + b + 1u;
                bnx = nx;
                bny = ny;
                bb[0] = bary[0];
                bb[1] = bary[1];
                bb[2] = bary[2];
            }
        }
    }
    if (live) {
        A.field_body[node] = best_body;
        FieldNode fn;
        fn.dist = best_body ? (float)best : (float)band;
        fn.nx = (float)bnx;
        fn.ny = (float)bny;
        fn.b0 = (float)bb[0];
        fn.b1 = (float)bb[1];
        fn.b2 = (float)bb[2];
        fn.ix = 0.0f;
        fn.iy = 0.0f;
        A.field[node] = fn;
    }
    if (err) raise_step_at(ctl, err, step, A.aux_pid[3u * bad_body]);
}

// trianglePairContact (reference/solver.ts:1804-1874) in FP64.
__device__ bool pair_contact64(const double* Aa, const double* Bb, double skin, double& nx, double& ny) {
    bool separated = false, have = false;
    double largest_gap = -INFINITY, smallest_overlap = INFINITY, axx = 0.0, axy = 0.0;
    for (int ti = 0; ti < 2; ++ti) {
        const double* T = ti == 0 ? Aa : Bb;
        for (int e = 0; e < 3; ++e) {
            const double ex = T[(2 * e + 2) % 6] - T[2 * e], ey = T[(2 * e + 3) % 6] - T[2 * e + 1];
            const double el = sqrt(ex * ex + ey * ey);
            if (el <= 1e-10) return false;
            const double ux = ey / el, uy = -ex / el;
            double mina = INFINITY, maxa = -INFINITY, minb = INFINITY, maxb = -INFINITY;
            for (int v = 0; v < 3; ++v) {
                const double a = Aa[2 * v] * ux + Aa[2 * v + 1] * uy, bv = Bb[2 * v] * ux + Bb[2 * v + 1] * uy;
                mina = fmin(mina, a);
                maxa = fmax(maxa, a);
                minb = fmin(minb, bv);
                maxb = fmax(maxb, bv);
            }
            const double gap = fmax(minb - maxa, mina - maxb);
            if (gap > 0) {
                separated = true;
                if (gap > largest_gap) {
                    largest_gap = gap;
                    axx = ux;
                    axy = uy;
                    have = true;
                }
            } else if (!separated && -gap < smallest_overlap) {
                smallest_overlap = -gap;
                axx = ux;
                axy = uy;
                have = true;
            }
        }
    }
    if (separated && largest_gap > skin) return false;
    if (!have) return false;
    const double dxc = (Aa[0] + Aa[2] + Aa[4]) / 3 - (Bb[0] + Bb[2] + Bb[4]) / 3;
    const double dyc = (Aa[1] + Aa[3] + Aa[5]) / 3 - (Bb[1] + Bb[3] + Bb[5]) / 3;
    const double sign = dxc * axx + dyc * axy >= 0 ? 1.0 : -1.0;
    nx = sign * axx;
    ny = sign * axy;
    const double pen = separated ? skin - largest_gap : skin + smallest_overlap;
    return pen > 0;
}

__device__ __forceinline__ void copy_slot(const SlotBuf& b, size_t from, size_t to) {
    b.pos[to] = b.pos[from];
    b.vel[to] = b.vel[from];
    b.C[to] = b.C[from];
    b.F[to] = b.F[from];
    b.pid[to] = b.pid[from];
    b.mat[to] = b.mat[from];
    b.S[to] = b.S[from];
    b.plastic[to] = b.plastic[from];
}

constexpr int BODY_THREADS = 1024;  // 16 waves: the body passes are chains of dependent loads
constexpr int FRACTURE_THREADS = 256;  // the endpoint stress cache needs more than 128 registers
constexpr uint32_t MAX_BODIES = 1024;  // bodies per scene (checked on the host)
constexpr uint32_t PAIR_CACHE_BODIES = 64;  // body counts whose pair tests are cached in LDS
constexpr uint32_t PAIR_CACHE_PAIRS = PAIR_CACHE_BODIES * (PAIR_CACHE_BODIES - 1) / 2;

// After G2P: applyTriangleContactImpulses, enforceBondedTriangleConstraints,
// resolveBondedTriangleContacts (reference/solver.ts:1592-1802;
// shaders.ts:2535-2955), the CFL check of the triangle vertices deferred from
// G2P (validateCflParticles, shaders.ts:2423-2454), and moving a vertex whose
// new position left its bucket.  One block.
__global__ __launch_bounds__(BODY_THREADS) void k_bodies(const SceneDesc* __restrict__ scenes, uint32_t step) {
    const SceneDesc& S = scenes[0];
    Control* ctl = S.ctl;
    if (block_status_set(ctl)) return;
    const Params& P = S.P;
    const AuxData& A = S.aux;
    const int in_buf = (int)((S.start_buf + step) & 1u), out_buf = in_buf ^ 1;
    const SlotBuf in = S.buf[in_buf];
    const SlotBuf out = S.buf[out_buf];
    const uint32_t* in_slot = A.slot[in_buf];
    uint32_t* out_slot = A.slot[out_buf];
    const uint32_t nb = A.nbodies, tid = threadIdx.x;
    __shared__ double s_tri[MAX_BODIES][6];
    __shared__ float s_vel[MAX_BODIES][2];
    __shared__ uint32_t s_cnt[MAX_BODIES];
    __shared__ uint32_t s_oid[MAX_BODIES];
    // up to PAIR_CACHE_BODIES bodies, each unordered pair (i < j) is tested
    // once: contact flag and the normal from i to j, as the float the
    // impulse uses
    __shared__ uint32_t s_pair_hit[PAIR_CACHE_PAIRS];
    __shared__ float s_pair_n[PAIR_CACHE_PAIRS][2];
    __shared__ uint32_t s_islot[3 * MAX_BODIES];  // vertex slots in the input and output buffers
    __shared__ uint32_t s_oslot[3 * MAX_BODIES];  // (the output ones change only in the final re-slot)
    __shared__ uint32_t s_moved[(3 * MAX_BODIES + 31) / 32];  // bit a: vertex a left its bucket
    __shared__ uint32_t s_part_n[BODY_THREADS];  // per-thread partial contact counts
    __shared__ float s_part_i[BODY_THREADS][2];  // per-thread partial contact impulses
    uint32_t err = 0u, err_index = NO_AUX;
    const double dx = P.dx, band = (double)P.dx * TRIANGLE_FIELD_BAND_CELLS;
    const float inv_mass = 1.0f / P.particle_mass;
    const uint32_t lane = tid & 63u, wave = tid >> 6;
    constexpr uint32_t BODY_WAVES = BODY_THREADS / 64;
    for (uint32_t a = tid; a < 3u * nb; a += BODY_THREADS) {
        s_islot[a] = in_slot[a];
        s_oslot[a] = out_slot[a];
    }
    __syncthreads();

    // field impulses on each body, barycentric-weighted: two bodies a wave,
    // 32 lanes each, every lane summing every 32nd node of its body's box,
    // then the 32 lanes in a fixed tree, so the sum's order is fixed
    const uint32_t half = lane >> 5, hl = lane & 31u;
    for (uint32_t b0 = 2u * wave; b0 < nb; b0 += 2u * BODY_WAVES) {
        const uint32_t b = b0 + half;
        const bool live = b < nb;
        float acc[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        // lane v < 3 of a half updates vertex v: its slot and velocity load during the walk
        uint32_t sl = 0u;
        float2 vel = make_float2(0.f, 0.f);
        if (live) {
            double t[6];
            load_triangle(in, s_islot, b, t);
            const double mnx = fmin(t[0], fmin(t[2], t[4])), mxx = fmax(t[0], fmax(t[2], t[4]));
            const double mny = fmin(t[1], fmin(t[3], t[5])), mxy = fmax(t[1], fmax(t[3], t[5]));
            const int x0 = (int)fmax(-1.0, floor((mnx - band) / dx) - 1.0);
            const int x1 = (int)fmin((double)P.cells_x + 1.0, ceil((mxx + band) / dx) + 1.0);
            const int y0 = (int)fmax(-1.0, floor((mny - band) / dx) - 1.0);
            const int y1 = (int)fmin((double)P.cells_y + 1.0, ceil((mxy + band) / dx) + 1.0);
            if (hl < 3u) {
                sl = s_oslot[3u * b + hl];
                vel = out.vel[sl];
            }
            if (x1 >= x0 && y1 >= y0) {
                const uint32_t w = (uint32_t)(x1 - x0 + 1), count = w * (uint32_t)(y1 - y0 + 1);
                // four nodes a lane in flight, both loads of a node issued
                // together; the sums still run in node order
                constexpr uint32_t U = 4u;
                for (uint32_t i0 = hl; i0 < count; i0 += 32u * U) {
                    uint32_t fb[U];
                    FieldNode fn[U];
#pragma unroll
                    for (uint32_t u = 0; u < U; ++u) {
                        const uint32_t i = i0 + 32u * u;
                        fb[u] = 0u;  // no body: the node adds nothing
                        fn[u] = FieldNode{};
                        if (i < count) {
                            const uint32_t node =
                                (uint32_t)(y0 + (int)(i / w) + 1) * P.node_w + (uint32_t)(x0 + (int)(i % w) + 1);
                            fb[u] = A.field_body[node];
                            fn[u] = A.field[node];
                        }
                    }
#pragma unroll
                    for (uint32_t u = 0; u < U; ++u) {
                        if (fb[u] == b + 1u) {
                            acc[0] += fn[u].b0 * fn[u].ix;
                            acc[1] += fn[u].b0 * fn[u].iy;
                            acc[2] += fn[u].b1 * fn[u].ix;
                            acc[3] += fn[u].b1 * fn[u].iy;
                            acc[4] += fn[u].b2 * fn[u].ix;
                            acc[5] += fn[u].b2 * fn[u].iy;
                        }
                    }
                }
            }
        }
        // each half's 32 lanes, in a fixed tree (every lane of the wave takes part)
#pragma unroll
        for (int off = 16; off > 0; off >>= 1)
#pragma unroll
            for (int k = 0; k < 6; ++k) acc[k] += __shfl_down(acc[k], off, 32);
#pragma unroll
        for (int k = 0; k < 6; ++k) acc[k] = __shfl(acc[k], 0, 32);
        if (live && hl < 3u) {
            const float ix = hl == 0u ? acc[0] : (hl == 1u ? acc[2] : acc[4]);
            const float iy = hl == 0u ? acc[1] : (hl == 1u ? acc[3] : acc[5]);
            vel.x = vel.x + ix * inv_mass;
            vel.y = vel.y + iy * inv_mass;
            if (!finite2(vel)) {
                err |= ERR_TRIANGLE_CONTACT;
                err_index = min(err_index, A.aux_pid[3u * b]);
            }
            out.vel[sl] = vel;
        }
    }
    __syncthreads();

    // two Gauss-Seidel passes over the edges of each triangle, in FP64 offsets
    // from the first vertex (exact differences of the float32 positions), so
    // the velocity change displacement / dt keeps no rounding of the large
    // world coordinates; then confinement
    const double ratio = (double)P.dt / (double)P.bond_ref_dt;
    double k = fmin((double)P.bond_stiffness * ratio * ratio, 1.0);
    k = isfinite(k) ? fmax(0.0, fmin(1.0, k)) : 0.0;
    const double rest = P.bond_rest;
    for (uint32_t b = tid; b < nb; b += BODY_THREADS) {
        uint32_t sl[3];
        float2 orig[3];
        double ox[3], oy[3], px[3], py[3];
        for (int v = 0; v < 3; ++v) {
            sl[v] = s_oslot[3u * b + (uint32_t)v];
            orig[v] = out.pos[sl[v]];
        }
        for (int v = 0; v < 3; ++v) {
            ox[v] = px[v] = (double)orig[v].x - (double)orig[0].x;
            oy[v] = py[v] = (double)orig[v].y - (double)orig[0].y;
        }
        bool ok = true;
        for (int it = 0; it < BONDED_TRIANGLE_SOLVER_ITERATIONS && ok; ++it)
            for (int e = 0; e < 3; ++e) {
                const int a = e, c = (e + 1) % 3;
                const double ddx = px[c] - px[a], ddy = py[c] - py[a];
                const double d = sqrt(ddx * ddx + ddy * ddy);
                if (!isfinite(d) || d <= 1e-10) {
                    ok = false;
                    break;
                }
                const double sc = ((d - rest) * k * 0.5) / d;
                px[a] += ddx * sc;
                py[a] += ddy * sc;
                px[c] -= ddx * sc;
                py[c] -= ddy * sc;
            }
        if (!ok) {
            err |= ERR_TRIANGLE_BOND;
            err_index = min(err_index, A.aux_pid[3u * b]);
            continue;
        }
        for (int v = 0; v < 3; ++v) {
            const double dxv = px[v] - ox[v], dyv = py[v] - oy[v];
            const float2 vel = out.vel[sl[v]];
            const float2 bx = confine_axis((float)((double)orig[v].x + dxv), (float)((double)vel.x + dxv / P.dt), P.dom_w);
            const float2 by = confine_axis((float)((double)orig[v].y + dyv), (float)((double)vel.y + dyv / P.dt), P.dom_h);
            out.pos[sl[v]] = make_float2(bx.x, by.x);
            out.vel[sl[v]] = make_float2(bx.y, by.y);
        }
    }
    __syncthreads();

    // pair contacts between bodies of different objects: triangles of the
    // input state, mean velocities after the edge solve
    if (nb >= 2u) {
        for (uint32_t b = tid; b < nb; b += BODY_THREADS) {
            double t[6];
            load_triangle(in, s_islot, b, t);
            for (int q = 0; q < 6; ++q) s_tri[b][q] = t[q];
            s_oid[b] = A.body_oid[b];
            float vx = 0.0f, vy = 0.0f;
            for (int v = 0; v < 3; ++v) {
                const float2 vel = out.vel[s_oslot[3u * b + (uint32_t)v]];
                vx += vel.x;
                vy += vel.y;
            }
            s_vel[b][0] = vx / 3.0f;
            s_vel[b][1] = vy / 3.0f;
        }
        __syncthreads();
        const double skin = (double)P.dx * TRIANGLE_CONTACT_SKIN_CELLS;
        const bool cached = nb <= PAIR_CACHE_BODIES;
        if (cached) {
            const uint32_t npairs = nb * (nb - 1u) / 2u;
            for (uint32_t pi = tid; pi < npairs; pi += BODY_THREADS) {
                uint32_t i = 0u, rest_p = pi;
                while (rest_p >= nb - 1u - i) {
                    rest_p -= nb - 1u - i;
                    ++i;
                }
                const uint32_t j = i + 1u + rest_p;
                double nx = 0.0, ny = 0.0;
                const bool hit = s_oid[i] != s_oid[j] && pair_contact64(s_tri[i], s_tri[j], skin, nx, ny);
                s_pair_hit[pi] = hit ? 1u : 0u;
                s_pair_n[pi][0] = (float)nx;
                s_pair_n[pi][1] = (float)ny;
            }
            __syncthreads();
        }
        // index of the unordered pair {b, j} (b != j) in the cache
        auto pair_index = [nb](uint32_t b, uint32_t j) {
            const uint32_t i = min(b, j), jj = max(b, j);
            return i * (2u * nb - i - 1u) / 2u + (jj - i - 1u);
        };
        // a team of threads per body, each testing every team-th other body;
        // the team's partial counts and impulses are then added in thread order
        const uint32_t team = max(1u, (uint32_t)BODY_THREADS / nb);
        const uint32_t teams = (uint32_t)BODY_THREADS / team;
        const uint32_t my_team = tid / team, t_in = tid % team;
        for (uint32_t b0 = 0; b0 < nb; b0 += teams) {
            const uint32_t b = b0 + my_team;
            uint32_t c = 0u;
            if (my_team < teams && b < nb) {
                for (uint32_t j = t_in; j < nb; j += team) {
                    if (j == b || s_oid[j] == s_oid[b]) continue;
                    double nx, ny;
                    if (cached ? s_pair_hit[pair_index(b, j)] != 0u
                               : (b < j ? pair_contact64(s_tri[b], s_tri[j], skin, nx, ny)
                                        : pair_contact64(s_tri[j], s_tri[b], skin, nx, ny)))
                        ++c;
                }
            }
            s_part_n[tid] = c;
            __syncthreads();
            if (my_team < teams && b < nb && t_in == 0u) {
                uint32_t sum = 0u;
                for (uint32_t q = 0; q < team; ++q) sum += s_part_n[tid + q];
                s_cnt[b] = sum;
            }
            __syncthreads();
        }
        const float body_mass = 3.0f * P.particle_mass;
        for (uint32_t b0 = 0; b0 < nb; b0 += teams) {
            const uint32_t b = b0 + my_team;
            float ix = 0.0f, iy = 0.0f;
            if (my_team < teams && b < nb && s_cnt[b] != 0u) {
                for (uint32_t j = t_in; j < nb; j += team) {
                    if (j == b || s_oid[j] == s_oid[b]) continue;
                    float fnx, fny;
                    if (cached) {
                        const uint32_t pi = pair_index(b, j);
                        if (s_pair_hit[pi] == 0u) continue;
                        fnx = s_pair_n[pi][0];
                        fny = s_pair_n[pi][1];
                    } else {
                        double nx, ny;
                        const bool hit = b < j ? pair_contact64(s_tri[b], s_tri[j], skin, nx, ny)
                                               : pair_contact64(s_tri[j], s_tri[b], skin, nx, ny);
                        if (!hit) continue;
                        fnx = (float)nx;
                        fny = (float)ny;
                    }
                    if (b > j) {  // the normal from b to j (negation is exact in either precision)
                        fnx = -fnx;
                        fny = -fny;
                    }
                    const float nx = fnx, ny = fny;
                    const float w = 1.0f / (float)max(s_cnt[b], s_cnt[j]);
                    const float rn = (s_vel[b][0] - s_vel[j][0]) * nx + (s_vel[b][1] - s_vel[j][1]) * ny;
                    if (rn < 0.0f) {
                        const float imp = -rn / (2.0f / body_mass);
                        const float iw = w * TRIANGLE_CONTACT_IMPULSE_RELAXATION;
                        ix += (iw * imp * (float)nx) / body_mass;
                        iy += (iw * imp * (float)ny) / body_mass;
                    }
                }
            }
            s_part_i[tid][0] = ix;
            s_part_i[tid][1] = iy;
            __syncthreads();
            if (my_team < teams && b < nb && t_in < 3u) {  // the team's first threads: the three vertices
                float tx = 0.0f, ty = 0.0f;
                for (uint32_t q = 0; q < team; ++q) {
                    tx += s_part_i[tid - t_in + q][0];
                    ty += s_part_i[tid - t_in + q][1];
                }
                for (uint32_t v = t_in; v < 3u; v += team) {
                    const uint32_t sl = s_oslot[3u * b + v];
                    const float2 p = out.pos[sl], vel = out.vel[sl];
                    const float2 bx = confine_axis(p.x, vel.x + tx, P.dom_w);
                    const float2 by = confine_axis(p.y, vel.y + ty, P.dom_h);
                    out.pos[sl] = make_float2(bx.x, by.x);
                    out.vel[sl] = make_float2(bx.y, by.y);
                }
            }
            __syncthreads();
        }
    }

    // CFL check of the vertices (their velocities changed after G2P)
    if (P.cfl_on >= 0.5f) {
        for (uint32_t a = tid; a < 3u * nb; a += BODY_THREADS) {
            const uint32_t sl = s_oslot[a];
            const float2 vel = out.vel[sl];
            const Material mt = S.mats[out.mat[sl]];
            const float speed = sqrtf(vel.x * vel.x + vel.y * vel.y);
            const float ws = wave_speed_bound<true>(P, mt.model, out.F[sl], mt.lambda, mt.mu, mt.density, mt.volume,
                                                    mt.param);
            if (!finite2(vel) || !finitef(speed) || !finitef(ws) || speed > P.vel_bound || ws > P.wave_bound) {
                record_failure(ctl, 3u, A.aux_pid[a], speed, ws, P.vel_bound, P.wave_bound);
                raise_step(ctl, ERR_CFL_BOUND, step);
            }
        }
    }
    if (err) raise_step_at(ctl, err, step, err_index);
    __syncthreads();

    // a vertex moved by the contacts keeps the bucket invariant: re-slot it.
    // Every thread tests its vertices; one thread then moves the few that
    // left their bucket, in vertex order.
    const uint32_t cap = S.cap;
    const uint32_t words = (3u * nb + 31u) / 32u;
    for (uint32_t q = tid; q < words; q += BODY_THREADS) s_moved[q] = 0u;
    __syncthreads();
    for (uint32_t a = tid; a < 3u * nb; a += BODY_THREADS) {
        const uint32_t src = s_oslot[a];
        Stencil st;
        const bool fits = make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, out.pos[src], st);
        if (!fits || bucket_of(st, S.bw) != src / cap) atomicOr(&s_moved[a >> 5], 1u << (a & 31u));
    }
    __syncthreads();
    if (tid == 0) {
        // the set bits in ascending vertex order; a failure ends the walk
        bool stop = false;
        for (uint32_t q = 0; q < words && !stop; ++q) {
            for (uint32_t bits = s_moved[q]; bits && !stop; bits &= bits - 1u) {
                const uint32_t a = 32u * q + (uint32_t)__builtin_ctz(bits);
                const uint32_t src = out_slot[a];
                Stencil st;
                if (!make_stencil(P.dx, P.inv_dx, P.cells_x, P.cells_y, out.pos[src], st)) {
                    raise_step(ctl, ERR_STENCIL, step);
                    stop = true;
                    continue;
                }
                const uint32_t nbk = bucket_of(st, S.bw), obk = src / cap;
                if (nbk == obk) continue;
                const uint32_t kk = out.cnt[nbk];
                if (kk >= cap) {
                    raise_step(ctl, ERR_INTERNAL_REBUCKET, step);
                    stop = true;
                    continue;
                }
                const uint32_t dst = nbk * cap + kk;
                out.cnt[nbk] = kk + 1u;
                copy_slot(out, src, dst);
                const uint32_t last = out.cnt[obk] - 1u;
                out.cnt[obk] = last;
                const uint32_t L = obk * cap + last;
                if (L != src) {
                    copy_slot(out, L, src);
                    const uint32_t moved = A.aux_of[out.pid[src]];
                    if (moved != NO_AUX) out_slot[moved] = src;
                }
                out_slot[a] = dst;
            }
        }
    }
}

// updateFractureBonds + applyFractureParticleDamage (shaders.ts:2957-3039;
// reference/solid-bonds.ts:68-108): bond damage from the tensile strain of
// the output positions, then each endpoint's damage = max(its damage, its
// bonds'), and the endpoint's stress cache with that damage.  One block.
__global__ __launch_bounds__(FRACTURE_THREADS) void k_fracture(const SceneDesc* __restrict__ scenes, uint32_t step) {
    const SceneDesc& S = scenes[0];
    Control* ctl = S.ctl;
    if (block_status_set(ctl)) return;
    const AuxData& A = S.aux;
    const int in_buf = (int)((S.start_buf + step) & 1u), out_buf = in_buf ^ 1;
    const SlotBuf out = S.buf[out_buf];
    const uint32_t* out_slot = A.slot[out_buf];
    const uint32_t tid = threadIdx.x;
    uint32_t err = 0u, err_index = NO_AUX;
    for (uint32_t k = tid; k < A.nbonds; k += FRACTURE_THREADS) {
        Bond bd = A.bonds[k];
        if (bd.broken != 0u) {
            bd.damage = 1.0f;
        } else {
            const float2 p1 = out.pos[out_slot[A.aux_of[bd.first]]], p2 = out.pos[out_slot[A.aux_of[bd.second]]];
            const float sx = p2.x - p1.x, sy = p2.y - p1.y;
            const float d = sqrtf(sx * sx + sy * sy);
            if (!finitef(d)) {
                err |= ERR_FRACTURE_BOND;
                err_index = min(err_index, bd.first);
                continue;
            }
            const float strain = fmaxf(0.0f, (d - bd.rest) / bd.rest);
            bd.damage = fmaxf(bd.damage, fminf(1.0f, strain / bd.failure));
            if (strain >= bd.failure) {
                bd.damage = 1.0f;
                bd.broken = 1u;
            }
        }
        A.bonds[k] = bd;
    }
    __syncthreads();
    for (uint32_t a = A.frac_aux0 + tid; a < A.naux; a += FRACTURE_THREADS) {
        float dmg = fminf(fmaxf(A.damage[in_buf][a], 0.0f), 1.0f);
        for (uint32_t q = A.bond_off[a]; q < A.bond_off[a + 1]; ++q) dmg = fmaxf(dmg, A.bonds[A.bond_inc[q]].damage);
        A.damage[out_buf][a] = dmg;
        const uint32_t sl = out_slot[a];
        Material mt = S.mats[out.mat[sl]];
        mt.param = dmg;
        const float4 F = out.F[sl];
        float4 Sc;
        const uint32_t pf = particle_stress<true>(mt, F, out.plastic[sl], det_comp(F), Sc);
        out.S[sl] = Sc;
        if (pf) {
            if (pf & ERR_P2G) atomicMin(&ctl->deferred_index[out_buf], A.aux_pid[a]);
            atomicOr(&ctl->deferred[out_buf], pf);
        }
    }
    if (err) raise_step_at(ctl, err, step, err_index);
}

// ------------------------------------------------------------- batch end
__global__ void k_commit_last(const SceneDesc* __restrict__ scenes, uint32_t steps) {
    Control* ctl = scenes[blockIdx.x].ctl;  // one block per scene
    const uint32_t seq = scenes[blockIdx.x].seq;
    if (threadIdx.x == 0) {
        const uint32_t done = min(ctl->fail_step, steps);
        ctl->completed = done;
        ctl->history[seq % CONTROL_HISTORY] = done;
    }
}

// On a failed microstep the committed state sits in the buffer of the last
// committed microstep; copy it to the batch's output buffer when their
// parity differs (commitParticles' whole-step rollback, shaders.ts:3067-3080).
__global__ __launch_bounds__(256) void k_fix_rollback(const SceneDesc* __restrict__ scenes, uint32_t steps) {
    const SceneDesc& S = scenes[blockIdx.y];  // one row of blocks per scene
    const uint32_t start_buf = S.start_buf;
    Control* ctl = S.ctl;
    const SlotBuf& b0 = S.buf[0];
    const SlotBuf& b1 = S.buf[1];
    const uint32_t nb = S.nb, cap = S.cap;
    const bool general = scene_general(S.scene_flags);
    if (ctl->status[0] == 0u) return;
    const uint32_t committed = (start_buf + ctl->completed) & 1u;
    const uint32_t final_buf = (start_buf + steps) & 1u;
    if (committed == final_buf) return;
    const SlotBuf& src = committed ? b1 : b0;
    const SlotBuf& dst = committed ? b0 : b1;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        ctl->deferred[final_buf] = ctl->deferred[committed];
        ctl->deferred_index[final_buf] = ctl->deferred_index[committed];
    }
    if (blockIdx.x == 0) {  // the aux particles keep their slots and damage
        const AuxData& A = S.aux;
        for (uint32_t a = threadIdx.x; a < A.naux; a += blockDim.x) {
            A.slot[final_buf][a] = A.slot[committed][a];
            A.damage[final_buf][a] = A.damage[committed][a];
        }
    }
    const size_t total = (size_t)nb * cap;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const uint32_t b = (uint32_t)g / cap, k = (uint32_t)g - b * cap;  // slot counts fit 32 bits
        const uint32_t c = src.cnt[b];
        if (k == 0) dst.cnt[b] = c;
        if (k >= c) continue;
        dst.pos[g] = src.pos[g];
        dst.vel[g] = src.vel[g];
        dst.C[g] = src.C[g];
        dst.F[g] = src.F[g];
        dst.pid[g] = src.pid[g];
        dst.mat[g] = src.mat[g];
        if (general) {
            dst.S[g] = src.S[g];
            dst.plastic[g] = src.plastic[g];
        }
    }
}

// measureParticleSpeedPartials + reduceParticleSpeedPartials
// (shaders.ts:2463-2533): wave max, then one atomic per block on the
// non-negative float's bit pattern.
__global__ __launch_bounds__(256) void k_measure_speed(const SceneDesc* __restrict__ scenes, uint32_t steps) {
    const SceneDesc& S = scenes[blockIdx.y];  // one row of blocks per scene
    const uint32_t which = (S.start_buf + steps) & 1u;  // the batch's final buffer
    Control* ctl = S.ctl;
    const SlotBuf& buf = S.buf[which];
    const uint32_t nb = S.nb, cap = S.cap;
    __shared__ float s_w[4];
    float m = 0.0f;
    // a thread a bucket: its count, then its used slots four at a time
    // (their loads issued together), so no thread waits on a chain of them
    for (uint32_t b = blockIdx.x * blockDim.x + threadIdx.x; b < nb; b += gridDim.x * blockDim.x) {
        const uint32_t c = min(buf.cnt[b], cap);
        const size_t base = (size_t)b * cap;
        for (uint32_t k0 = 0; k0 < c; k0 += 4u) {
            float2 v[4];
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) v[u] = k0 + u < c ? buf.vel[base + k0 + u] : make_float2(0.f, 0.f);
#pragma unroll
            for (uint32_t u = 0; u < 4u; ++u) {
                const float sp = sqrtf(v[u].x * v[u].x + v[u].y * v[u].y);
                m = fmaxf(m, finitef(sp) ? sp : FLT_MAX);
            }
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        const float b = fmaxf(fmaxf(s_w[0], s_w[1]), fmaxf(s_w[2], s_w[3]));
        atomicMax(&ctl->max_speed_bits, __float_as_uint(b));
    }
}

// Particle mass and momentum (totalMass / totalMomentum,
// reference/solver.ts:2505-2526) in FP64 for the scenes with diagnostics on,
// of the state `after` microsteps into the batch (0 or its length):
// block b of a scene sums its buckets [b*nb/B, (b+1)*nb/B) in slot order
// into diag_part[3b..3b+2].
__global__ __launch_bounds__(256) void k_particle_sums(const SceneDesc* __restrict__ scenes, uint32_t after) {
    const SceneDesc& S = scenes[blockIdx.y];  // one row of blocks per scene
    if (!S.diag) return;
    const uint32_t which = (S.start_buf + after) & 1u;  // the buffer `after` microsteps into the batch
    const SlotBuf& buf = S.buf[which];
    const uint32_t nb = S.nb, cap = S.cap;
    const uint32_t b0 = (uint32_t)((uint64_t)blockIdx.x * nb / gridDim.x);
    const uint32_t b1 = (uint32_t)((uint64_t)(blockIdx.x + 1) * nb / gridDim.x);
    double m = 0.0, px = 0.0, py = 0.0;
    for (uint32_t b = b0; b < b1; ++b) {
        const uint32_t c = min(buf.cnt[b], cap);
        for (uint32_t k = threadIdx.x; k < c; k += blockDim.x) {
            const size_t g = (size_t)b * cap + k;
            const double mass = (double)S.mats[buf.mat[g]].mass;
            const float2 v = buf.vel[g];
            m += mass;
            px += mass * (double)v.x;
            py += mass * (double)v.y;
        }
    }
    block_sum3_f64(m, px, py, S.diag_part + 3 * blockIdx.x);
}

// One block per scene with diagnostics on: the partial triples in a fixed
// order into diag[slot], diag[slot + 3], diag[slot + 6] (mass, momentum x,
// momentum y).  grid != 0 sums the per-tile grid partials of the last P2G,
// otherwise the `parts` particle partials.
__global__ __launch_bounds__(256) void k_finish_sums(const SceneDesc* __restrict__ scenes, uint32_t grid,
                                                     uint32_t parts, uint32_t slot) {
    const SceneDesc& S = scenes[blockIdx.x];
    if (!S.diag) return;
    const double* part = grid ? S.diag_grid : S.diag_part;
    const uint32_t n = grid ? S.ptx * S.pty : parts;
    double a = 0.0, b = 0.0, c = 0.0;
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
        a += part[3 * i];
        b += part[3 * i + 1];
        c += part[3 * i + 2];
    }
    __shared__ double s_out[3];
    block_sum3_f64(a, b, c, s_out);
    __syncthreads();
    if (threadIdx.x == 0) {
        S.diag[slot] = s_out[0];
        S.diag[slot + 3] = s_out[1];
        S.diag[slot + 6] = s_out[2];
    }
}

inline uint32_t blocks_for(size_t n, uint32_t t) { return (uint32_t)((n + t - 1) / t); }

// Kernel and block size of the step kernels for a launch shape.  Scenes
// with bodies or bonds (AUX) run the default shape only.
// The counters of both microstep sets, at the start of a step launch.
__global__ __launch_bounds__(64) void k_queue_reset(TileQueue q) {
    for (uint32_t i = threadIdx.x; i < 2u * QUEUE_SET_WORDS; i += blockDim.x) q.sets[i] = 0u;
}

// Active-tile lists of one microstep, for step kernels that take their tiles
// from them: a P2G tile is active when a bucket of its region holds a
// particle, a G2P tile when one of its own buckets does (the tiles the step
// kernels would otherwise end at once).  An inactive P2G tile's duties are
// done here: clearing the next counts of the buckets it owns, the scene's
// deferred stress-cache flags when it is the scene's first tile, and its
// diagnostic grid sums.  Scenes whose status is set are passed over, as the
// step kernels pass over them.  The next microstep's counters are cleared.
// A list's order depends on the order of the atomic adds; the results do
// not, since every tile computes the same sums whichever block runs it.
template <int TI, int TB>
__global__ __launch_bounds__(256) void k_compact_tiles(const SceneDesc* __restrict__ scenes,
                                                       const uint32_t* __restrict__ p2g_scene,
                                                       const uint32_t* __restrict__ g2p_scene, uint32_t nscenes,
                                                       uint32_t step, TileQueue q) {
    using T = P2GTile<TI>;
    __shared__ uint32_t s_n[2 * QUEUE_XCDS];     // this block's tiles per list, then their offset in it
    const uint32_t g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < QUEUE_SET_WORDS) q.sets[((step + 1u) & 1u) * QUEUE_SET_WORDS + g] = 0u;
    if (threadIdx.x < 2u * QUEUE_XCDS) s_n[threadIdx.x] = 0u;
    __syncthreads();
    uint32_t kind = 0u, t = 0u, slot = 0xffffffffu;  // slot: list entry of an active tile
    if (g < q.tiles[0] + q.tiles[1]) {
        kind = g < q.tiles[0] ? 0u : 1u;
        t = kind ? g - q.tiles[0] : g;
        const SceneDesc& S = scenes[nscenes == 1u ? 0u : (kind ? g2p_scene : p2g_scene)[t]];
        Control* ctl = S.ctl;
        if (load_status(ctl) == 0u) {
            const uint32_t tile = t - (kind ? S.g2p_base : S.p2g_base);
            const int in_buf = (int)((S.start_buf + step) & 1u);
            const uint32_t* cnt = S.buf[in_buf].cnt;
            const int bw = (int)S.bw, bh = (int)S.bh;
            uint32_t any = 0u;  // every count is loaded, then they are tested together
            if (kind == 0u) {
                const int tx = (int)(tile % S.ptx), ty = (int)(tile / S.ptx);
                const int rbx0 = tx * T::OWN - 1, rby0 = ty * T::OWN - 1;
#pragma unroll
                for (int b = 0; b < T::NB; ++b) {
                    const int gbx = rbx0 + b % T::RB, gby = rby0 + b / T::RB;
                    if (gbx >= 0 && gby >= 0 && gbx < bw && gby < bh) any |= cnt[gby * bw + gbx];
                }
                if (any == 0u) {
                    uint32_t* cnt_next = S.buf[in_buf ^ 1].cnt;
                    for (int k = 0; k < T::OWN * T::OWN; ++k) {
                        const int obx = T::OWN * tx + k % T::OWN, oby = T::OWN * ty + k / T::OWN;
                        if (obx < bw && oby < bh) cnt_next[oby * bw + obx] = 0u;
                    }
                    if (tile == 0u) scene_deferred_flags(ctl, in_buf, step);
                    if (S.P.diag_on) {
                        double* dg = S.diag_grid;
                        dg[3 * tile] = 0.0;
                        dg[3 * tile + 1] = 0.0;
                        dg[3 * tile + 2] = 0.0;
                    }
                }
            } else {
                const int tbx = (int)(tile % S.gtx) * TB, tby = (int)(tile / S.gtx) * TB;
#pragma unroll
                for (int b = 0; b < TB * TB; ++b) {
                    const int gbx = tbx + b % TB, gby = tby + b / TB;
                    if (gbx < bw && gby < bh) any |= cnt[gby * bw + gbx];
                }
            }
            if (any != 0u) slot = atomicAdd(&s_n[kind * QUEUE_XCDS + xcd_of_tile(q.tiles[kind], t)], 1u);
        }
    }
    __syncthreads();
    // one add per list and block to the microstep's counters
    uint32_t* set = q.sets + (step & 1u) * QUEUE_SET_WORDS;
    if (threadIdx.x < 2u * QUEUE_XCDS) {
        const uint32_t c = s_n[threadIdx.x];
        s_n[threadIdx.x] = c != 0u ? atomicAdd(&set[threadIdx.x], c) : 0u;
    }
    __syncthreads();
    if (slot != 0xffffffffu) {
        const uint32_t x = xcd_of_tile(q.tiles[kind], t);
        q.list[kind][xcd_range_start(q.tiles[kind], x) + s_n[kind * QUEUE_XCDS + x] + slot] = t;
    }
}

using CompactKernel = void (*)(const SceneDesc*, const uint32_t*, const uint32_t*, uint32_t, uint32_t, TileQueue);
CompactKernel pick_compact(const Shape& shape) {
    const bool small = shape.g2p_tile == 4u;
    switch (shape.p2g_tile) {
        case 20u: return small ? k_compact_tiles<20, 4> : k_compact_tiles<20, 8>;
        case 12u: return small ? k_compact_tiles<12, 4> : k_compact_tiles<12, 8>;
        default: return small ? k_compact_tiles<28, 4> : k_compact_tiles<28, 8>;
    }
}

// Blocks of a queued step launch: K for each of the 8 XCDs, K the device's
// compute units (eight blocks a compute unit, more than any step kernel holds
// at once) or, for a small launch, the largest XCD range of tiles, so that
// each block then takes at most one tile.
uint32_t queued_blocks(uint32_t tiles) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
        return 0u;
    const uint32_t range = (tiles + QUEUE_XCDS - 1u) / QUEUE_XCDS;
    return QUEUE_XCDS * std::max<uint32_t>(1u, std::min<uint32_t>((uint32_t)cus, range));
}

using StepKernel = void (*)(const SceneDesc*, const uint32_t*, uint32_t, uint32_t, TileQueue);
struct StepChoice {
    StepKernel fn;
    uint32_t threads;
};

template <int TI>
StepChoice p2g_choice(bool fixed_point, bool general, bool solids) {
    StepKernel fn;
    if (fixed_point)
        fn = general ? (solids ? k_p2g_fixed<TI, true, true> : k_p2g_fixed<TI, true, false>)
                     : (solids ? k_p2g_fixed<TI, false, true> : k_p2g_fixed<TI, false, false>);
    else
        fn = general ? k_p2g_gather<TI, true> : k_p2g_gather<TI, false>;
    return {fn, (uint32_t)P2GTile<TI>::THREADS};
}

StepChoice pick_p2g(const StepBatch& sb, bool general, bool aux) {
    const bool solids = (sb.scene_flags & SCENE_SOLID) != 0u;
    if (aux)
        return {sb.fixed_point ? k_p2g_fixed<28, true, true, true> : k_p2g_gather<28, true, true>,
                (uint32_t)P2GTile<28>::THREADS};
    switch (sb.shape.p2g_tile) {
        case 20u: return p2g_choice<20>(sb.fixed_point != 0u, general, solids);
        case 12u: return p2g_choice<12>(sb.fixed_point != 0u, general, solids);
        default: return p2g_choice<28>(sb.fixed_point != 0u, general, solids);
    }
}

// G2P variant: plain liquids; general scenes without solids (Tait liquids);
// general scenes with solids
template <int TB, int THREADS>
StepChoice g2p_choice(bool general, bool solids) {
    StepKernel fn = !general ? k_g2p<TB, THREADS, false>
                             : (solids ? k_g2p<TB, THREADS, true> : k_g2p<TB, THREADS, true, false, false>);
    return {fn, (uint32_t)THREADS};
}

StepChoice pick_g2p(const StepBatch& sb, bool general, bool aux) {
    if (aux) return sb.shape.g2p_tile == 4u ? StepChoice{k_g2p<4, 256, true, true>, 256u}
                                            : StepChoice{k_g2p<8, 256, true, true>, 256u};
    const bool solids = (sb.scene_flags & SCENE_SOLID) != 0u;
    if (sb.shape.g2p_tile == 4u) return sb.shape.g2p_threads == 128u ? g2p_choice<4, 128>(general, solids)
                                                                     : g2p_choice<4, 256>(general, solids);
    switch (sb.shape.g2p_threads) {
        case 128u: return g2p_choice<8, 128>(general, solids);
        case 512u: return g2p_choice<8, 512>(general, solids);
        default: return g2p_choice<8, 256>(general, solids);
    }
}

}  // namespace

#define MPM_LAUNCH(name, ...)                                                     \
    do {                                                                          \
        TimedRegion _tr(lc.timing, name, lc.stream);                              \
        hipLaunchKernelGGL(__VA_ARGS__);                                          \
        hipError_t _e = hipGetLastError();                                        \
        if (_e != hipSuccess) return _e;                                          \
    } while (0)

hipError_t launch_steps(const StepBatch& sb, uint32_t steps, bool measure) {
    const bool general = scene_general(sb.scene_flags);
    const uint32_t ns = sb.nscenes;
    struct {
        hipStream_t stream;
        KernelTimes* timing;
    } lc = {sb.stream, sb.timing};
    // particle partial blocks per scene: one scene uses the whole GPU
    const uint32_t parts = ns == 1 ? DIAG_PART_BLOCKS : 32u;
    MPM_LAUNCH("batch_begin", k_batch_begin, dim3(ns), dim3(64), 0, lc.stream, sb.d_scenes);
    if (sb.diag) {
        MPM_LAUNCH("particle_sums", k_particle_sums, dim3(parts, ns), dim3(256), 0, lc.stream, sb.d_scenes, 0u);
        MPM_LAUNCH("finish_sums", k_finish_sums, dim3(ns), dim3(256), 0, lc.stream, sb.d_scenes, 0u, parts, 0u);
    }
    const bool aux = (sb.scene_flags & SCENE_AUX) != 0u;
    if (!aux && !shape_supported(sb.shape)) return hipErrorInvalidValue;
    const StepChoice p2g = pick_p2g(sb, general, aux);
    const StepChoice g2p = pick_g2p(sb, general, aux);
    const char* p2g_name = sb.fixed_point ? "p2g_grid_fixed" : "p2g_grid";
    // with a tile queue the step kernels run as the blocks the device holds
    // at once and take the active tiles from the queue
    const TileQueue q = aux ? TileQueue{} : sb.queue;
    uint32_t p2g_blocks = sb.p2g_tiles, g2p_blocks = sb.g2p_tiles;
    CompactKernel compact = nullptr;
    if (q.sets) {
        p2g_blocks = queued_blocks(sb.p2g_tiles);
        g2p_blocks = queued_blocks(sb.g2p_tiles);
        if (p2g_blocks == 0u || g2p_blocks == 0u) return hipErrorInvalidValue;
        compact = pick_compact(sb.shape);
        MPM_LAUNCH("queue_reset", k_queue_reset, dim3(1), dim3(64), 0, lc.stream, q);
    }
    for (uint32_t s = 0; s < steps; ++s) {
        if (aux && sb.aux_bodies)
            MPM_LAUNCH("tri_field", k_tri_field, dim3(blocks_for(sb.node_count, FIELD_THREADS)), dim3(FIELD_THREADS),
                       0, lc.stream, sb.d_scenes, s);
        if (compact)
            MPM_LAUNCH("compact_tiles", compact, dim3(blocks_for(sb.p2g_tiles + sb.g2p_tiles, 256)), dim3(256), 0,
                       lc.stream, sb.d_scenes, sb.p2g_scene, sb.g2p_scene, sb.nscenes, s, q);
        MPM_LAUNCH(p2g_name, p2g.fn, dim3(p2g_blocks), dim3(p2g.threads), 0, lc.stream, sb.d_scenes, sb.p2g_scene,
                   sb.nscenes, s, q);
        MPM_LAUNCH("g2p", g2p.fn, dim3(g2p_blocks), dim3(g2p.threads), 0, lc.stream, sb.d_scenes, sb.g2p_scene,
                   sb.nscenes, s, q);
        if (aux && sb.aux_bodies)
            MPM_LAUNCH("bodies", k_bodies, dim3(1), dim3(BODY_THREADS), 0, lc.stream, sb.d_scenes, s);
        if (aux && sb.aux_bonds)
            MPM_LAUNCH("fracture", k_fracture, dim3(1), dim3(FRACTURE_THREADS), 0, lc.stream, sb.d_scenes, s);
    }
    MPM_LAUNCH("commit_last", k_commit_last, dim3(ns), dim3(64), 0, lc.stream, sb.d_scenes, steps);
    const uint32_t fb = std::max<uint32_t>(1u, std::min<uint32_t>(blocks_for(sb.max_slots, 256), ns == 1 ? 2048u : 64u));
    MPM_LAUNCH("fix_rollback", k_fix_rollback, dim3(fb, ns), dim3(256), 0, lc.stream, sb.d_scenes, steps);
    if (sb.diag) {
        MPM_LAUNCH("finish_sums", k_finish_sums, dim3(ns), dim3(256), 0, lc.stream, sb.d_scenes, 1u, parts, 1u);
        MPM_LAUNCH("particle_sums", k_particle_sums, dim3(parts, ns), dim3(256), 0, lc.stream, sb.d_scenes, steps);
        MPM_LAUNCH("finish_sums", k_finish_sums, dim3(ns), dim3(256), 0, lc.stream, sb.d_scenes, 0u, parts, 2u);
    }
    if (measure) {
        const uint32_t mb = std::max<uint32_t>(1u, std::min<uint32_t>(blocks_for(sb.max_slots, 256),
                                                                       ns == 1 ? 1024u : 32u));
        MPM_LAUNCH("measure_speed", k_measure_speed, dim3(mb, ns), dim3(256), 0, lc.stream, sb.d_scenes, steps);
    }
    return hipSuccess;
}

hipError_t launch_clear_counts(const LaunchCtx& lc, int b) {
    MPM_LAUNCH("clear_counts", k_clear_counts, dim3(std::min<uint32_t>(blocks_for(lc.nb, 256), 1024u)), dim3(256), 0,
               lc.stream, lc.buf[b].cnt, lc.nb);
    return hipSuccess;
}

hipError_t launch_bucket(const LaunchCtx& lc, int b, uint32_t first, uint32_t count) {
    if (count == 0) return hipSuccess;
    const float dx = lc.cell_size;
    auto kern = scene_general(lc.scene_flags) ? k_bucket<true> : k_bucket<false>;
    MPM_LAUNCH("bucket", kern, dim3(blocks_for(count, 256)), dim3(256), 0, lc.stream, lc.ctl, lc.home, lc.mats,
               lc.buf[b], first, count, dx, lc.cells_x, lc.cells_y, lc.bw, lc.cap, lc.aux, b);
    return hipSuccess;
}

hipError_t launch_unbucket(const LaunchCtx& lc, int b) {
    const size_t slots = (size_t)lc.nb * lc.cap;
    MPM_LAUNCH("unbucket", k_unbucket, dim3(std::min<uint32_t>(blocks_for(slots, 256), 2048u)), dim3(256), 0,
               lc.stream, lc.buf[b], lc.home, lc.nb, lc.cap, scene_general(lc.scene_flags) ? 1u : 0u);
    return hipSuccess;
}

// ------------------------------------------------- interaction operations
namespace {

// removeParticlesInOutlet (shaders.ts:3041-3065) on the caller-order arrays:
// a water or oil particle whose position lies in the closed box becomes a
// removed tombstone.  One counter add per wave.
__global__ __launch_bounds__(256) void k_outlet(Home home, uint32_t n, float4 box, uint32_t* removed) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    bool hit = false;
    if (i < n) {
        const uint32_t m = home.model[i];
        if (m == MODEL_WATER || m == MODEL_OIL) {
            const float2 p = home.pos[i];
            hit = p.x >= box.x && p.y >= box.y && p.x <= box.z && p.y <= box.w;
            if (hit) home.model[i] = MODEL_REMOVED;
        }
    }
    const uint64_t mask = __ballot(hit);
    if ((threadIdx.x & 63u) == 0u && mask != 0u) atomicAdd(removed, (uint32_t)__popcll(mask));
}

// pickParticle (shaders.ts:3082-3118) as a parallel arg-min over the
// committed slots: the key (squared distance bits, particle index) orders
// candidates exactly as the reference's serial scan (smaller distance first,
// the lower index on a tie; non-negative float bits order like the floats).
__global__ __launch_bounds__(256) void k_pick(SlotBuf b, uint32_t nb, uint32_t cap, float px, float py, float r2,
                                              unsigned long long* best) {
    unsigned long long key = ~0ull;
    const size_t total = (size_t)nb * cap;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const uint32_t bk = (uint32_t)g / cap;  // slot counts fit 32 bits
        if ((uint32_t)g - bk * cap >= b.cnt[bk]) continue;
        const float2 p = b.pos[g];
        const float ox = p.x - px, oy = p.y - py;
        const float d2 = ox * ox + oy * oy;
        if (d2 <= r2) {
            const unsigned long long k = ((unsigned long long)__float_as_uint(d2) << 32) | b.pid[g];
            key = k < key ? k : key;
        }
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        const uint32_t lo = __shfl_down((uint32_t)key, off, 64), hi = __shfl_down((uint32_t)(key >> 32), off, 64);
        const unsigned long long o = ((unsigned long long)hi << 32) | lo;
        key = o < key ? o : key;
    }
    if ((threadIdx.x & 63u) == 0u && key != ~0ull) atomicMin(best, key);
}

// Position of the picked particle (the one committed slot holding its index).
__global__ __launch_bounds__(256) void k_pick_where(SlotBuf b, uint32_t nb, uint32_t cap,
                                                    const unsigned long long* best, float2* where) {
    const unsigned long long key = *best;
    if (key == ~0ull) return;
    const uint32_t pid = (uint32_t)key;
    const size_t total = (size_t)nb * cap;
    for (size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const uint32_t bk = (uint32_t)g / cap;  // slot counts fit 32 bits
        if ((uint32_t)g - bk * cap >= b.cnt[bk]) continue;
        if (b.pid[g] == pid) *where = b.pos[g];
    }
}

// cloneObject (shaders.ts:3120-3147) on the caller-order arrays: the copy of
// particle src + k goes to dst + k with its position translated and clamped
// to the domain (clamp(p + offset, 0, domain)).
__global__ __launch_bounds__(256) void k_clone(Home home, uint32_t src, uint32_t count, uint32_t dst, float ox,
                                               float oy, float dom_w, float dom_h) {
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= count) return;
    const uint32_t s = src + k, d = dst + k;
    const float2 p = home.pos[s];
    home.pos[d] = make_float2(fminf(fmaxf(p.x + ox, 0.0f), dom_w), fminf(fmaxf(p.y + oy, 0.0f), dom_h));
    home.vel[d] = home.vel[s];
    home.C[d] = home.C[s];
    home.F[d] = home.F[s];
    home.plastic[d] = home.plastic[s];
    home.model[d] = home.model[s];
    home.mat[d] = home.mat[s];
}

}  // namespace

hipError_t launch_occupancy(hipStream_t stream, const SceneDesc* d_scene, uint32_t tiles, bool general,
                            KernelTimes* timing) {
    struct {
        hipStream_t stream;
        KernelTimes* timing;
    } lc = {stream, timing};
    auto fn = general ? k_p2g_gather<28, true, false, true> : k_p2g_gather<28, false, false, true>;
    MPM_LAUNCH("occupancy", fn, dim3(tiles), dim3(P2GTile<28>::THREADS), 0, lc.stream, d_scene, nullptr, 1u, 0u, TileQueue{});
    return hipSuccess;
}

hipError_t launch_outlet(const LaunchCtx& lc, float4 box, uint32_t* removed) {
    if (lc.n == 0) return hipSuccess;
    MPM_LAUNCH("outlet", k_outlet, dim3(blocks_for(lc.n, 256)), dim3(256), 0, lc.stream, lc.home, lc.n, box, removed);
    return hipSuccess;
}

hipError_t launch_pick(const LaunchCtx& lc, int b, float px, float py, float r2, unsigned long long* best,
                       float2* where) {
    const size_t slots = (size_t)lc.nb * lc.cap;
    const uint32_t blocks = std::max<uint32_t>(1u, std::min<uint32_t>(blocks_for(slots, 256), 2048u));
    MPM_LAUNCH("pick", k_pick, dim3(blocks), dim3(256), 0, lc.stream, lc.buf[b], lc.nb, lc.cap, px, py, r2, best);
    MPM_LAUNCH("pick_where", k_pick_where, dim3(blocks), dim3(256), 0, lc.stream, lc.buf[b], lc.nb, lc.cap, best,
               where);
    return hipSuccess;
}

hipError_t launch_clone(const LaunchCtx& lc, uint32_t src, uint32_t count, uint32_t dst, float ox, float oy,
                        float dom_w, float dom_h) {
    if (count == 0) return hipSuccess;
    MPM_LAUNCH("clone", k_clone, dim3(blocks_for(count, 256)), dim3(256), 0, lc.stream, lc.home, src, count, dst, ox,
               oy, dom_w, dom_h);
    return hipSuccess;
}

hipError_t launch_keys(const LaunchCtx& lc, uint32_t* keys) {
    if (lc.n == 0) return hipSuccess;
    MPM_LAUNCH("keys", k_keys, dim3(blocks_for(lc.n, 256)), dim3(256), 0, lc.stream, lc.home, keys, lc.n,
               lc.cell_size, lc.cells_x, lc.cells_y, lc.node_w);
    return hipSuccess;
}

}  // namespace mpm
