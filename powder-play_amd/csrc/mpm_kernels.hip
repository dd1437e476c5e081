// gfx950 kernels of the 2D MLS-MPM microstep.
//
// One microstep is six launches on one stream:
//   scan_reduce -> scan_blocks -> scan_final   bin starts from per-node counts
//   scatter                                    particle indices grouped by bin
//   p2g_grid<FIXED>                            LDS-staged gather P2G + grid update
//   g2p                                        G2P, F update, projections, CFL guard,
//                                              and the next microstep's bin counts
// framed per batch by batch_begin + bin (first counts) and commit_last +
// fix_rollback + measure_speed.  Numerical semantics follow the WGSL kernels
// of src/lib/mpm/kernels/shaders.ts (cited per function); the fixed-point
// path reproduces their f32 operation order exactly (the library is built
// with -ffp-contract=off) so its int32 accumulators are bit-exact.
#include <float.h>
#include <math.h>

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
constexpr float MAX_DEFORMATION_INCREMENT = 0.25f;
constexpr uint32_t MAX_DEFORMATION_SUBSTEPS = 64u;

// P2G tile geometry: a 1024-thread block computes node sums for a 32x32
// patch (a one-node ring around its 30x30 output tile, needed by the liquid
// contact's volume gradient) from the particles binned to the 34x34 base
// cells that reach it, staged through LDS in chunks of CAP particles.
constexpr int TS = 32;
constexpr int TI = TS - 2;
constexpr int RC = TS + 2;
constexpr int CAP = 1280;
constexpr int P2G_THREADS = TS * TS;

constexpr int SCAN_THREADS = 1024;
constexpr int SCAN_ITEMS = 4;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

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
__device__ __forceinline__ void raise_flags(Control* c, uint32_t bits) { atomicOr(&c->status[0], bits); }
__device__ __forceinline__ void raise_at(Control* c, uint32_t bits, uint32_t index) {
    atomicMin(&c->status[1], index);
    atomicOr(&c->status[0], bits);
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

// prepareParticle (shaders.ts:808-963): the per-particle affine momentum
// A = m*C + (-dt*V*4/dx^2) * P*F^T plus (mass, volume, det F, phase) for P2G.
// GENERAL = false handles scenes of tension-free liquids only (pure f32);
// GENERAL = true adds Tait liquids (FP64 pow) and the solid models (FP64).
// Returns the status bits to raise.
template <bool GENERAL>
__device__ uint32_t prepare_particle(const Params& P, uint32_t model, float4 C, float4 F, float lambda, float mu,
                                     float density, float volume, float param, float pv, float4& A, float4& aux) {
    float det = det_comp(F);
    if (!finitef(det) || det <= 0.0f || !finitef(volume) || volume <= 0.0f) return ERR_DEFORMATION;
    float4 S = make_float4(0.f, 0.f, 0.f, 0.f);
    const bool liquid = model == MODEL_WATER || model == MODEL_OIL;
    if (liquid) {
        if (!finitef(density) || density <= 0.0f || !finitef(lambda) || lambda <= 0.0f || !finitef(param) ||
            param < 0.0f)
            return ERR_P2G | ERR_INVALID_MATERIAL;
        if (param > 0.0f) {  // Tait: -J * B * (J^-n - 1)
            if constexpr (!GENERAL) return ERR_P2G | ERR_INVALID_MATERIAL;
            float inv = 1.0f / det;
            float pw = (float)pow((double)inv, (double)param);
            float pressure = lambda * (pw - 1.0f);
            float stress = -det * pressure;
            S.x = stress; S.w = stress;
        } else {             // tension-free linear: lambda * min(J - 1, 0)
            float jm1 = det - 1.0f;
            float pressure = lambda * (jm1 < 0.0f ? jm1 : 0.0f);
            S.x = pressure; S.w = pressure;
        }
    } else {
        if constexpr (!GENERAL) {
            return ERR_P2G | ERR_INVALID_MATERIAL;
        } else {
            if (model == MODEL_NEO_HOOKEAN) {
                if (!finitef(param) || param < 0.0f || param > 1.0f) return ERR_P2G | ERR_INVALID_MATERIAL;
                if (!solid_stress(model, F, lambda, mu, param, pv, S)) return ERR_P2G | ERR_NONFINITE_PARTICLE;
            } else if (model == MODEL_SNOW) {
                if (!finitef(pv) || pv < 0.6f || pv > 20.0f) return ERR_P2G | ERR_INVALID_MATERIAL;
                if (!solid_stress(model, F, lambda, mu, param, pv, S)) return ERR_P2G | ERR_INVALID_MATERIAL;
            } else if (model == MODEL_POWDER) {
                if (!finitef(lambda) || !finitef(mu) || mu <= 0.0f || !finitef(param) || param < 0.0f)
                    return ERR_P2G | ERR_INVALID_MATERIAL;
                if (!solid_stress(model, F, lambda, mu, param, pv, S)) return ERR_P2G | ERR_NONFINITE_PARTICLE;
            } else {
                return ERR_P2G | ERR_INVALID_MATERIAL;
            }
        }
    }
    const float mass = liquid ? density * volume : P.particle_mass;
    const float ss = (-P.dt * volume) * P.grad_scale;
    A.x = mass * C.x + ss * S.x;
    A.y = mass * C.y + ss * S.y;
    A.z = mass * C.z + ss * S.z;
    A.w = mass * C.w + ss * S.w;
    if (!finite4(S) || !finite4(A)) return ERR_P2G | ERR_NONFINITE_PARTICLE;
    const uint32_t phase = model == MODEL_WATER ? 1u : (model == MODEL_OIL ? 2u : 0u);
    aux = make_float4(mass, volume, det, __uint_as_float(phase));
    return 0;
}

// Stencil base and fraction (shaders.ts:667-676, :1023-1028).
struct Stencil { int bx, by; float fx, fy; };
__device__ __forceinline__ bool make_stencil(const Params& P, float2 p, Stencil& s) {
    float gx = p.x / P.dx, gy = p.y / P.dx;
    if (!finitef(gx) || !finitef(gy)) return false;
    s.bx = (int)floorf(gx - 0.5f);
    s.by = (int)floorf(gy - 0.5f);
    // stencilFitsHalo (shaders.ts:511-516)
    if (s.bx < -1 || s.by < -1 || s.bx + 2 > (int)P.cells_x + 1 || s.by + 2 > (int)P.cells_y + 1) return false;
    s.fx = gx - (float)s.bx;
    s.fy = gy - (float)s.by;
    return true;
}
// quadraticWeights (shaders.ts:500-509), one component.
__device__ __forceinline__ float qweight(float f, int o) {
    float l = 1.5f - f, c = f - 1.0f, r = f - 0.5f;
    float w0 = (0.5f * l) * l, w1 = 0.75f - c * c, w2 = (0.5f * r) * r;
    return o == 0 ? w0 : (o == 1 ? w1 : w2);
}

// ------------------------------------------------------------ batch frame
__global__ void k_batch_begin(Control* ctl, uint32_t* count, uint32_t nodes) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    for (uint32_t k = i; k < nodes; k += gridDim.x * blockDim.x) count[k] = 0u;
    if (i == 0) {
        ctl->completed = 0u;
        ctl->max_speed_bits = 0u;
        ctl->deferred = 0u;
        ctl->deferred_index = 0xffffffffu;
    }
}

// insertParticleBucket binning (shaders.ts:657-693) as a counting sort:
// each live particle counts itself into the bin of its stencil base node and
// fills its prepared affine (prepareAndBuildBuckets, shaders.ts:965-969).
template <bool GENERAL>
__global__ __launch_bounds__(256) void k_bin(const Params* __restrict__ Pp, Control* ctl, ParticleBuf in,
                                             ParticleStatic st, Bins bins, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (load_status(ctl)) return;
    const Params& P = *Pp;
    const uint32_t model = st.model[i];
    if (model == MODEL_REMOVED) { bins.key[i] = INVALID_KEY; return; }
    float4 A, aux;
    const uint32_t pf = prepare_particle<GENERAL>(P, model, in.C[i], in.F[i], st.lambda[i], st.mu[i], st.density[i],
                                                  st.volume[i], st.param[i], in.plastic[i], A, aux);
    if (pf) {
        bins.key[i] = INVALID_KEY;
        if (pf & ERR_P2G) raise_at(ctl, pf, i); else raise_flags(ctl, pf);
        return;
    }
    bins.affine[i] = A;
    bins.aux[i] = aux;
    Stencil s;
    if (!make_stencil(P, in.pos[i], s)) { bins.key[i] = INVALID_KEY; raise_flags(ctl, ERR_STENCIL); return; }
    uint32_t k = (uint32_t)(s.by + 1) * P.node_w + (uint32_t)(s.bx + 1);
    bins.key[i] = k;
    bins.rank[i] = atomicAdd(&bins.count[k], 1u);
}

// ------------------------------------------------------------------- scan
__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t t = __shfl_up(v, off, 64);
        if (lane >= off) v += t;
    }
    return v;
}
__device__ __forceinline__ uint32_t wave_sum(uint32_t v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
    return v;
}

// Per-tile totals of the bin counts; block 0 also commits the preceding
// microstep (commitParticles' progress counter, shaders.ts:3067-3080).
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(Control* ctl, const uint32_t* __restrict__ count,
                                                               uint32_t* __restrict__ block_sums, uint32_t nodes,
                                                               uint32_t commit_prev) {
    __shared__ uint32_t s_w[SCAN_THREADS / 64];
    if (block_status_set(ctl)) return;
    if (commit_prev && blockIdx.x == 0 && threadIdx.x == 0) {
        // the preceding microstep committed; failures its G2P found while
        // preparing this microstep's particles belong to this microstep
        ctl->completed += 1u;
        const uint32_t d = ctl->deferred;
        if (d) {
            if (d & ERR_P2G) atomicMin(&ctl->status[1], ctl->deferred_index);
            atomicOr(&ctl->status[0], d);
        }
    }
    uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t v = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k)
        if (base + k < nodes) v += count[base + k];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s_w[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t t = threadIdx.x < SCAN_THREADS / 64 ? s_w[threadIdx.x] : 0u;
        t = wave_sum(t);
        if (threadIdx.x == 0) block_sums[blockIdx.x] = t;
    }
}

// Exclusive scan of the tile totals (one block).
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_blocks(Control* ctl, uint32_t* block_sums, uint32_t nb) {
    __shared__ uint32_t s_w[SCAN_THREADS / 64];
    __shared__ uint32_t s_carry;
    if (block_status_set(ctl)) return;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (uint32_t base = 0; base < nb; base += SCAN_THREADS) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < nb ? block_sums[i] : 0u;
        uint32_t inc = wave_inclusive_scan(v);
        if ((threadIdx.x & 63) == 63) s_w[threadIdx.x >> 6] = inc;
        __syncthreads();
        if (threadIdx.x < 64) {
            uint32_t t = threadIdx.x < SCAN_THREADS / 64 ? s_w[threadIdx.x] : 0u;
            uint32_t ti = wave_inclusive_scan(t);
            if (threadIdx.x < SCAN_THREADS / 64) s_w[threadIdx.x] = ti - t;
        }
        __syncthreads();
        uint32_t carry = s_carry;
        uint32_t excl = carry + s_w[threadIdx.x >> 6] + inc - v;
        if (i < nb) block_sums[i] = excl;
        __syncthreads();
        if (threadIdx.x == SCAN_THREADS - 1) s_carry = excl + v;
        __syncthreads();
    }
}

// Bin starts: exclusive scan inside each tile plus the tile offset; clears
// the counts for the next binning pass.
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_final(Control* ctl, uint32_t* __restrict__ count,
                                                              const uint32_t* __restrict__ block_sums,
                                                              uint32_t* __restrict__ start, uint32_t nodes) {
    __shared__ uint32_t s_w[SCAN_THREADS / 64];
    if (block_status_set(ctl)) return;
    uint32_t base = blockIdx.x * SCAN_TILE + threadIdx.x * SCAN_ITEMS;
    uint32_t c[SCAN_ITEMS];
    uint32_t local = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        c[k] = base + k < nodes ? count[base + k] : 0u;
        local += c[k];
    }
    uint32_t inc = wave_inclusive_scan(local);
    if ((threadIdx.x & 63) == 63) s_w[threadIdx.x >> 6] = inc;
    __syncthreads();
    if (threadIdx.x < 64) {
        uint32_t t = threadIdx.x < SCAN_THREADS / 64 ? s_w[threadIdx.x] : 0u;
        uint32_t ti = wave_inclusive_scan(t);
        if (threadIdx.x < SCAN_THREADS / 64) s_w[threadIdx.x] = ti - t;
    }
    __syncthreads();
    uint32_t run = block_sums[blockIdx.x] + s_w[threadIdx.x >> 6] + inc - local;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; ++k) {
        if (base + k < nodes) {
            start[base + k] = run;
            count[base + k] = 0u;
        }
        run += c[k];
    }
    // the thread covering the last node also writes the end sentinel
    if (nodes > 0 && base <= nodes - 1 && nodes - 1 < base + SCAN_ITEMS) start[nodes] = run;
}

__global__ __launch_bounds__(256) void k_scatter(Control* ctl, const uint32_t* __restrict__ key,
                                                 const uint32_t* __restrict__ rank, const uint32_t* __restrict__ start,
                                                 uint32_t* __restrict__ perm, uint32_t n) {
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (load_status(ctl)) return;
    uint32_t k = key[i];
    if (k != INVALID_KEY) perm[start[k] + rank[i]] = i;
}

// ------------------------------------------------------------ P2G + grid
// finishGridVelocity (shaders.ts:1748-1777)
__device__ __forceinline__ float2 finish_velocity(const Params& P, float m, float qx, float qy, int nx, int ny) {
    if (m <= 0.0f) return make_float2(0.f, 0.f);
    float2 v = make_float2(qx / m + P.dt * P.gx, qy / m + P.dt * P.gy);
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
// applyWallCushion / applyWallTangentialFriction (shaders.ts:1438-1457)
__device__ __forceinline__ float wall_cushion(float v, int coord, int extent) {
    float cushion = fminf(WALL_CUSHION_CELLS, (float)extent / 4.0f);
    float dist = v < 0.0f ? (float)coord : (float)(extent - coord);
    return v * fminf(fmaxf(dist / cushion, 0.0f), 1.0f);
}
__device__ __forceinline__ float wall_friction(float v, int coord, int extent, float friction) {
    if (friction <= 0.0f) return v;
    float dist = fmaxf(0.0f, fminf((float)coord, (float)(extent - coord)));
    float cushion = fminf(WALL_CUSHION_CELLS, (float)extent / 4.0f);
    float infl = 1.0f - fminf(fmaxf(dist / cushion, 0.0f), 1.0f);
    return v * (1.0f - friction * infl);
}

__device__ __forceinline__ uint32_t fixed_round(float value, int32_t& out) {
    float scaled = rintf(value * FIXED_POINT_SCALE);  // addFixedPoint (shaders.ts:971-985)
    if (!finitef(scaled) || fabsf(scaled) > FIXED_POINT_LIMIT) { out = 0; return ERR_FIXED_POINT; }
    out = (int32_t)scaled;
    return 0u;
}

struct Staged {
    float fx[CAP], fy[CAP], m[CAP], vx[CAP], vy[CAP], a0[CAP], a1[CAP], a2[CAP], a3[CAP], vol[CAP], det[CAP];
    uint32_t tag[CAP];  // region column | phase << 8 (phase 0 solid, 1 water, 2 oil)
};

template <bool FIXED>
__global__ __launch_bounds__(P2G_THREADS) void k_p2g_grid(const Params* __restrict__ Pp, Control* ctl, ParticleBuf in,
                                                        ParticleStatic st, Bins bins, Grid grid, uint32_t tiles_x,
                                                        uint32_t keep_fixed, uint32_t scene_flags) {
    __shared__ uint32_t s_start[RC][RC + 1];
    __shared__ uint32_t s_rowoff[RC + 1];
    __shared__ Staged s;
    __shared__ float s_vw[TS * TS];
    __shared__ float s_vo[TS * TS];
    __shared__ uint32_t s_cut;

    if (block_status_set(ctl)) return;
    const Params& P = *Pp;
    const int tid = threadIdx.x;
    const int lx = tid % TS, ly = tid / TS;
    const int X0 = (int)(blockIdx.x % tiles_x) * TI, Y0 = (int)(blockIdx.x / tiles_x) * TI;
    const int W = (int)P.node_w, H = (int)P.node_h;
    const int rx0 = X0 - 3, ry0 = Y0 - 3;
    const int cx_lo = max(rx0, 0), cx_hi = min(rx0 + RC, W);
    const int sx = X0 - 1 + lx, sy = Y0 - 1 + ly;  // this thread's node (storage coords)
    const bool node_ok = sx >= 0 && sx < W && sy >= 0 && sy < H;

    for (int e = tid; e < RC * (RC + 1); e += P2G_THREADS) {
        int r = e / (RC + 1), c = e % (RC + 1);
        int cy = ry0 + r;
        uint32_t v = 0;
        if (cy >= 0 && cy < H && cx_lo < cx_hi) {
            int cx = min(max(rx0 + c, cx_lo), cx_hi);
            v = bins.start[(uint32_t)cy * (uint32_t)W + (uint32_t)cx];
        }
        s_start[r][c] = v;
    }
    __syncthreads();
    if (tid == 0) {
        uint32_t acc = 0;
        for (int r = 0; r < RC; ++r) {
            s_rowoff[r] = acc;
            acc += s_start[r][RC] - s_start[r][0];
        }
        s_rowoff[RC] = acc;
    }
    __syncthreads();
    const uint32_t total = s_rowoff[RC];

    // accumulators: solid (m, qx, qy), water (m, qx, qy, V), oil (m, qx, qy, V)
    float f_sm = 0, f_sqx = 0, f_sqy = 0, f_wm = 0, f_wqx = 0, f_wqy = 0, f_wv = 0, f_om = 0, f_oqx = 0, f_oqy = 0,
          f_ov = 0;
    long long i_acc[11];
#pragma unroll
    for (int k = 0; k < 11; ++k) i_acc[k] = 0;
    uint32_t flags = 0;
    const float mref = P.particle_mass, vref = P.particle_volume, dx = P.dx;

    for (uint32_t c0 = 0; c0 < total;) {
        uint32_t c1 = total;
        bool split = false;
        if (total - c0 > (uint32_t)CAP) {
            if (tid == 0) s_cut = c0;
            __syncthreads();
            for (int e = tid; e < RC * RC; e += P2G_THREADS) {
                int r = e / RC, c = e % RC;
                uint32_t bnd = s_rowoff[r] + (s_start[r][c + 1] - s_start[r][0]);
                if (bnd > c0 && bnd <= c0 + (uint32_t)CAP) atomicMax(&s_cut, bnd);
            }
            __syncthreads();
            c1 = s_cut;
            if (c1 == c0) { c1 = c0 + (uint32_t)CAP; split = true; }
            __syncthreads();
        }
        // stage [c0, c1): each live particle lands at its bin's LDS range in
        // ascending particle-index order, which fixes the summation order.
        for (uint32_t j = c0 + tid; j < c1; j += P2G_THREADS) {
            int lo = 0, hi = RC - 1;
            while (lo < hi) {  // last row r with s_rowoff[r] <= j
                int mid = (lo + hi + 1) >> 1;
                if (s_rowoff[mid] <= j) lo = mid; else hi = mid - 1;
            }
            const int r = lo;
            const uint32_t slot = s_start[r][0] + (j - s_rowoff[r]);
            const uint32_t idx = bins.perm[slot];
            const float2 p = in.pos[idx];
            Stencil sc = {0, 0, 0.0f, 0.0f};
            uint32_t li = j - c0;
            uint32_t tag = 0xffffu;  // contributes nothing unless staged below
            if (!make_stencil(P, p, sc) || sc.by + 1 - ry0 != r || sc.bx + 1 - rx0 < 0 || sc.bx + 1 - rx0 >= RC) {
                flags |= ERR_STENCIL;
            } else {
                int c = sc.bx + 1 - rx0;
                if (!split) {
                    const uint32_t key = (uint32_t)(sc.by + 1) * (uint32_t)W + (uint32_t)(sc.bx + 1);
                    const uint32_t cs = bins.start[key], ce = bins.start[key + 1];
                    uint32_t lr = 0;
                    for (uint32_t q = cs; q < ce; ++q) lr += bins.perm[q] < idx ? 1u : 0u;
                    li = s_rowoff[r] + (cs - s_start[r][0]) + lr - c0;
                    if (li >= c1 - c0) {  // bin and position disagree: never expected
                        flags |= ERR_STENCIL;
                        li = j - c0;
                        c = 0xff;
                    }
                }
                const float2 v = in.vel[idx];
                const float4 A = bins.affine[idx];
                const float4 aux = bins.aux[idx];
                tag = c == 0xff ? 0xffffu : ((uint32_t)c | (__float_as_uint(aux.w) << 8));
                s.vx[li] = v.x;
                s.vy[li] = v.y;
                s.m[li] = aux.x;
                s.vol[li] = aux.y;
                s.det[li] = aux.z;
                s.a0[li] = A.x; s.a1[li] = A.y; s.a2[li] = A.z; s.a3[li] = A.w;
            }
            s.fx[li] = sc.fx;
            s.fy[li] = sc.fy;
            s.tag[li] = tag;
        }
        __syncthreads();

        if (node_ok) {
#pragma unroll 1
            for (int dy = 0; dy < 3; ++dy) {
                const int r = ly + dy;
                const uint32_t rb = s_rowoff[r] - s_start[r][0];
                uint32_t b = rb + s_start[r][lx], e = rb + s_start[r][lx + 3];
                b = max(b, c0);
                e = min(e, c1);
                const int oy = 2 - dy;
#pragma unroll 1
                for (uint32_t q = b; q < e; ++q) {
                    const uint32_t li = q - c0;
                    const uint32_t tag = s.tag[li];
                    const int c = (int)(tag & 0xffu);
                    if (c == 0xff) continue;
                    const uint32_t ph = tag >> 8;
                    const int ox = lx - c + 2;
                    const float fx = s.fx[li], fy = s.fy[li];
                    const float w = qweight(fx, ox) * qweight(fy, oy);
                    const float dpx = ((float)ox - fx) * dx;
                    const float dpy = ((float)oy - fy) * dx;
                    const float m = s.m[li];
                    const float qx = (m * s.vx[li] + s.a0[li] * dpx) + s.a1[li] * dpy;
                    const float qy = (m * s.vy[li] + s.a2[li] * dpx) + s.a3[li] * dpy;
                    if (!finitef(w) || w < 0.0f || !finitef(qx) || !finitef(qy)) {
                        flags |= ERR_P2G | ERR_NONFINITE_PARTICLE;
                        continue;
                    }
                    if constexpr (FIXED) {
                        int32_t im, iqx, iqy, iv = 0;
                        flags |= fixed_round((w * m) / mref, im);
                        flags |= fixed_round((w * qx) / mref, iqx);
                        flags |= fixed_round((w * qy) / mref, iqy);
                        if (ph != 0u) flags |= fixed_round((w * (s.vol[li] * s.det[li])) / vref, iv);
                        const int base = ph == 0u ? 0 : (ph == 1u ? 3 : 7);
#pragma unroll
                        for (int k = 0; k < 11; ++k) {
                            const int rel = k - base;
                            long long add = rel == 0 ? im : rel == 1 ? iqx : rel == 2 ? iqy : rel == 3 ? iv : 0;
                            if (rel < 0 || rel > 3 || (ph == 0u && rel == 3)) add = 0;
                            i_acc[k] += add;
                        }
                    } else {
                        const float wm = w * m, wqx = w * qx, wqy = w * qy;
                        if (ph == 0u) {
                            f_sm += wm; f_sqx += wqx; f_sqy += wqy;
                        } else {
                            const float wv = (w * s.vol[li]) * s.det[li];
                            if (ph == 1u) { f_wm += wm; f_wqx += wqx; f_wqy += wqy; f_wv += wv; }
                            else { f_om += wm; f_oqx += wqx; f_oqy += wqy; f_ov += wv; }
                        }
                    }
                }
            }
        }
        __syncthreads();
        c0 = c1;
    }

    if constexpr (FIXED) {
        // gatherAtomicGrid decode (shaders.ts:1045-1070) of the wrapped words;
        // an exact sum outside int32 is the accumulator overflow of 65536.
        float vals[11];
#pragma unroll
        for (int k = 0; k < 11; ++k) {
            if (i_acc[k] > 2147483647LL || i_acc[k] < -2147483648LL) flags |= ERR_FIXED_POINT;
            const int32_t word = (int32_t)(uint32_t)(unsigned long long)i_acc[k];
            vals[k] = (float)word / FIXED_POINT_SCALE;
        }
        f_sm = vals[0] * mref; f_sqx = vals[1] * mref; f_sqy = vals[2] * mref;
        f_wm = vals[3] * mref; f_wqx = vals[4] * mref; f_wqy = vals[5] * mref; f_wv = vals[6] * vref;
        f_om = vals[7] * mref; f_oqx = vals[8] * mref; f_oqy = vals[9] * mref; f_ov = vals[10] * vref;
    }
    s_vw[tid] = node_ok ? f_wv : 0.0f;
    s_vo[tid] = node_ok ? f_ov : 0.0f;
    __syncthreads();

    const bool interior = node_ok && lx >= 1 && lx <= TI && ly >= 1 && ly <= TI;
    if (interior) {
        const uint32_t node = (uint32_t)sy * (uint32_t)W + (uint32_t)sx;
        if constexpr (FIXED) {
            if (keep_fixed) {
#pragma unroll
                for (int k = 0; k < 11; ++k)
                    grid.fixed[(size_t)node * 11 + k] = (int32_t)(uint32_t)(unsigned long long)i_acc[k];
            }
        }
        if (!finitef(f_sm) || !finitef(f_sqx) || !finitef(f_sqy) || !finitef(f_wm) || !finitef(f_wqx) ||
            !finitef(f_wqy) || !finitef(f_wv) || !finitef(f_om) || !finitef(f_oqx) || !finitef(f_oqy) ||
            !finitef(f_ov)) {
            flags |= FIXED ? ERR_GRID : (ERR_P2G | ERR_NONFINITE_ACCUMULATION);
        } else if (f_sm < 0.0f || f_wm < 0.0f || f_wv < 0.0f || f_om < 0.0f || f_ov < 0.0f) {
            flags |= ERR_GRID;
        } else {
            // updateGrid (shaders.ts:1485-1746)
            const int nx = sx - 1, ny = sy - 1;
            float2 vs = finish_velocity(P, f_sm, f_sqx, f_sqy, nx, ny);
            float2 vw = finish_velocity(P, f_wm, f_wqx, f_wqy, nx, ny);
            float2 vo = finish_velocity(P, f_om, f_oqx, f_oqy, nx, ny);
            if (nx >= 0 && ny >= 0 && nx <= (int)P.cells_x && ny <= (int)P.cells_y && f_wm > 0.0f && f_om > 0.0f &&
                f_wv > 0.0f && f_ov > 0.0f) {
                // phaseVolumeGradient (shaders.ts:1779-1799) from the LDS ring
                const float wgx = s_vw[tid + 1] - s_vw[tid - 1], wgy = s_vw[tid + TS] - s_vw[tid - TS];
                const float ogx = s_vo[tid + 1] - s_vo[tid - 1], ogy = s_vo[tid + TS] - s_vo[tid - TS];
                const float sepx = ogx - wgx, sepy = ogy - wgy;
                const float len = sqrtf(sepx * sepx + sepy * sepy);
                if (finitef(len) && len > 1e-8f) {
                    float nrx = sepx / len, nry = sepy / len;
                    const float wd = f_wm / f_wv, od = f_om / f_ov;
                    const float dscale = fmaxf(wd, od);
                    const float contrast = (od - wd) / dscale;
                    const float gmag = sqrtf(P.gx * P.gx + P.gy * P.gy);
                    float align = 0.0f;
                    if (!finitef(wd) || wd <= 0.0f || !finitef(od) || od <= 0.0f || !finitef(contrast) ||
                        !finitef(gmag)) {
                        flags |= ERR_GRID;
                    } else {
                        if (gmag > 0.0f) align = (nrx * P.gx + nry * P.gy) / gmag;
                        if (!finitef(align)) flags |= ERR_GRID;
                        if (contrast * align < -LIQUID_STABLE_ORDER_EPSILON) { nrx = -nrx; nry = -nry; }
                        const float rel = (vo.x - vw.x) * nrx + (vo.y - vw.y) * nry;
                        const float overlap =
                            fminf(fmaxf(2.0f * fminf(f_wv, f_ov) / (f_wv + f_ov), 0.0f), 1.0f);
                        const float minsep = fmaxf(
                            LIQUID_CONTACT_SEPARATION_FRACTION * overlap * dx / LIQUID_CONTACT_RELAXATION_SECONDS,
                            fmaxf(-rel, 0.0f) * LIQUID_CONTACT_APPROACH_SPEED_MULTIPLIER);
                        if (rel < minsep) {
                            const float imp = (minsep - rel) / (1.0f / f_wm + 1.0f / f_om);
                            const float ix = imp * nrx, iy = imp * nry;
                            vw.x = vw.x - ix / f_wm; vw.y = vw.y - iy / f_wm;
                            vo.x = vo.x + ix / f_om; vo.y = vo.y + iy / f_om;
                        }
                    }
                }
            }
            if (!finite2(vs) || !finite2(vw) || !finite2(vo)) {
                flags |= ERR_GRID | ERR_TRIANGLE_CONTACT;
            } else {
                const int cxn = (int)P.cells_x, cyn = (int)P.cells_y;
                vs.x = wall_cushion(vs.x, nx, cxn); vs.y = wall_cushion(vs.y, ny, cyn);
                vw.x = wall_cushion(vw.x, nx, cxn); vw.y = wall_cushion(vw.y, ny, cyn);
                vo.x = wall_cushion(vo.x, nx, cxn); vo.y = wall_cushion(vo.y, ny, cyn);
                const float fr = P.wall_friction;
                vs.x = wall_friction(vs.x, ny, cyn, fr); vs.y = wall_friction(vs.y, nx, cxn, fr);
                vw.x = wall_friction(vw.x, ny, cyn, fr); vw.y = wall_friction(vw.y, nx, cxn, fr);
                vo.x = wall_friction(vo.x, ny, cyn, fr); vo.y = wall_friction(vo.y, nx, cxn, fr);
                if (scene_flags & SCENE_SOLID) grid.v_solid[node] = vs;
                if (scene_flags & SCENE_WATER) grid.v_water[node] = vw;
                if (scene_flags & SCENE_OIL) grid.v_oil[node] = vo;
                if (scene_flags & (SCENE_WATER | SCENE_OIL)) grid.liquid[node] = make_float2(f_wm + f_om, f_wv + f_ov);
            }
        }
    }
    if (flags) raise_flags(ctl, flags);
}

// -------------------------------------------------------------------- G2P
// integrateDeformation (shaders.ts:1459-1483)
__device__ __forceinline__ float4 integrate_deformation(float4 F, float4 C, float dt, uint32_t substeps) {
    const float h = dt / (float)substeps;
    const float4 u = make_float4(1.0f + h * C.x, h * C.y, h * C.z, 1.0f + h * C.w);
    float4 r = F;
    for (uint32_t k = 0; k < substeps; ++k) {
        r = make_float4(fmaf(u.x, r.x, u.y * r.z), fmaf(u.x, r.y, u.y * r.w), fmaf(u.z, r.x, u.w * r.z),
                        fmaf(u.z, r.y, u.w * r.w));
    }
    return r;
}

// projectLiquidDeformation (shaders.ts:474-488)
__device__ __forceinline__ float4 project_liquid(float4 m) {
    float nx = m.x + m.w, ny = m.z - m.y;
    float sc = fmaxf(fabsf(nx), fabsf(ny));
    float ax = nx / sc, ay = ny / sc;
    float len = sqrtf(ax * ax + ay * ay);
    float rx = ax / len, ry = ay / len;
    float vol = det_comp(m);
    float bv = fminf(fmaxf(vol, MIN_LIQUID_DETERMINANT), MAX_LIQUID_DETERMINANT);
    float s = sqrtf(bv);
    return make_float4(s * rx, s * -ry, s * ry, s * rx);
}

// particleWaveSpeedBound (shaders.ts:268-353)
template <bool GENERAL>
__device__ float wave_speed_bound(const Params& P, uint32_t model, float4 F, float lambda, float mu, float density,
                                  float volume, float param) {
    if (model == MODEL_WATER || model == MODEL_OIL) {
        if (!finitef(lambda) || !finitef(density) || lambda <= 0.0f || density <= 0.0f) return FLT_MAX;
        float ws2 = lambda / density;
        if (param > 0.0f) {
            if constexpr (!GENERAL) return FLT_MAX;
            float det = det_comp(F);
            if (!finitef(param) || param <= 0.0f || !finitef(det) || det <= 0.0f) return FLT_MAX;
            ws2 = param * ws2 * (float)pow((double)(1.0f / det), (double)(param - 1.0f));
        }
        if (!finitef(ws2) || ws2 < 0.0f) return FLT_MAX;
        return sqrtf(ws2);
    }
    if constexpr (!GENERAL) return FLT_MAX;
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

__device__ __forceinline__ float2 confine_axis(float x, float v, float maximum) {  // shaders.ts:522-534
    if (x < 0.0f || (x == 0.0f && v < 0.0f)) return make_float2(0.0f, 0.0f);
    if (x > maximum || (x == maximum && v > 0.0f)) return make_float2(maximum, 0.0f);
    return make_float2(x, v);
}

// g2p (shaders.ts:2114-2421) + validateCflParticle (:2423-2454), then the
// bin count of the updated particle for the next microstep.
template <bool GENERAL>
__global__ __launch_bounds__(256) void k_g2p(const Params* __restrict__ Pp, Control* ctl, ParticleBuf in,
                                             ParticleBuf out, ParticleStatic st, Grid grid, Bins bins, uint32_t n,
                                             uint32_t bin_next) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (load_status(ctl)) return;
    const Params& P = *Pp;
    const uint32_t model = st.model[i];
    if (model == MODEL_REMOVED) {
        out.pos[i] = in.pos[i];
        out.vel[i] = in.vel[i];
        out.C[i] = in.C[i];
        out.F[i] = in.F[i];
        out.plastic[i] = in.plastic[i];
        if (bin_next) bins.key[i] = INVALID_KEY;
        return;
    }
    const float2 p = in.pos[i];
    const float4 F = in.F[i];
    Stencil s;
    if (!make_stencil(P, p, s)) { raise_flags(ctl, ERR_STENCIL); return; }
    const bool water = model == MODEL_WATER, oil = model == MODEL_OIL, liquid = water || oil;
    const float2* __restrict__ gv = water ? grid.v_water : (oil ? grid.v_oil : grid.v_solid);
    const uint32_t W = P.node_w;
    float vx = 0.f, vy = 0.f, a0 = 0.f, a1 = 0.f, a2 = 0.f, a3 = 0.f, pm = 0.f, pvol = 0.f;
    const float dx = P.dx;
#pragma unroll
    for (int oy = 0; oy < 3; ++oy) {
        const float wy = qweight(s.fy, oy);
        const float dpy = ((float)oy - s.fy) * dx;
        const uint32_t row = (uint32_t)(s.by + oy + 1) * W + (uint32_t)(s.bx + 1);
#pragma unroll
        for (int ox = 0; ox < 3; ++ox) {
            const uint32_t node = row + (uint32_t)ox;
            const float w = qweight(s.fx, ox) * wy;
            const float dpx = ((float)ox - s.fx) * dx;
            const float2 g = gv[node];
            if (!finitef(w) || w < 0.0f || !finitef(dpx) || !finitef(dpy) || !finite2(g)) {
                raise_flags(ctl, ERR_G2P);
                return;
            }
            vx = vx + w * g.x;
            vy = vy + w * g.y;
            if (liquid) {
                const float2 lq = grid.liquid[node];
                if (!finitef(lq.x) || lq.x < 0.0f || !finitef(lq.y) || lq.y < 0.0f) {
                    raise_at(ctl, ERR_G2P, i);
                    return;
                }
                pm = pm + w * lq.x;
                pvol = pvol + w * lq.y;
            }
            a0 = a0 + (w * g.x) * dpx;
            a1 = a1 + (w * g.x) * dpy;
            a2 = a2 + (w * g.y) * dpx;
            a3 = a3 + (w * g.y) * dpy;
        }
    }
    if (liquid) {
        const float pdet = det_comp(F);
        if (!finitef(pm) || pm <= 0.0f || !finitef(pvol) || pvol <= 0.0f || !finitef(pdet) || pdet <= 0.0f) {
            raise_at(ctl, ERR_G2P, i);
            return;
        }
        const float pd = st.density[i] / pdet;
        const float ld = pm / pvol;
        const float rel = (pd - ld) / pd;
        vx = vx + P.dt * (rel * P.gx);
        vy = vy + P.dt * (rel * P.gy);
        if (!finitef(pd) || pd <= 0.0f || !finitef(ld) || ld <= 0.0f || !finitef(rel) || !finitef(vx) ||
            !finitef(vy)) {
            raise_at(ctl, ERR_G2P, i);
            return;
        }
    }
    const float gs = P.grad_scale, dt = P.dt;
    const float4 C = make_float4(a0 * gs, a1 * gs, a2 * gs, a3 * gs);
    const float2 np = make_float2(p.x + dt * vx, p.y + dt * vy);
    if (!finitef(vx) || !finitef(vy) || !finite4(C) || !finite2(np)) { raise_flags(ctl, ERR_G2P); return; }
    const float rowsum = fmaxf(fabsf(C.x) + fabsf(C.y), fabsf(C.z) + fabsf(C.w));
    const float increment = dt * rowsum;
    const float required = ceilf(increment / MAX_DEFORMATION_INCREMENT);
    if (!finitef(required) || required > (float)MAX_DEFORMATION_SUBSTEPS) {
        raise_at(ctl, ERR_SUBSTEP_LIMIT, i);
        return;
    }
    uint32_t substeps = required < 1.0f ? 1u : (uint32_t)required;
    float4 trial = integrate_deformation(F, C, dt, substeps);
    if (!finite4(trial)) { raise_flags(ctl, ERR_G2P); return; }
    double tdet = det_f64(trial);
    // A near-singular F can lose orientation to f32 rounding; retry that update
    // with doubled substeps (shaders.ts:2274-2286).  The sign test is FP64.
    while (isfinite(tdet) && tdet <= 0.0 && substeps < MAX_DEFORMATION_SUBSTEPS) {
        substeps = min(substeps * 2u, MAX_DEFORMATION_SUBSTEPS);
        trial = integrate_deformation(F, C, dt, substeps);
        if (!finite4(trial)) { raise_flags(ctl, ERR_G2P); return; }
        tdet = det_f64(trial);
    }
    if (!isfinite(tdet) || tdet <= 0.0) {
        record_failure(ctl, 1u, i, (float)tdet, increment, (float)det_f64(F), 0.0f);
        raise_flags(ctl, ERR_DEFORMATION);
        return;
    }
    float4 Fn = trial;
    float pvn = in.plastic[i];
    if (liquid) {
        Fn = project_liquid(trial);
    } else if constexpr (!GENERAL) {
        raise_flags(ctl, ERR_G2P);
        return;
    } else if (model == MODEL_SNOW) {
        Decomp d;
        if (!decompose({trial.x, trial.y, trial.z, trial.w}, d)) { raise_flags(ctl, ERR_G2P); return; }
        const double lo = 1 - SNOW_COMPRESSION_YIELD, hi = 1 + SNOW_STRETCH_YIELD;
        const double s1 = fmax(lo, fmin(hi, d.s1)), s2 = fmax(lo, fmin(hi, d.s2));
        const double upv = fmax(SNOW_MIN_PLASTIC_VOLUME, fmin(SNOW_MAX_PLASTIC_VOLUME, ((double)pvn * tdet) / (s1 * s2)));
        const M2 r = reconstruct(d, s1, s2);
        Fn = make_float4((float)r.a, (float)r.b, (float)r.c, (float)r.d);
        pvn = (float)upv;
    } else if (model == MODEL_POWDER) {
        const float lambda = st.lambda[i], mu = st.mu[i], fric = st.param[i];
        if (!finitef(lambda) || !finitef(mu) || mu <= 0.0f || !finitef(fric) || fric < 0.0f) {
            raise_flags(ctl, ERR_G2P);
            return;
        }
        Decomp d;
        if (!decompose({trial.x, trial.y, trial.z, trial.w}, d)) { raise_flags(ctl, ERR_G2P); return; }
        const double l1 = log(d.s1), l2 = log(d.s2), lv = l1 + l2;
        const double minlv = log(MIN_POWDER_DETERMINANT);
        const double blv = fmax(minlv, lv);
        const double pressure = fmax(0.0, -((double)lambda + (double)mu) * blv);
        const double dev = 0.5 * (l1 - l2);
        const double maxdev = ((double)fric * pressure) / (2 * (double)mu);
        if (!(dev <= maxdev && lv >= minlv)) {
            const double pdev = fmin(dev, maxdev), mean = 0.5 * blv;
            const M2 r = reconstruct(d, exp(mean + pdev), exp(mean - pdev));
            Fn = make_float4((float)r.a, (float)r.b, (float)r.c, (float)r.d);
        }
    } else if (model == MODEL_NEO_HOOKEAN) {
        const float tr = trial.x * trial.x + trial.y * trial.y + trial.z * trial.z + trial.w * trial.w;
        if (!finitef(tr)) { raise_flags(ctl, ERR_G2P); return; }
        if (tdet < MIN_NEO_HOOKEAN_STRETCH * sqrt((double)tr) ||
            (double)tr > MAX_NEO_HOOKEAN_STRETCH * MAX_NEO_HOOKEAN_STRETCH) {
            Decomp d;
            if (!decompose({trial.x, trial.y, trial.z, trial.w}, d)) { raise_flags(ctl, ERR_G2P); return; }
            const double s1 = fmax(MIN_NEO_HOOKEAN_STRETCH, fmin(MAX_NEO_HOOKEAN_STRETCH, d.s1));
            const double s2 = fmax(MIN_NEO_HOOKEAN_STRETCH, fmin(MAX_NEO_HOOKEAN_STRETCH, d.s2));
            const M2 r = reconstruct(d, s1, s2);
            Fn = make_float4((float)r.a, (float)r.b, (float)r.c, (float)r.d);
        }
    } else {
        raise_flags(ctl, ERR_G2P);
        return;
    }
    const double ndet = det_f64(Fn);
    if (!finite4(Fn) || !finitef(pvn) || pvn < 0.6f || pvn > 20.0f) { raise_flags(ctl, ERR_G2P); return; }
    if (!isfinite(ndet) || ndet <= 0.0) {
        record_failure(ctl, 2u, i, Fn.x, Fn.y, Fn.z, Fn.w);
        raise_flags(ctl, ERR_PROJECTED_DEFORMATION);
        return;
    }
    // projectLiquidFromTriangles is the identity without bonded bodies; the
    // second confineAxis pass is then idempotent (shaders.ts:2376-2408).
    const float2 bxv = confine_axis(np.x, vx, P.dom_w);
    const float2 byv = confine_axis(np.y, vy, P.dom_h);
    const float2 fp = make_float2(bxv.x, byv.x), fv = make_float2(bxv.y, byv.y);
    out.pos[i] = fp;
    out.vel[i] = fv;
    out.C[i] = C;
    out.F[i] = Fn;
    out.plastic[i] = pvn;
    if (P.cfl_on >= 0.5f && P.body_count == 0u) {
        const float speed = sqrtf(fv.x * fv.x + fv.y * fv.y);
        const float ws = wave_speed_bound<GENERAL>(P, model, Fn, st.lambda[i], st.mu[i], st.density[i], st.volume[i], st.param[i]);
        if (!finite2(fv) || !finitef(speed) || !finitef(ws) || speed > P.vel_bound || ws > P.wave_bound) {
            record_failure(ctl, 3u, i, speed, ws, P.vel_bound, P.wave_bound);
            raise_flags(ctl, ERR_CFL_BOUND);
            return;
        }
    }
    if (bin_next) {
        // prepare + bin the committed candidate for the next microstep; its
        // failures are deferred so this microstep can still commit
        float4 A, aux;
        const uint32_t pf = prepare_particle<GENERAL>(P, model, C, Fn, st.lambda[i], st.mu[i], st.density[i],
                                                      st.volume[i], st.param[i], pvn, A, aux);
        if (pf) {
            bins.key[i] = INVALID_KEY;
            if (pf & ERR_P2G) atomicMin(&ctl->deferred_index, i);
            atomicOr(&ctl->deferred, pf);
            return;
        }
        bins.affine[i] = A;
        bins.aux[i] = aux;
        Stencil ns;
        if (!make_stencil(P, fp, ns)) {
            bins.key[i] = INVALID_KEY;
            raise_flags(ctl, ERR_STENCIL);
            return;
        }
        const uint32_t k = (uint32_t)(ns.by + 1) * W + (uint32_t)(ns.bx + 1);
        bins.key[i] = k;
        bins.rank[i] = atomicAdd(&bins.count[k], 1u);
    }
}

// ------------------------------------------------------------- batch end
__global__ void k_commit_last(Control* ctl) {
    if (threadIdx.x == 0 && blockIdx.x == 0 && ctl->status[0] == 0u) ctl->completed += 1u;
}

// On a failed microstep the committed state sits in the buffer of the last
// committed step; copy it to the batch's output buffer when their parity
// differs (commitParticles' whole-step rollback, shaders.ts:3067-3080).
__global__ __launch_bounds__(256) void k_fix_rollback(const Control* ctl, ParticleBuf b0, ParticleBuf b1, uint32_t n,
                                                      uint32_t start_buf, uint32_t steps) {
    if (ctl->status[0] == 0u) return;
    const uint32_t committed = (start_buf + ctl->completed) & 1u;
    const uint32_t final_buf = (start_buf + steps) & 1u;
    if (committed == final_buf) return;
    const ParticleBuf& src = committed ? b1 : b0;
    const ParticleBuf& dst = committed ? b0 : b1;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        dst.pos[i] = src.pos[i];
        dst.vel[i] = src.vel[i];
        dst.C[i] = src.C[i];
        dst.F[i] = src.F[i];
        dst.plastic[i] = src.plastic[i];
    }
}

// measureParticleSpeedPartials + reduceParticleSpeedPartials
// (shaders.ts:2463-2533) as one pass: wave max, then one atomic per block on
// the non-negative float's bit pattern.
__global__ __launch_bounds__(256) void k_measure_speed(Control* ctl, const float2* __restrict__ vel,
                                                       const uint32_t* __restrict__ model, uint32_t n) {
    __shared__ float s_w[4];
    float m = 0.0f;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (model[i] != MODEL_REMOVED) {
            const float2 v = vel[i];
            const float sp = sqrtf(v.x * v.x + v.y * v.y);
            m = fmaxf(m, finitef(sp) ? sp : FLT_MAX);
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

inline uint32_t blocks_for(uint32_t n, uint32_t t) { return (n + t - 1) / t; }

}  // namespace

size_t scan_block_count(uint32_t node_count) { return (node_count + SCAN_TILE - 1) / SCAN_TILE; }

#define MPM_LAUNCH(name, grid, block, ...)                                        \
    do {                                                                          \
        TimedRegion _tr(lc.timing, name, lc.stream);                              \
        hipLaunchKernelGGL(__VA_ARGS__);                                          \
        hipError_t _e = hipGetLastError();                                        \
        if (_e != hipSuccess) return _e;                                          \
    } while (0)

hipError_t launch_batch(const LaunchCtx& lc, int start_buf, uint32_t steps, bool measure) {
    const uint32_t n = lc.n, nodes = lc.node_count;
    const uint32_t pb = blocks_for(n > 0 ? n : 1, 256);
    const uint32_t nb = (uint32_t)scan_block_count(nodes);
    const uint32_t tiles_x = (lc.node_w + TI - 1) / TI, tiles_y = (lc.node_h + TI - 1) / TI;
    const uint32_t zb = min(blocks_for(nodes, 256), 2048u);
    MPM_LAUNCH("batch_begin", zb, 256, k_batch_begin, dim3(zb), dim3(256), 0, lc.stream, lc.ctl, lc.bins.count, nodes);
    const bool general = scene_general(lc.scene_flags);
    if (n > 0) {
        if (general)
            MPM_LAUNCH("bin", pb, 256, k_bin<true>, dim3(pb), dim3(256), 0, lc.stream, lc.params, lc.ctl,
                       lc.buf[start_buf], lc.st, lc.bins, n);
        else
            MPM_LAUNCH("bin", pb, 256, k_bin<false>, dim3(pb), dim3(256), 0, lc.stream, lc.params, lc.ctl,
                       lc.buf[start_buf], lc.st, lc.bins, n);
    }
    int cur = start_buf;
    for (uint32_t s = 0; s < steps; ++s) {
        const ParticleBuf& in = lc.buf[cur];
        const ParticleBuf& out = lc.buf[cur ^ 1];
        MPM_LAUNCH("scan_reduce", nb, SCAN_THREADS, k_scan_reduce, dim3(nb), dim3(SCAN_THREADS), 0, lc.stream, lc.ctl,
                   lc.bins.count, lc.bins.block_sums, nodes, s > 0 ? 1u : 0u);
        MPM_LAUNCH("scan_blocks", 1, SCAN_THREADS, k_scan_blocks, dim3(1), dim3(SCAN_THREADS), 0, lc.stream, lc.ctl,
                   lc.bins.block_sums, nb);
        MPM_LAUNCH("scan_final", nb, SCAN_THREADS, k_scan_final, dim3(nb), dim3(SCAN_THREADS), 0, lc.stream, lc.ctl,
                   lc.bins.count, lc.bins.block_sums, lc.bins.start, nodes);
        if (n > 0)
            MPM_LAUNCH("scatter", pb, 256, k_scatter, dim3(pb), dim3(256), 0, lc.stream, lc.ctl, lc.bins.key,
                       lc.bins.rank, lc.bins.start, lc.bins.perm, n);
        if (lc.fixed_point)
            MPM_LAUNCH("p2g_grid_fixed", tiles_x * tiles_y, P2G_THREADS, k_p2g_grid<true>, dim3(tiles_x * tiles_y),
                       dim3(P2G_THREADS), 0, lc.stream, lc.params, lc.ctl, in, lc.st, lc.bins, lc.grid, tiles_x,
                       lc.keep_fixed, lc.scene_flags);
        else
            MPM_LAUNCH("p2g_grid", tiles_x * tiles_y, P2G_THREADS, k_p2g_grid<false>, dim3(tiles_x * tiles_y),
                       dim3(P2G_THREADS), 0, lc.stream, lc.params, lc.ctl, in, lc.st, lc.bins, lc.grid, tiles_x, 0u,
                       lc.scene_flags);
        if (n > 0) {
            if (general)
                MPM_LAUNCH("g2p", pb, 256, k_g2p<true>, dim3(pb), dim3(256), 0, lc.stream, lc.params, lc.ctl, in, out,
                           lc.st, lc.grid, lc.bins, n, s + 1 < steps ? 1u : 0u);
            else
                MPM_LAUNCH("g2p", pb, 256, k_g2p<false>, dim3(pb), dim3(256), 0, lc.stream, lc.params, lc.ctl, in, out,
                           lc.st, lc.grid, lc.bins, n, s + 1 < steps ? 1u : 0u);
        }
        cur ^= 1;
    }
    if (steps > 0) MPM_LAUNCH("commit_last", 1, 64, k_commit_last, dim3(1), dim3(64), 0, lc.stream, lc.ctl);
    if (n > 0) {
        const uint32_t fb = min(pb, 1024u);
        MPM_LAUNCH("fix_rollback", fb, 256, k_fix_rollback, dim3(fb), dim3(256), 0, lc.stream, lc.ctl, lc.buf[0],
                   lc.buf[1], n, (uint32_t)start_buf, steps);
        if (measure) {
            const uint32_t mb = min(pb, 512u);
            MPM_LAUNCH("measure_speed", mb, 256, k_measure_speed, dim3(mb), dim3(256), 0, lc.stream, lc.ctl,
                       lc.buf[cur].vel, lc.st.model, n);
        }
    }
    return hipSuccess;
}

}  // namespace mpm
