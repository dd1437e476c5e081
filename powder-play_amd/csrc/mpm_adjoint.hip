// Differentiable rollout of the 2D MLS-MPM step for a fixed-corotated elastic
// solid on gfx950: a float32 forward with FP64 polar/stress, segment
// checkpoints, and a hand-written reverse pass that recomputes one segment at
// a time.
//
// Forward step (the solid branch of stepMpm, src/lib/mpm/reference/solver.ts:
// 396-824): quadratic B-spline stencil (:2017-2034), stress
// P F^T = 2 mu (F - R) F^T + lambda J (J - 1) I (snow branch of
// constitutiveStressTimesFt with hardening factor 1, :2171-2203), APIC P2G
// (:418-479), grid velocity with wall cushion and tangential friction
// (:1507-1524, :2542-2559), G2P (:610-773) without plastic projection,
// F <- (I + h C)^s F with the reference's substep count (:2561-2592), and
// confineAxis (:2528-2540).  The reference has no differentiation code; the
// reverse pass is this library's own.
//
// Layout in HBM (n particles, T steps, segment length L, N grid nodes):
//   state level    3 float4 per particle: (x, y, vx, vy), C, F       48 B
//   checkpoints    ceil(T / L) levels + the final state
//   segment        L levels of states + L grid levels (v float2, (m, vh) float4)
//   acc            float4 per node (m, qx, qy, -), float atomics, cleared by
//                  the node pass that consumes it
//   adjacc         float2 per node (dL/dv), float atomics, cleared likewise
//   qm             float4 per node (dL/dq, dL/dm)
//   node boxes     int4 (inclusive node range) of each step's P2G; the node
//                  passes visit only that box
//   partial        2 float4 per particle: dL/dx and dL/dF carried from the
//                  G2P reverse to the P2G reverse of the same step
//   grad lambda/mu double per particle, accumulated over all steps
//
// Per step the forward is two launches — nodes (grid update) and particles
// (G2P of this step fused with P2G of the next) — and so is the reverse:
// particles (P2G reverse of step t+1 fused with G2P reverse of step t, one
// thread per particle carries the adjoint between them in registers) and
// nodes (grid-update reverse).  Both kernel sequences are captured once into
// HIP graphs and replayed; the status reset, the gradient clears and the
// host copies stay outside the graphs.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <string>

#include "mpm355.h"

namespace {

constexpr int PT = 64;   // particles per workgroup: one wave, so the LDS scatter of a block spreads over many CUs
constexpr int NT = 256;  // nodes per workgroup
constexpr float CUSHION_CELLS = 8.0f;     // WALL_CUSHION_CELLS, solver.ts:266-283
constexpr float MAX_INCREMENT = 0.25f;    // MAX_DEFORMATION_INCREMENT
constexpr int MAX_SUBSTEPS = 64;          // MAX_DEFORMATION_SUBSTEPS

struct DiffConst {
    float dt, dx, gs, pm, ss, gdx, gdy, friction;
    float ext_x, ext_y, cush_x, cush_y, wd, hd;
    int cells_x, cells_y, W, nodes;
    uint32_t n;
    uint32_t* status;
};

struct Stencil {
    int bx, by;
    float fx, fy;
    float wx[3], wy[3], gx[3], gy[3];
};

__device__ inline bool make_stencil(const DiffConst& P, float x, float y, Stencil& s) {
    const float gx = x / P.dx, gy = y / P.dx;
    const float bxf = floorf(gx - 0.5f), byf = floorf(gy - 0.5f);
    if (!(bxf >= -1.0f && byf >= -1.0f && bxf <= (float)(P.cells_x - 1) && byf <= (float)(P.cells_y - 1)))
        return false;
    s.bx = (int)bxf;
    s.by = (int)byf;
    const float fx = gx - bxf, fy = gy - byf;
    s.fx = fx;
    s.fy = fy;
    s.wx[0] = 0.5f * (1.5f - fx) * (1.5f - fx);
    s.wx[1] = 0.75f - (fx - 1.0f) * (fx - 1.0f);
    s.wx[2] = 0.5f * (fx - 0.5f) * (fx - 0.5f);
    s.wy[0] = 0.5f * (1.5f - fy) * (1.5f - fy);
    s.wy[1] = 0.75f - (fy - 1.0f) * (fy - 1.0f);
    s.wy[2] = 0.5f * (fy - 0.5f) * (fy - 0.5f);
    s.gx[0] = (fx - 1.5f) / P.dx;
    s.gx[1] = -2.0f * (fx - 1.0f) / P.dx;
    s.gx[2] = (fx - 0.5f) / P.dx;
    s.gy[0] = (fy - 1.5f) / P.dx;
    s.gy[1] = -2.0f * (fy - 1.0f) / P.dx;
    s.gy[2] = (fy - 0.5f) / P.dx;
    return true;
}

__device__ inline int node_key(const DiffConst& P, const Stencil& s, int ox, int oy) {
    return (s.by + oy + 1) * P.W + (s.bx + ox + 1);
}

// Fixed-corotated P F^T in FP64: R is the polar rotation of F through the
// angle atan2(F10 - F01, F00 + F11).  Returns false for det F <= 0.
__device__ inline bool corot(float4 Ff, float lamf, float muf, float* K) {
    const double a = Ff.x, b = Ff.y, c = Ff.z, d = Ff.w;
    const double J = a * d - b * c;
    const double x = a + d, y = c - b, r = sqrt(x * x + y * y);
    if (!(J > 0.0) || !(r > 0.0)) {
        K[0] = K[1] = K[2] = K[3] = 0.0f;
        return false;
    }
    const double co = x / r, si = y / r;
    const double m0 = a - co, m1 = b + si, m2 = c - si, m3 = d - co;
    const double s2 = 2.0 * (double)muf, vol = (double)lamf * J * (J - 1.0);
    K[0] = (float)(s2 * (m0 * a + m1 * b) + vol);
    K[1] = (float)(s2 * (m0 * c + m1 * d));
    K[2] = (float)(s2 * (m2 * a + m3 * b));
    K[3] = (float)(s2 * (m2 * c + m3 * d) + vol);
    return true;
}

// Reverse of corot in FP64: Fb += dL/dF, lamb/mub += dL/dlambda, dL/dmu.
__device__ inline void corot_bwd(float4 Ff, float lamf, float muf, const float* Kbf, double* Fb, double& lamb,
                                 double& mub) {
    const double a = Ff.x, b = Ff.y, c = Ff.z, d = Ff.w;
    const double J = a * d - b * c;
    const double x = a + d, y = c - b, r2 = x * x + y * y, r = sqrt(r2);
    if (!(J > 0.0) || !(r > 0.0)) return;
    const double co = x / r, si = y / r;
    const double m0 = a - co, m1 = b + si, m2 = c - si, m3 = d - co;
    const double s2 = 2.0 * (double)muf;
    const double k0 = Kbf[0], k1 = Kbf[1], k2 = Kbf[2], k3 = Kbf[3];
    mub += 2.0 * (k0 * (m0 * a + m1 * b) + k1 * (m0 * c + m1 * d) + k2 * (m2 * a + m3 * b) + k3 * (m2 * c + m3 * d));
    const double tr = k0 + k3;
    lamb += J * (J - 1.0) * tr;
    const double mb0 = s2 * (k0 * a + k1 * c), mb1 = s2 * (k0 * b + k1 * d);
    const double mb2 = s2 * (k2 * a + k3 * c), mb3 = s2 * (k2 * b + k3 * d);
    Fb[0] += s2 * (k0 * m0 + k2 * m2) + mb0;
    Fb[1] += s2 * (k0 * m1 + k2 * m3) + mb1;
    Fb[2] += s2 * (k1 * m0 + k3 * m2) + mb2;
    Fb[3] += s2 * (k1 * m1 + k3 * m3) + mb3;
    const double jv = (double)lamf * (2.0 * J - 1.0) * tr;
    Fb[0] += jv * d;
    Fb[1] -= jv * c;
    Fb[2] -= jv * b;
    Fb[3] += jv * a;
    const double thb = mb0 * si + mb1 * co - mb2 * co + mb3 * si;
    const double xb = -thb * y / r2, yb = thb * x / r2;
    Fb[0] += xb;
    Fb[3] += xb;
    Fb[2] += yb;
    Fb[1] -= yb;
}

// applyWallCushion factor (solver.ts:2542-2551), chosen by the sign of v.
__device__ inline float cushion_factor(float v, float coord, float extent, float cushion) {
    const float dist = v < 0.0f ? coord : extent - coord;
    return fmaxf(0.0f, fminf(1.0f, dist / cushion));
}
// applyWallTangentialFriction factor (solver.ts:2552-2559).
__device__ inline float friction_factor(float coord, float extent, float cushion, float friction) {
    if (friction == 0.0f) return 1.0f;
    const float dist = fmaxf(0.0f, fminf(coord, extent - coord));
    return 1.0f - friction * (1.0f - fmaxf(0.0f, fminf(1.0f, dist / cushion)));
}

// Substep count of updateDeformationGradient; 0 when it exceeds the limit.
__device__ inline int substeps(const float* C, float dt) {
    const float row = fmaxf(fabsf(C[0]) + fabsf(C[1]), fabsf(C[2]) + fabsf(C[3]));
    const float inc = dt * row;
    if (!isfinite(inc)) return 0;
    const float s = fmaxf(1.0f, ceilf(inc / MAX_INCREMENT));
    return s > (float)MAX_SUBSTEPS ? 0 : (int)s;
}

// confineAxis (solver.ts:2528-2540): true when the axis is clamped.
__device__ inline bool clamped(float x, float v, float maximum) {
    return x < 0.0f || (x == 0.0f && v < 0.0f) || x > maximum || (x == maximum && v > 0.0f);
}

__device__ inline void mat_mul(const float* a, const float* b, float* r) {
    const float r0 = a[0] * b[0] + a[1] * b[2], r1 = a[0] * b[1] + a[1] * b[3];
    const float r2 = a[2] * b[0] + a[3] * b[2], r3 = a[2] * b[1] + a[3] * b[3];
    r[0] = r0;
    r[1] = r1;
    r[2] = r2;
    r[3] = r3;
}

// G2P gather of one particle: velocity and C = (4 / dx^2) sum w v dp^T.
__device__ inline void gather(const DiffConst& P, const Stencil& s, const float2* __restrict__ nodeV, float* v,
                              float* C) {
    float vx = 0.0f, vy = 0.0f, c0 = 0.0f, c1 = 0.0f, c2 = 0.0f, c3 = 0.0f;
#pragma unroll
    for (int oy = 0; oy < 3; ++oy)
#pragma unroll
        for (int ox = 0; ox < 3; ++ox) {
            const float w = s.wx[ox] * s.wy[oy];
            const float2 nv = nodeV[node_key(P, s, ox, oy)];
            const float dpx = ((float)ox - s.fx) * P.dx, dpy = ((float)oy - s.fy) * P.dx;
            vx += w * nv.x;
            vy += w * nv.y;
            c0 += w * nv.x * dpx;
            c1 += w * nv.x * dpy;
            c2 += w * nv.y * dpx;
            c3 += w * nv.y * dpy;
        }
    v[0] = vx;
    v[1] = vy;
    C[0] = c0 * P.gs;
    C[1] = c1 * P.gs;
    C[2] = c2 * P.gs;
    C[3] = c3 * P.gs;
}

// ---- block-local pre-reduction of the node scatters ----------------------
//
// Device-scope float atomics are the cost of a naive scatter (each one a
// round trip past the L2).  A workgroup first sums its particles'
// contributions into an LDS tile spanning the bounding box of their stencils
// and then adds each touched node once to HBM.  Particles of one workgroup
// are neighbours in the state order; a workgroup whose box exceeds the tile
// falls back to direct atomics, which is correct for any order.

constexpr int EMPTY_LO = 0x3fffffff, EMPTY_HI = -0x3fffffff;

struct ScatterTile {
    int x0, y0, tw, th;  // box of stencil bases; tile is (tw + 2) x (th + 2) nodes
    bool tiled, empty;
};

__device__ inline int wave_min(int v) {
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ inline int wave_max(int v) {
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// Block-wide box of the stencil bases of valid particles (all threads call).
__device__ inline ScatterTile block_tile(bool ok, int bx, int by, int (*red)[4], int cap_nodes) {
    int v[4] = {ok ? bx : EMPTY_LO, ok ? by : EMPTY_LO, ok ? bx : EMPTY_HI, ok ? by : EMPTY_HI};
    v[0] = wave_min(v[0]);
    v[1] = wave_min(v[1]);
    v[2] = wave_max(v[2]);
    v[3] = wave_max(v[3]);
    const int wave = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
        for (int k = 0; k < 4; ++k) red[wave][k] = v[k];
    __syncthreads();
    int x0 = red[0][0], y0 = red[0][1], x1 = red[0][2], y1 = red[0][3];
    for (int w = 1; w < PT / 64; ++w) {
        x0 = min(x0, red[w][0]);
        y0 = min(y0, red[w][1]);
        x1 = max(x1, red[w][2]);
        y1 = max(y1, red[w][3]);
    }
    ScatterTile t;
    t.empty = x0 > x1;
    t.x0 = x0;
    t.y0 = y0;
    t.tw = t.empty ? 0 : x1 - x0 + 3;
    t.th = t.empty ? 0 : y1 - y0 + 3;
    t.tiled = !t.empty && t.tw * t.th <= cap_nodes;
    return t;
}

template <int NC>
__device__ inline void tile_clear(const ScatterTile& t, float* tile) {
    if (!t.tiled) return;
    for (int i = threadIdx.x; i < NC * t.tw * t.th; i += PT) tile[i] = 0.0f;
}

// Adds NC values of one stencil node: into the LDS tile or straight to HBM.
template <int NC>
__device__ inline void tile_add(const ScatterTile& t, float* tile, float* global_node, int bx, int by, int ox,
                                int oy, const float* val) {
    if (t.tiled) {
        const int i = (by - t.y0 + oy) * t.tw + (bx - t.x0 + ox), plane = t.tw * t.th;
#pragma unroll
        for (int c = 0; c < NC; ++c) atomicAdd(&tile[c * plane + i], val[c]);
    } else {
#pragma unroll
        for (int c = 0; c < NC; ++c) unsafeAtomicAdd(global_node + c, val[c]);
    }
}

// Adds the tile to HBM, one atomic per touched node and channel.  Node i of
// the tile is grid node (y0 + i / tw + 1) * W + (x0 + i % tw + 1).
template <int NC, int STRIDE>
__device__ inline void tile_flush(const DiffConst& P, const ScatterTile& t, const float* tile, float* global) {
    if (!t.tiled) return;
    const int plane = t.tw * t.th;
    for (int i = threadIdx.x; i < plane; i += PT) {
        float v[NC];
        bool any = false;
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            v[c] = tile[c * plane + i];
            any |= v[c] != 0.0f;
        }
        if (!any) continue;
        const int key = (t.y0 + i / t.tw + 1) * P.W + (t.x0 + i % t.tw + 1);
#pragma unroll
        for (int c = 0; c < NC; ++c)
            if (v[c] != 0.0f) unsafeAtomicAdd(global + (size_t)key * STRIDE + c, v[c]);
    }
}

// Widens the step's node box (inclusive node coordinates) by this block's.
__device__ inline void box_update(const ScatterTile& t, int4* box) {
    if (threadIdx.x != 0 || t.empty) return;
    atomicMin(&box->x, t.x0 + 1);
    atomicMin(&box->y, t.y0 + 1);
    atomicMax(&box->z, t.x0 + t.tw);
    atomicMax(&box->w, t.y0 + t.th);
}

constexpr int TILE_NODES = 2048;  // LDS tile capacity in nodes (3 channels: 24 KB)

// Forward particle pass: G2P of step t (src -> dst, may alias) and/or P2G of
// the resulting state into acc, whose node box is widened in `box`.
template <bool G2P, bool P2G>
__global__ __launch_bounds__(PT) void k_diff_particles(DiffConst P, const float4* src, float4* dst,
                                                       const float2* __restrict__ nodeV, float4* __restrict__ acc,
                                                       const float* __restrict__ lam, const float* __restrict__ mu,
                                                       int4* __restrict__ box) {
    __shared__ float tile[3 * TILE_NODES];
    __shared__ int red[PT / 64][4];
    const uint32_t p = blockIdx.x * PT + threadIdx.x;
    const bool live = p < P.n;
    float4 xv = make_float4(0.0f, 0.0f, 0.0f, 0.0f), C = xv, F = xv;
    if (live) {
        xv = src[3 * p];
        C = src[3 * p + 1];
        F = src[3 * p + 2];
    }
    uint32_t err = 0;
    if (G2P && live) {
        Stencil s;
        if (!make_stencil(P, xv.x, xv.y, s)) {
            err |= MPM_DIFF_STENCIL;
        } else {
            float v[2], Cn[4];
            gather(P, s, nodeV, v, Cn);
            const int ns = substeps(Cn, P.dt);
            if (ns == 0 || !isfinite(v[0]) || !isfinite(v[1])) {
                err |= ns == 0 ? MPM_DIFF_SUBSTEPS : MPM_DIFF_NONFINITE;
            } else {
                const float h = P.dt / (float)ns;
                const float U[4] = {1.0f + h * Cn[0], h * Cn[1], h * Cn[2], 1.0f + h * Cn[3]};
                float Fm[4] = {F.x, F.y, F.z, F.w};
                for (int k = 0; k < ns; ++k) mat_mul(U, Fm, Fm);
                if (!(Fm[0] * Fm[3] - Fm[1] * Fm[2] > 0.0f)) err |= MPM_DIFF_DEFORMATION;
                const float nx = xv.x + P.dt * v[0], ny = xv.y + P.dt * v[1];
                const bool cx = clamped(nx, v[0], P.wd), cy = clamped(ny, v[1], P.hd);
                if (cx || cy) err |= MPM_DIFF_CONFINED;
                xv.x = cx ? ((nx < 0.0f || (nx == 0.0f && v[0] < 0.0f)) ? 0.0f : P.wd) : nx;
                xv.y = cy ? ((ny < 0.0f || (ny == 0.0f && v[1] < 0.0f)) ? 0.0f : P.hd) : ny;
                xv.z = cx ? 0.0f : v[0];
                xv.w = cy ? 0.0f : v[1];
                C = make_float4(Cn[0], Cn[1], Cn[2], Cn[3]);
                F = make_float4(Fm[0], Fm[1], Fm[2], Fm[3]);
            }
        }
        dst[3 * p] = xv;
        dst[3 * p + 1] = C;
        dst[3 * p + 2] = F;
    }
    if (P2G) {
        Stencil s;
        const bool ok = live && make_stencil(P, xv.x, xv.y, s);
        if (live && !ok) err |= MPM_DIFF_STENCIL;
        const ScatterTile t = block_tile(ok, ok ? s.bx : 0, ok ? s.by : 0, red, TILE_NODES);
        tile_clear<3>(t, tile);
        if (t.tiled) __syncthreads();
        if (ok) {
            float K[4];
            if (!corot(F, lam[p], mu[p], K)) err |= MPM_DIFF_DEFORMATION;
            const float A0 = C.x * P.pm + K[0] * P.ss, A1 = C.y * P.pm + K[1] * P.ss;
            const float A2 = C.z * P.pm + K[2] * P.ss, A3 = C.w * P.pm + K[3] * P.ss;
            const float mvx = P.pm * xv.z, mvy = P.pm * xv.w;
#pragma unroll
            for (int oy = 0; oy < 3; ++oy)
#pragma unroll
                for (int ox = 0; ox < 3; ++ox) {
                    const float w = s.wx[ox] * s.wy[oy];
                    const float dpx = ((float)ox - s.fx) * P.dx, dpy = ((float)oy - s.fy) * P.dx;
                    const float val[3] = {w * P.pm, w * (mvx + A0 * dpx + A1 * dpy), w * (mvy + A2 * dpx + A3 * dpy)};
                    tile_add<3>(t, tile, reinterpret_cast<float*>(acc + node_key(P, s, ox, oy)), s.bx, s.by, ox, oy,
                                val);
                }
        }
        if (t.tiled) __syncthreads();
        tile_flush<3, 4>(P, t, tile, reinterpret_cast<float*>(acc));
        box_update(t, box);
    }
    if (err) atomicOr(P.status, err);
}

constexpr int NODE_BLOCKS = 16;  // node passes loop over the step's box

// Grid node of the i-th position of a box (row-major over the box).
__device__ inline int box_node(const int4& b, int i, const DiffConst& P) {
    const int bw = b.z - b.x + 1;
    return (b.y + i / bw) * P.W + (b.x + i % bw);
}

// Forward node pass over the box of the step's P2G: consumes and clears acc,
// writes the node velocity after the walls (nodeV) and, when recording a
// segment, (m, vh) for the reverse (nodeB) and the box itself (box_save).
// Resets the other box for the next P2G.
__global__ __launch_bounds__(NT) void k_diff_nodes(DiffConst P, float4* __restrict__ acc, float2* __restrict__ nodeV,
                                                   float4* __restrict__ nodeB, const int4* __restrict__ box_in,
                                                   int4* __restrict__ box_reset, int4* __restrict__ box_save) {
    const int4 b = *box_in;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *box_reset = make_int4(EMPTY_LO, EMPTY_LO, EMPTY_HI, EMPTY_HI);
        if (box_save) *box_save = b;
    }
    if (b.x > b.z) return;
    const int count = (b.z - b.x + 1) * (b.w - b.y + 1);
    for (int i = blockIdx.x * NT + threadIdx.x; i < count; i += gridDim.x * NT) {
        const int k = box_node(b, i, P);
        const float4 a = acc[k];
        acc[k] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        const float xi = (float)(k % P.W - 1), yi = (float)(k / P.W - 1);
        float vhx = 0.0f, vhy = 0.0f;
        if (a.x > 0.0f) {
            vhx = a.y / a.x + P.gdx;
            vhy = a.z / a.x + P.gdy;
        }
        const float vx = (vhx * cushion_factor(vhx, xi, P.ext_x, P.cush_x)) *
                         friction_factor(yi, P.ext_y, P.cush_y, P.friction);
        const float vy = (vhy * cushion_factor(vhy, yi, P.ext_y, P.cush_y)) *
                         friction_factor(xi, P.ext_x, P.cush_x, P.friction);
        nodeV[k] = make_float2(vx, vy);
        if (nodeB) nodeB[k] = make_float4(a.x, vhx, vhy, 0.0f);
        if (!isfinite(vx) || !isfinite(vy)) atomicOr(P.status, (uint32_t)MPM_DIFF_NONFINITE);
    }
}

// Reverse of the grid update of one step, over that step's box: consumes and
// clears adjacc (dL/dv), writes (dL/dq, dL/dm) per node.
__global__ __launch_bounds__(NT) void k_diff_nodes_bwd(DiffConst P, float2* __restrict__ adjacc,
                                                       const float4* __restrict__ nodeB, float4* __restrict__ qm,
                                                       const int4* __restrict__ box) {
    const int4 b = *box;
    if (b.x > b.z) return;
    const int count = (b.z - b.x + 1) * (b.w - b.y + 1);
    for (int i = blockIdx.x * NT + threadIdx.x; i < count; i += gridDim.x * NT) {
        const int k = box_node(b, i, P);
        const float2 vb = adjacc[k];
        adjacc[k] = make_float2(0.0f, 0.0f);
        const float4 nb = nodeB[k];
        float4 out = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (nb.x > 0.0f) {
            const float xi = (float)(k % P.W - 1), yi = (float)(k / P.W - 1);
            const float ax = cushion_factor(nb.y, xi, P.ext_x, P.cush_x) * friction_factor(yi, P.ext_y, P.cush_y, P.friction);
            const float ay = cushion_factor(nb.z, yi, P.ext_y, P.cush_y) * friction_factor(xi, P.ext_x, P.cush_x, P.friction);
            const float hbx = ax * vb.x, hby = ay * vb.y;
            out.x = hbx / nb.x;
            out.y = hby / nb.x;
            out.z = -(hbx * (nb.y - P.gdx) + hby * (nb.z - P.gdy)) / nb.x;
        }
        qm[k] = out;
    }
}

__global__ void k_diff_box_reset(int4* __restrict__ boxes) {
    if (threadIdx.x < 2) boxes[threadIdx.x] = make_int4(EMPTY_LO, EMPTY_LO, EMPTY_HI, EMPTY_HI);
}

// Reverse of F' = U^s F (U = I + h C'): Fb = (U^T)^s G and
// Ub = sum_k (U^T)^k G F^T (U^T)^(s-1-k), for G = dL/dF'.
__device__ inline void substep_bwd(const float* U, float4 F, int ns, const float* G, float* Fb, float* Ub) {
    const float Ut[4] = {U[0], U[2], U[1], U[3]};
    float Q[4] = {F.x, F.z, F.y, F.w};
    float T[4];
    mat_mul(G, Q, T);
    for (int j = 1; j < ns; ++j) {
        float GQ[4];
        mat_mul(Q, Ut, Q);
        mat_mul(Ut, T, T);
        mat_mul(G, Q, GQ);
        T[0] += GQ[0];
        T[1] += GQ[1];
        T[2] += GQ[2];
        T[3] += GQ[3];
    }
    for (int k = 0; k < 4; ++k) {
        Ub[k] = T[k];
        Fb[k] = G[k];
    }
    for (int j = 0; j < ns; ++j) mat_mul(Ut, Fb, Fb);
}

// Reverse particle pass.  B3: reverse of P2G + stress of step t+1 (state
// s_next, node adjoints qm, carried partials) -> dL/d(state t+1) in
// registers; without B3 that adjoint is read from bar_in.  B1: reverse of
// G2P of step t (state s_cur, node velocities nodeV) -> scatter dL/dv into
// adjacc, write the partials of state t; without B1 the adjoint of state t+1
// is written to bar_out.
template <bool B3, bool B1>
__global__ __launch_bounds__(PT) void k_diff_particles_bwd(
    DiffConst P, const float4* __restrict__ s_next, const float4* __restrict__ qm, const float4* __restrict__ s_cur,
    const float2* __restrict__ nodeV, float2* __restrict__ adjacc, float4* __restrict__ partial,
    const float4* __restrict__ bar_in, float4* __restrict__ bar_out, const float* __restrict__ lam,
    const float* __restrict__ mu, double* __restrict__ glam, double* __restrict__ gmu) {
    __shared__ float tile[2 * TILE_NODES];
    __shared__ int red[PT / 64][4];
    const uint32_t p = blockIdx.x * PT + threadIdx.x;
    const bool live = p < P.n;
    float ax = 0.0f, ay = 0.0f, avx = 0.0f, avy = 0.0f;
    float aC[4] = {0.0f, 0.0f, 0.0f, 0.0f}, aF[4] = {0.0f, 0.0f, 0.0f, 0.0f};
    uint32_t err = 0;
    if (B3 && live) {
        const float4 xv = s_next[3 * p], C = s_next[3 * p + 1], F = s_next[3 * p + 2];
        const float4 px = partial[2 * p], pF = partial[2 * p + 1];
        ax = px.x;
        ay = px.y;
        double Fb[4] = {pF.x, pF.y, pF.z, pF.w};
        float Ab[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        Stencil s;
        float K[4];
        if (!make_stencil(P, xv.x, xv.y, s)) {
            err |= MPM_DIFF_STENCIL;
        } else {
            corot(F, lam[p], mu[p], K);
            const float A0 = C.x * P.pm + K[0] * P.ss, A1 = C.y * P.pm + K[1] * P.ss;
            const float A2 = C.z * P.pm + K[2] * P.ss, A3 = C.w * P.pm + K[3] * P.ss;
            const float mvx = P.pm * xv.z, mvy = P.pm * xv.w;
#pragma unroll
            for (int oy = 0; oy < 3; ++oy)
#pragma unroll
                for (int ox = 0; ox < 3; ++ox) {
                    const float w = s.wx[ox] * s.wy[oy];
                    const float dpx = ((float)ox - s.fx) * P.dx, dpy = ((float)oy - s.fy) * P.dx;
                    const float4 g = qm[node_key(P, s, ox, oy)];
                    const float momx = mvx + A0 * dpx + A1 * dpy, momy = mvy + A2 * dpx + A3 * dpy;
                    const float wb = P.pm * g.z + g.x * momx + g.y * momy;
                    avx += w * P.pm * g.x;
                    avy += w * P.pm * g.y;
                    Ab[0] += w * g.x * dpx;
                    Ab[1] += w * g.x * dpy;
                    Ab[2] += w * g.y * dpx;
                    Ab[3] += w * g.y * dpy;
                    ax += wb * s.gx[ox] * s.wy[oy] - w * (A0 * g.x + A2 * g.y);
                    ay += wb * s.wx[ox] * s.gy[oy] - w * (A1 * g.x + A3 * g.y);
                }
            float Kb[4];
            for (int k = 0; k < 4; ++k) Kb[k] = P.ss * Ab[k];
            double lb = 0.0, mb = 0.0;
            corot_bwd(F, lam[p], mu[p], Kb, Fb, lb, mb);
            glam[p] += lb;
            gmu[p] += mb;
        }
        for (int k = 0; k < 4; ++k) {
            aC[k] = P.pm * Ab[k];
            aF[k] = (float)Fb[k];
        }
    } else if (!B3 && live) {
        const float4 b0 = bar_in[3 * p], b1 = bar_in[3 * p + 1], b2 = bar_in[3 * p + 2];
        ax = b0.x;
        ay = b0.y;
        avx = b0.z;
        avy = b0.w;
        aC[0] = b1.x; aC[1] = b1.y; aC[2] = b1.z; aC[3] = b1.w;
        aF[0] = b2.x; aF[1] = b2.y; aF[2] = b2.z; aF[3] = b2.w;
    }
    if (B1) {
        float4 xv = make_float4(0.0f, 0.0f, 0.0f, 0.0f), F = xv;
        if (live) {
            xv = s_cur[3 * p];
            F = s_cur[3 * p + 2];
        }
        float xb[2] = {0.0f, 0.0f}, Fb[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        Stencil s;
        const bool ok = live && make_stencil(P, xv.x, xv.y, s);
        if (live && !ok) err |= MPM_DIFF_STENCIL;
        const ScatterTile t = block_tile(ok, ok ? s.bx : 0, ok ? s.by : 0, red, TILE_NODES);
        tile_clear<2>(t, tile);
        if (t.tiled) __syncthreads();
        if (ok) {
            float v[2], Cn[4];
            gather(P, s, nodeV, v, Cn);
            int ns = substeps(Cn, P.dt);
            if (ns == 0) {
                err |= MPM_DIFF_SUBSTEPS;
                ns = 1;
            }
            const float h = P.dt / (float)ns;
            const float U[4] = {1.0f + h * Cn[0], h * Cn[1], h * Cn[2], 1.0f + h * Cn[3]};
            const float nx = xv.x + P.dt * v[0], ny = xv.y + P.dt * v[1];
            const bool cx = clamped(nx, v[0], P.wd), cy = clamped(ny, v[1], P.hd);
            const float vgx = cx ? 0.0f : avx + P.dt * ax, vgy = cy ? 0.0f : avy + P.dt * ay;
            xb[0] = cx ? 0.0f : ax;
            xb[1] = cy ? 0.0f : ay;
            float Ub[4];
            substep_bwd(U, F, ns, aF, Fb, Ub);
            const float S0 = P.gs * (aC[0] + h * Ub[0]), S1 = P.gs * (aC[1] + h * Ub[1]);
            const float S2 = P.gs * (aC[2] + h * Ub[2]), S3 = P.gs * (aC[3] + h * Ub[3]);
#pragma unroll
            for (int oy = 0; oy < 3; ++oy)
#pragma unroll
                for (int ox = 0; ox < 3; ++ox) {
                    const float w = s.wx[ox] * s.wy[oy];
                    const float dpx = ((float)ox - s.fx) * P.dx, dpy = ((float)oy - s.fy) * P.dx;
                    const int key = node_key(P, s, ox, oy);
                    const float2 nv = nodeV[key];
                    const float ux = S0 * dpx + S1 * dpy, uy = S2 * dpx + S3 * dpy;
                    const float val[2] = {w * (vgx + ux), w * (vgy + uy)};
                    tile_add<2>(t, tile, reinterpret_cast<float*>(adjacc + key), s.bx, s.by, ox, oy, val);
                    const float wb = nv.x * (vgx + ux) + nv.y * (vgy + uy);
                    xb[0] += wb * s.gx[ox] * s.wy[oy] - w * (S0 * nv.x + S2 * nv.y);
                    xb[1] += wb * s.wx[ox] * s.gy[oy] - w * (S1 * nv.x + S3 * nv.y);
                }
        }
        if (t.tiled) __syncthreads();
        tile_flush<2, 2>(P, t, tile, reinterpret_cast<float*>(adjacc));
        if (live) {
            partial[2 * p] = make_float4(xb[0], xb[1], 0.0f, 0.0f);
            partial[2 * p + 1] = make_float4(Fb[0], Fb[1], Fb[2], Fb[3]);
        }
    } else if (live) {
        bar_out[3 * p] = make_float4(ax, ay, avx, avy);
        bar_out[3 * p + 1] = make_float4(aC[0], aC[1], aC[2], aC[3]);
        bar_out[3 * p + 2] = make_float4(aF[0], aF[1], aF[2], aF[3]);
    }
    if (err) atomicOr(P.status, err);
}

}  // namespace

struct mpm_diff {
    int device = 0;
    hipStream_t stream = nullptr;
    mpm_diff_config cfg{};
    DiffConst P{};
    uint32_t steps = 0, seg = 0, nseg = 0;
    size_t level_f4 = 0;  // float4 per state level (3 n)
    float4* ckpt = nullptr;     // nseg levels
    float4* fin = nullptr;      // final state
    float4* scratch = nullptr;  // one level for the forward
    float4* segs = nullptr;     // seg levels (level 0 unused: it is the checkpoint)
    float2* segV = nullptr;     // seg grid levels: node velocity
    float4* segB = nullptr;     // seg grid levels: (m, vh)
    int4* segBox = nullptr;     // seg grid levels: node box of the step
    int4* box = nullptr;        // two node boxes: the pending P2G's and the next one's
    float2* fwdV = nullptr;     // forward node velocity
    float4* acc = nullptr;
    float2* adjacc = nullptr;
    float4* qm = nullptr;
    float4* partial = nullptr;
    float4* bar = nullptr;      // staging of dL/d(final) and dL/d(state0)
    float* lam = nullptr;
    float* mu = nullptr;
    double* glam = nullptr;
    double* gmu = nullptr;
    float* gout = nullptr;      // float copies of the lambda/mu gradients
    uint32_t* status = nullptr;
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    bool use_graphs = true;
    hipGraphExec_t graph[2] = {nullptr, nullptr};  // captured forward / reverse sequences
    bool have_forward = false;
    uint32_t last_status = 0;
    std::string error;
};

namespace {

thread_local std::string g_create_error;  // message of the last failed mpm_diff_create

int diff_fail(mpm_diff* h, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    (h ? h->error : g_create_error) = buf;
    return code;
}

int diff_hip(mpm_diff* h, hipError_t e, const char* what) {
    return diff_fail(h, MPM_EDEVICE, "%s: %s", what, hipGetErrorString(e));
}

#define DIFF_CHECK(expr)                                         \
    do {                                                         \
        hipError_t _e = (expr);                                  \
        if (_e != hipSuccess) return diff_hip(h, _e, #expr);     \
    } while (0)

void release(mpm_diff* h) {
    void* bufs[] = {h->ckpt, h->fin, h->scratch, h->segs, h->segV, h->segB, h->segBox, h->box, h->fwdV, h->acc,
                    h->adjacc, h->qm,
                    h->partial, h->bar, h->lam, h->mu, h->glam, h->gmu, h->gout, h->status};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t& e : h->ev)
        if (e) (void)hipEventDestroy(e);
    for (hipGraphExec_t& g : h->graph)
        if (g) (void)hipGraphExecDestroy(g);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

inline unsigned grid_for(uint32_t count, int per) { return (count + per - 1) / per; }

float4* ckpt_level(mpm_diff* h, uint32_t j) { return h->ckpt + (size_t)j * h->level_f4; }
float4* seg_level(mpm_diff* h, uint32_t j, uint32_t k) {
    return k == 0 ? ckpt_level(h, j) : h->segs + (size_t)k * h->level_f4;
}

// State level of step index t (0 <= t <= T) while segment j is recorded.
float4* state_at(mpm_diff* h, uint32_t j, uint32_t t) {
    const uint32_t t0 = j * h->seg;
    if (t == h->steps) return h->fin;
    if (t >= t0 + h->seg) return ckpt_level(h, t / h->seg);
    return seg_level(h, j, t - t0);
}

hipError_t launch_particles(mpm_diff* h, bool g2p, bool p2g, const float4* src, float4* dst, const float2* nodeV,
                            int4* box) {
    const dim3 grid(grid_for(h->P.n, PT)), block(PT);
    if (g2p && p2g)
        k_diff_particles<true, true><<<grid, block, 0, h->stream>>>(h->P, src, dst, nodeV, h->acc, h->lam, h->mu, box);
    else if (g2p)
        k_diff_particles<true, false><<<grid, block, 0, h->stream>>>(h->P, src, dst, nodeV, h->acc, h->lam, h->mu, box);
    else
        k_diff_particles<false, true><<<grid, block, 0, h->stream>>>(h->P, src, dst, nodeV, h->acc, h->lam, h->mu, box);
    return hipGetLastError();
}

// Node pass of the P2G whose box is box[q]; resets box[q ^ 1] for the next.
hipError_t launch_nodes(mpm_diff* h, float2* nodeV, float4* nodeB, int q, int4* box_save) {
    k_diff_nodes<<<NODE_BLOCKS, NT, 0, h->stream>>>(h->P, h->acc, nodeV, nodeB, h->box + q, h->box + (q ^ 1),
                                                     box_save);
    return hipGetLastError();
}

hipError_t reset_boxes(mpm_diff* h) {
    k_diff_box_reset<<<1, 64, 0, h->stream>>>(h->box);
    return hipGetLastError();
}

}  // namespace

extern "C" {

int mpm_diff_create(const mpm_diff_config* cfg, mpm_diff** out) {
    if (!cfg || !out) return diff_fail(nullptr, MPM_EINVAL, "null argument");
    *out = nullptr;
    const mpm_diff_config c = *cfg;
    if (c.particle_count == 0 || c.particle_count > (1u << 24) || c.cells_x < 1 || c.cells_y < 1 ||
        c.cells_x > 16384 || c.cells_y > 16384 || c.steps < 1 || !(c.cell_size > 0.0f) || !(c.dt > 0.0f) ||
        !(c.particle_mass > 0.0f) || !(c.particle_volume > 0.0f) || !(c.wall_friction >= 0.0f) ||
        !(c.wall_friction <= 1.0f) || !std::isfinite(c.gravity_x) || !std::isfinite(c.gravity_y) ||
        !std::isfinite(c.cell_size) || !std::isfinite(c.dt) || (c.flags & ~(uint32_t)MPM_DIFF_NO_GRAPHS))
        return diff_fail(nullptr, MPM_EINVAL, "invalid rollout configuration");
    mpm_diff* h = new (std::nothrow) mpm_diff();
    if (!h) return diff_fail(nullptr, MPM_EDEVICE, "out of host memory");
    auto bail = [&](int code) {
        release(h);
        delete h;
        return code;
    };
    h->cfg = c;
    h->use_graphs = (c.flags & MPM_DIFF_NO_GRAPHS) == 0;
    h->steps = c.steps;
    h->seg = c.segment_steps ? (c.segment_steps < c.steps ? c.segment_steps : c.steps) : (c.steps < 20 ? c.steps : 20);
    h->nseg = (h->steps + h->seg - 1) / h->seg;
    const uint32_t n = c.particle_count;
    h->level_f4 = (size_t)3 * n;
    DiffConst& P = h->P;
    P.dt = c.dt;
    P.dx = c.cell_size;
    P.gs = 4.0f / (c.cell_size * c.cell_size);
    P.pm = c.particle_mass;
    P.ss = -c.dt * c.particle_volume * P.gs;
    P.gdx = c.dt * c.gravity_x;
    P.gdy = c.dt * c.gravity_y;
    P.friction = c.wall_friction;
    P.ext_x = (float)c.cells_x;
    P.ext_y = (float)c.cells_y;
    P.cush_x = fminf(CUSHION_CELLS, P.ext_x / 4.0f);
    P.cush_y = fminf(CUSHION_CELLS, P.ext_y / 4.0f);
    P.wd = (float)c.cells_x * c.cell_size;
    P.hd = (float)c.cells_y * c.cell_size;
    P.cells_x = (int)c.cells_x;
    P.cells_y = (int)c.cells_y;
    P.W = (int)c.cells_x + 3;
    P.nodes = P.W * ((int)c.cells_y + 3);
    P.n = n;

    hipError_t e;
    if (c.device >= 0 && (e = hipSetDevice(c.device)) != hipSuccess) return bail(diff_hip(nullptr, e, "hipSetDevice"));
    if ((e = hipGetDevice(&h->device)) != hipSuccess) return bail(diff_hip(nullptr, e, "hipGetDevice"));
    if ((e = hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(diff_hip(nullptr, e, "hipStreamCreateWithFlags"));
    for (hipEvent_t& ev : h->ev)
        if ((e = hipEventCreate(&ev)) != hipSuccess) return bail(diff_hip(nullptr, e, "hipEventCreate"));
    const size_t lvl = h->level_f4 * sizeof(float4), nodes = (size_t)P.nodes;
#define DIFF_ALLOC(ptr, bytes)                                                   \
    do {                                                                         \
        if ((e = hipMalloc((void**)&(ptr), (bytes))) != hipSuccess) return bail(diff_hip(nullptr, e, "hipMalloc")); \
    } while (0)
    DIFF_ALLOC(h->ckpt, lvl * h->nseg);
    DIFF_ALLOC(h->fin, lvl);
    DIFF_ALLOC(h->scratch, lvl);
    DIFF_ALLOC(h->segs, lvl * h->seg);
    DIFF_ALLOC(h->segV, nodes * h->seg * sizeof(float2));
    DIFF_ALLOC(h->segB, nodes * h->seg * sizeof(float4));
    DIFF_ALLOC(h->segBox, h->seg * sizeof(int4));
    DIFF_ALLOC(h->box, 2 * sizeof(int4));
    DIFF_ALLOC(h->fwdV, nodes * sizeof(float2));
    DIFF_ALLOC(h->acc, nodes * sizeof(float4));
    DIFF_ALLOC(h->adjacc, nodes * sizeof(float2));
    DIFF_ALLOC(h->qm, nodes * sizeof(float4));
    DIFF_ALLOC(h->partial, (size_t)2 * n * sizeof(float4));
    DIFF_ALLOC(h->bar, lvl);
    DIFF_ALLOC(h->lam, n * sizeof(float));
    DIFF_ALLOC(h->mu, n * sizeof(float));
    DIFF_ALLOC(h->glam, n * sizeof(double));
    DIFF_ALLOC(h->gmu, n * sizeof(double));
    DIFF_ALLOC(h->gout, (size_t)2 * n * sizeof(float));
    DIFF_ALLOC(h->status, sizeof(uint32_t));
#undef DIFF_ALLOC
    P.status = h->status;
    if (hipMemsetAsync(h->acc, 0, nodes * sizeof(float4), h->stream) != hipSuccess ||
        hipMemsetAsync(h->adjacc, 0, nodes * sizeof(float2), h->stream) != hipSuccess ||
        hipMemsetAsync(h->status, 0, sizeof(uint32_t), h->stream) != hipSuccess ||
        hipStreamSynchronize(h->stream) != hipSuccess)
        return bail(diff_fail(nullptr, MPM_EDEVICE, "clearing the rollout buffers failed"));
    *out = h;
    return MPM_OK;
}

}  // extern "C"

namespace {

__global__ void k_diff_grad_out(uint32_t n, const double* __restrict__ glam, const double* __restrict__ gmu,
                                float* __restrict__ out) {
    const uint32_t p = blockIdx.x * NT + threadIdx.x;
    if (p >= n) return;
    out[p] = (float)glam[p];
    out[n + p] = (float)gmu[p];
}

int finish(mpm_diff* h, const char* what) {
    uint32_t st = 0;
    hipError_t e = hipMemcpyAsync(&st, h->status, sizeof st, hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) return diff_hip(h, e, what);
    h->last_status = st;
    const uint32_t errors = st & ~(uint32_t)MPM_DIFF_CONFINED;
    if (errors)
        return diff_fail(h, MPM_ESTATE, "%s: status 0x%x (%s%s%s%s)", what, st,
                         (st & MPM_DIFF_STENCIL) ? "stencil outside the grid " : "",
                         (st & MPM_DIFF_DEFORMATION) ? "non-positive det F " : "",
                         (st & MPM_DIFF_NONFINITE) ? "non-finite value " : "",
                         (st & MPM_DIFF_SUBSTEPS) ? "deformation substep limit " : "");
    return MPM_OK;
}

// Kernel sequence of the forward rollout: P2G of state 0, then per step the
// node pass and the particle pass (G2P fused with the next P2G).  Every
// pointer is fixed for the life of the rollout, so the sequence can be
// captured once and replayed.
hipError_t enqueue_forward(mpm_diff* h) {
    const uint32_t T = h->steps, L = h->seg;
    int q = 0;
    hipError_t e = reset_boxes(h);
    if (e == hipSuccess) e = launch_particles(h, false, true, h->ckpt, nullptr, nullptr, h->box + q);
    const float4* src = h->ckpt;
    for (uint32_t t = 0; t < T && e == hipSuccess; ++t) {
        e = launch_nodes(h, h->fwdV, nullptr, q, nullptr);
        q ^= 1;
        const uint32_t t1 = t + 1;
        float4* dst = t1 == T ? h->fin : (t1 % L == 0 ? ckpt_level(h, t1 / L) : h->scratch);
        if (e == hipSuccess) e = launch_particles(h, true, t1 < T, src, dst, h->fwdV, h->box + q);
        src = dst;
    }
    return e;
}

// Kernel sequence of the reverse pass: for each segment, last first, record
// its states and grids from the checkpoint, then reverse its steps.
hipError_t enqueue_backward(mpm_diff* h) {
    const uint32_t n = h->P.n, T = h->steps, L = h->seg;
    const size_t nodes = (size_t)h->P.nodes;
    const dim3 pg(grid_for(n, PT)), pb(PT);
    hipError_t e = hipSuccess;
    for (int j = (int)h->nseg - 1; j >= 0 && e == hipSuccess; --j) {
        const uint32_t t0 = (uint32_t)j * L, len = (T - t0) < L ? (T - t0) : L;
        int q = 0;
        e = reset_boxes(h);
        if (e == hipSuccess) e = launch_particles(h, false, true, ckpt_level(h, j), nullptr, nullptr, h->box + q);
        if (e == hipSuccess) e = launch_nodes(h, h->segV, h->segB, q, h->segBox);
        q ^= 1;
        for (uint32_t k = 1; k < len && e == hipSuccess; ++k) {
            e = launch_particles(h, true, true, seg_level(h, j, k - 1), seg_level(h, j, k), h->segV + (k - 1) * nodes,
                                 h->box + q);
            if (e == hipSuccess) e = launch_nodes(h, h->segV + k * nodes, h->segB + k * nodes, q, h->segBox + k);
            q ^= 1;
        }
        for (int k = (int)len - 1; k >= 0 && e == hipSuccess; --k) {
            const uint32_t t = t0 + (uint32_t)k;
            const float4* s_cur = seg_level(h, j, (uint32_t)k);
            const float2* nv = h->segV + (size_t)k * nodes;
            if (t + 1 == T)
                k_diff_particles_bwd<false, true><<<pg, pb, 0, h->stream>>>(
                    h->P, nullptr, nullptr, s_cur, nv, h->adjacc, h->partial, h->bar, nullptr, h->lam, h->mu,
                    h->glam, h->gmu);
            else
                k_diff_particles_bwd<true, true><<<pg, pb, 0, h->stream>>>(
                    h->P, state_at(h, j, t + 1), h->qm, s_cur, nv, h->adjacc, h->partial, nullptr, nullptr, h->lam,
                    h->mu, h->glam, h->gmu);
            e = hipGetLastError();
            if (e != hipSuccess) break;
            k_diff_nodes_bwd<<<NODE_BLOCKS, NT, 0, h->stream>>>(h->P, h->adjacc, h->segB + (size_t)k * nodes, h->qm,
                                                                h->segBox + k);
            e = hipGetLastError();
        }
    }
    if (e != hipSuccess) return e;
    k_diff_particles_bwd<true, false><<<pg, pb, 0, h->stream>>>(h->P, h->ckpt, h->qm, nullptr, nullptr, nullptr,
                                                               h->partial, nullptr, h->bar, h->lam, h->mu, h->glam,
                                                               h->gmu);
    return hipGetLastError();
}

// Runs sequence 0 (forward) or 1 (reverse): replays its graph, capturing it
// on first use, or launches the kernels one by one when graphs are off.
hipError_t run_sequence(mpm_diff* h, int which) {
    if (!h->use_graphs) return which == 0 ? enqueue_forward(h) : enqueue_backward(h);
    if (!h->graph[which]) {
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return e;
        const hipError_t le = which == 0 ? enqueue_forward(h) : enqueue_backward(h);
        e = hipStreamEndCapture(h->stream, &g);
        if (le != hipSuccess || e != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return le != hipSuccess ? le : e;
        }
        e = hipGraphInstantiate(&h->graph[which], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) {
            h->graph[which] = nullptr;
            return e;
        }
    }
    return hipGraphLaunch(h->graph[which], h->stream);
}

}  // namespace

extern "C" {

int mpm_diff_forward(mpm_diff* h, const float* state0, const float* lambda, const float* mu, float* state_final) {
    if (!h || !state0 || !lambda || !mu) return MPM_EINVAL;
    DIFF_CHECK(hipSetDevice(h->device));
    const uint32_t n = h->P.n;
    const size_t lvl = h->level_f4 * sizeof(float4);
    h->have_forward = false;
    DIFF_CHECK(hipMemsetAsync(h->status, 0, sizeof(uint32_t), h->stream));
    DIFF_CHECK(hipMemcpyAsync(h->ckpt, state0, lvl, hipMemcpyDefault, h->stream));
    DIFF_CHECK(hipMemcpyAsync(h->lam, lambda, n * sizeof(float), hipMemcpyDefault, h->stream));
    DIFF_CHECK(hipMemcpyAsync(h->mu, mu, n * sizeof(float), hipMemcpyDefault, h->stream));
    DIFF_CHECK(hipEventRecord(h->ev[0], h->stream));
    DIFF_CHECK(run_sequence(h, 0));
    DIFF_CHECK(hipEventRecord(h->ev[1], h->stream));
    if (state_final) DIFF_CHECK(hipMemcpyAsync(state_final, h->fin, lvl, hipMemcpyDefault, h->stream));
    const int rc = finish(h, "forward");
    h->have_forward = rc == MPM_OK;
    return rc;
}

int mpm_diff_backward(mpm_diff* h, const float* bar_final, float* bar_state0, float* grad_lambda, float* grad_mu) {
    if (!h || !bar_final) return MPM_EINVAL;
    if (!h->have_forward) return diff_fail(h, MPM_ESTATE, "backward needs a successful forward first");
    DIFF_CHECK(hipSetDevice(h->device));
    const uint32_t n = h->P.n;
    const size_t lvl = h->level_f4 * sizeof(float4);
    DIFF_CHECK(hipMemsetAsync(h->status, 0, sizeof(uint32_t), h->stream));
    DIFF_CHECK(hipMemcpyAsync(h->bar, bar_final, lvl, hipMemcpyDefault, h->stream));
    DIFF_CHECK(hipMemsetAsync(h->glam, 0, n * sizeof(double), h->stream));
    DIFF_CHECK(hipMemsetAsync(h->gmu, 0, n * sizeof(double), h->stream));
    DIFF_CHECK(hipEventRecord(h->ev[2], h->stream));
    DIFF_CHECK(run_sequence(h, 1));
    DIFF_CHECK(hipEventRecord(h->ev[3], h->stream));
    k_diff_grad_out<<<grid_for(n, NT), NT, 0, h->stream>>>(n, h->glam, h->gmu, h->gout);
    DIFF_CHECK(hipGetLastError());
    if (bar_state0) DIFF_CHECK(hipMemcpyAsync(bar_state0, h->bar, lvl, hipMemcpyDefault, h->stream));
    if (grad_lambda) DIFF_CHECK(hipMemcpyAsync(grad_lambda, h->gout, n * sizeof(float), hipMemcpyDefault, h->stream));
    if (grad_mu) DIFF_CHECK(hipMemcpyAsync(grad_mu, h->gout + n, n * sizeof(float), hipMemcpyDefault, h->stream));
    return finish(h, "backward");
}

int mpm_diff_read_status(mpm_diff* h, uint32_t* status) {
    if (!h || !status) return MPM_EINVAL;
    *status = h->last_status;
    return MPM_OK;
}

int mpm_diff_read_timing(mpm_diff* h, double* forward_ms, double* backward_ms) {
    if (!h) return MPM_EINVAL;
    float f = 0.0f, b = 0.0f;
    if (forward_ms) {
        if (hipEventElapsedTime(&f, h->ev[0], h->ev[1]) != hipSuccess) f = -1.0f;
        *forward_ms = f;
    }
    if (backward_ms) {
        if (hipEventElapsedTime(&b, h->ev[2], h->ev[3]) != hipSuccess) b = -1.0f;
        *backward_ms = b;
    }
    return MPM_OK;
}

int mpm_diff_get_layout(mpm_diff* h, uint32_t* segment_steps, uint32_t* checkpoints, size_t* device_bytes) {
    if (!h) return MPM_EINVAL;
    if (segment_steps) *segment_steps = h->seg;
    if (checkpoints) *checkpoints = h->nseg;
    if (device_bytes) {
        const size_t lvl = h->level_f4 * sizeof(float4), nodes = (size_t)h->P.nodes, n = h->P.n;
        *device_bytes = lvl * (h->nseg + 3 + h->seg) + nodes * h->seg * (sizeof(float2) + sizeof(float4)) +
                        (h->seg + 2) * sizeof(int4) +
                        nodes * (sizeof(float2) * 2 + sizeof(float4) * 2) + n * (2 * sizeof(float4) + 2 * sizeof(float) +
                        2 * sizeof(double) + 2 * sizeof(float)) + sizeof(uint32_t);
    }
    return MPM_OK;
}

void mpm_diff_destroy(mpm_diff* h) {
    if (!h) return;
    (void)hipSetDevice(h->device);
    (void)hipStreamSynchronize(h->stream);
    release(h);
    delete h;
}

const char* mpm_diff_last_error(mpm_diff* h) { return h ? h->error.c_str() : g_create_error.c_str(); }

}  // extern "C"
