// C-ABI host layer: context lifetime, validation, uploads, readbacks.
// Replaces the WebGPU runtime below MpmMlsGpuSolver
// (src/lib/mpm/kernels/solver.ts:408-1230, :1407-1599, :2388-2731, :2950-2964).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <new>
#include <string>
#include <vector>

#include "../../include/mpm355.h"
#include "mpm_internal.h"
#include "mpm_timing.h"

using namespace mpm;

namespace {

constexpr int PARAM_RING = 16;

thread_local std::string g_create_error = "no error";

struct Allocation {
    std::vector<void*> ptrs;
    hipError_t alloc(void** p, size_t bytes) {
        *p = nullptr;
        if (bytes == 0) bytes = 16;
        hipError_t e = hipMalloc(p, bytes);
        if (e == hipSuccess) ptrs.push_back(*p);
        return e;
    }
    void release() {
        for (void* p : ptrs) (void)hipFree(p);
        ptrs.clear();
    }
};

uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

}  // namespace

struct mpm_ctx {
    mpm_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    Allocation mem;
    LaunchCtx lc{};
    Params* d_params = nullptr;
    Control* d_ctl = nullptr;
    Params* h_param_ring = nullptr;       // pinned
    hipEvent_t ring_events[PARAM_RING] = {};
    bool ring_used[PARAM_RING] = {};
    int ring_next = 0;
    Control* h_ctl = nullptr;             // pinned readback
    uint32_t* h_object_ids = nullptr;     // host copy of object ids
    float interaction[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t n = 0;
    uint32_t node_w = 0, node_h = 0, node_count = 0;
    int cur = 0;                          // buffer holding the committed state
    std::string error = "no error";
    KernelTimes* times = nullptr;
};

namespace {

int fail(mpm_ctx* c, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (c) c->error = buf; else g_create_error = buf;
    return code;
}

int hip_fail(mpm_ctx* c, hipError_t e, const char* what) {
    return fail(c, MPM_EDEVICE, "%s failed: %s", what, hipGetErrorString(e));
}

#define HIP_TRY(ctx, expr)                                   \
    do {                                                     \
        hipError_t _e = (expr);                              \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #expr); \
    } while (0)

// validateMpmMlsGpuInput per-record checks at the float32 level
// (layout.ts:300-437): model code, finite state, positive volume/density,
// damage and plastic-volume ranges, domain bounds and a positive packed det.
int validate_records(mpm_ctx* c, const mpm_config& cfg, const uint32_t* rec, uint32_t n) {
    const float dom_w = (float)cfg.cells_x * cfg.cell_size;
    const float dom_h = (float)cfg.cells_y * cfg.cell_size;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t* r = rec + 20ull * i;
        const uint32_t model = r[13];
        if (model > 4u)
            return fail(c, MPM_EINVAL, "particles[%u].model code %u is not a supported material", i, model);
        if (r[19] != 0u)
            return fail(c, MPM_EINVAL, "particles[%u] is a bonded-triangle vertex; bonded bodies are not supported", i);
        for (int k = 0; k < 19; ++k) {
            if (k == 13) continue;
            if (!isfinite(u2f(r[k]))) return fail(c, MPM_EINVAL, "particles[%u] word %d must be finite", i, k);
        }
        const float x = u2f(r[0]), y = u2f(r[1]);
        if (x < 0.0f || x > dom_w || y < 0.0f || y > dom_h)
            return fail(c, MPM_EINVAL, "particles[%u].position is outside the float32 GPU grid domain", i);
        const float det = u2f(f2u(u2f(r[8]) * u2f(r[11]))) - u2f(r[9]) * u2f(r[10]);
        if (!(det > 0.0f))
            return fail(c, MPM_EINVAL, "particles[%u].F must retain a finite positive determinant in float32", i);
        const float pv = u2f(r[12]), volume = u2f(r[17]), density = u2f(r[16]), param = u2f(r[18]);
        if (pv < 0.6f || pv > 20.0f)
            return fail(c, MPM_EINVAL, "particles[%u].plasticVolume must be between 0.6 and 20", i);
        if (!(volume > 0.0f)) return fail(c, MPM_EINVAL, "particles[%u].volume must be greater than zero", i);
        if ((model == 2u || model == 3u) && !(density > 0.0f))
            return fail(c, MPM_EINVAL, "particles[%u].material.density must be greater than zero", i);
        if (model == 0u && (param < 0.0f || param > 1.0f))
            return fail(c, MPM_EINVAL, "particles[%u].damage must be between zero and one", i);
    }
    return MPM_OK;
}

void unpack_records(const uint32_t* rec, uint32_t n, std::vector<float2>& pos, std::vector<float2>& vel,
                    std::vector<float4>& C, std::vector<float4>& F, std::vector<float>& pl, std::vector<uint32_t>& model,
                    std::vector<float>& lambda, std::vector<float>& mu, std::vector<float>& density,
                    std::vector<float>& volume, std::vector<float>& param) {
    pos.resize(n); vel.resize(n); C.resize(n); F.resize(n); pl.resize(n); model.resize(n);
    lambda.resize(n); mu.resize(n); density.resize(n); volume.resize(n); param.resize(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t* r = rec + 20ull * i;
        pos[i] = make_float2(u2f(r[0]), u2f(r[1]));
        vel[i] = make_float2(u2f(r[2]), u2f(r[3]));
        C[i] = make_float4(u2f(r[4]), u2f(r[5]), u2f(r[6]), u2f(r[7]));
        F[i] = make_float4(u2f(r[8]), u2f(r[9]), u2f(r[10]), u2f(r[11]));
        pl[i] = u2f(r[12]);
        model[i] = r[13];
        lambda[i] = u2f(r[14]);
        mu[i] = u2f(r[15]);
        density[i] = u2f(r[16]);
        volume[i] = u2f(r[17]);
        param[i] = u2f(r[18]);
    }
}

template <class T>
hipError_t upload(T* dst, const std::vector<T>& src, uint32_t offset, hipStream_t s) {
    if (src.empty()) return hipSuccess;
    return hipMemcpyAsync(dst + offset, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice, s);
}

int write_particles(mpm_ctx* c, const uint32_t* rec, const uint32_t* ids, uint32_t n, uint32_t offset) {
    std::vector<float2> pos, vel;
    std::vector<float4> C, F;
    std::vector<float> pl, lambda, mu, density, volume, param;
    std::vector<uint32_t> model;
    unpack_records(rec, n, pos, vel, C, F, pl, model, lambda, mu, density, volume, param);
    const ParticleBuf& b = c->lc.buf[c->cur];
    HIP_TRY(c, upload(b.pos, pos, offset, c->stream));
    HIP_TRY(c, upload(b.vel, vel, offset, c->stream));
    HIP_TRY(c, upload(b.C, C, offset, c->stream));
    HIP_TRY(c, upload(b.F, F, offset, c->stream));
    HIP_TRY(c, upload(b.plastic, pl, offset, c->stream));
    HIP_TRY(c, upload(c->lc.st.model, model, offset, c->stream));
    HIP_TRY(c, upload(c->lc.st.lambda, lambda, offset, c->stream));
    HIP_TRY(c, upload(c->lc.st.mu, mu, offset, c->stream));
    HIP_TRY(c, upload(c->lc.st.density, density, offset, c->stream));
    HIP_TRY(c, upload(c->lc.st.volume, volume, offset, c->stream));
    HIP_TRY(c, upload(c->lc.st.param, param, offset, c->stream));
    for (uint32_t i = 0; i < n; ++i) {
        c->h_object_ids[offset + i] = ids ? ids[i] : 0u;
        const uint32_t m = model[i];
        c->lc.scene_flags |= m == 2u ? SCENE_WATER : (m == 3u ? SCENE_OIL : SCENE_SOLID);
        if ((m == 2u || m == 3u) && param[i] > 0.0f) c->lc.scene_flags |= SCENE_TAIT;
    }
    HIP_TRY(c, hipStreamSynchronize(c->stream));  // host vectors go out of scope
    return MPM_OK;
}

int sync(mpm_ctx* c) {
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (c->times) c->times->resolve();
    return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_create(const mpm_config* cfg, const uint32_t* packed20, const uint32_t* object_ids, uint32_t n,
               mpm_ctx** out) {
    if (!out) return fail(nullptr, MPM_EINVAL, "out must not be NULL");
    *out = nullptr;
    if (!cfg) return fail(nullptr, MPM_EINVAL, "config must not be NULL");
    if (n > 0 && !packed20) return fail(nullptr, MPM_EINVAL, "packed particles must not be NULL");
    const mpm_config c = *cfg;
    if (c.cells_x < 1 || c.cells_y < 1 || c.cells_x > 65535u || c.cells_y > 65535u)
        return fail(nullptr, MPM_EINVAL, "grid.cellsX and grid.cellsY must be in [1, 65535]");
    if (!(c.cell_size > 0.0f) || !isfinite(c.cell_size))
        return fail(nullptr, MPM_EINVAL, "grid.cellSize must be greater than zero");
    if (!(c.particle_mass > 0.0f) || !isfinite(c.particle_mass))
        return fail(nullptr, MPM_EINVAL, "particleMass must be greater than zero");
    if (!(c.particle_volume > 0.0f) || !isfinite(c.particle_volume))
        return fail(nullptr, MPM_EINVAL, "particleVolume must be greater than zero");
    if (!(c.wall_friction >= 0.0f && c.wall_friction <= 1.0f))
        return fail(nullptr, MPM_EINVAL, "wallFriction must be between zero and one");
    if (c.p2g_mode > MPM_P2G_FIXED_POINT) return fail(nullptr, MPM_EINVAL, "p2g_mode must be 0 or 1");
    const uint32_t capacity = c.capacity == 0 ? n : c.capacity;
    if (capacity < n) return fail(nullptr, MPM_EINVAL, "capacity must be at least the initial particle count");
    if (capacity > (1u << 30)) return fail(nullptr, MPM_EINVAL, "capacity exceeds the supported 2^30 particle slots");
    const uint64_t nodes64 = (uint64_t)(c.cells_x + 3) * (uint64_t)(c.cells_y + 3);
    if (nodes64 >= 0xffffffffull) return fail(nullptr, MPM_EINVAL, "halo grid dimensions exceed the GPU index range");
    {
        int v = validate_records(nullptr, c, packed20, n);
        if (v) return v;
    }

    mpm_ctx* ctx = new (std::nothrow) mpm_ctx();
    if (!ctx) return fail(nullptr, MPM_EDEVICE, "out of host memory");
    ctx->cfg = c;
    ctx->cfg.capacity = capacity;
    ctx->node_w = c.cells_x + 3;
    ctx->node_h = c.cells_y + 3;
    ctx->node_count = (uint32_t)nodes64;
    auto bail = [&](int code) {
        g_create_error = ctx->error;
        mpm_destroy(ctx);
        return code;
    };
    hipError_t e;
    if (c.device >= 0) {
        if ((e = hipSetDevice(c.device)) != hipSuccess) return bail(hip_fail(ctx, e, "hipSetDevice"));
    }
    if ((e = hipGetDevice(&ctx->device)) != hipSuccess) return bail(hip_fail(ctx, e, "hipGetDevice"));
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipStreamCreateWithFlags"));

    const size_t cap = capacity ? capacity : 1;
    const size_t nodes = ctx->node_count;
    LaunchCtx& lc = ctx->lc;
    Allocation& m = ctx->mem;
#define ALLOC(ptr, bytes)                                                                \
    do {                                                                                 \
        if ((e = m.alloc((void**)&(ptr), (bytes))) != hipSuccess)                        \
            return bail(fail(ctx, MPM_EDEVICE, "hipMalloc(%zu bytes) failed: %s", (size_t)(bytes), hipGetErrorString(e))); \
    } while (0)
    for (int b = 0; b < 2; ++b) {
        ALLOC(lc.buf[b].pos, cap * sizeof(float2));
        ALLOC(lc.buf[b].vel, cap * sizeof(float2));
        ALLOC(lc.buf[b].C, cap * sizeof(float4));
        ALLOC(lc.buf[b].F, cap * sizeof(float4));
        ALLOC(lc.buf[b].plastic, cap * sizeof(float));
    }
    ALLOC(lc.st.model, cap * sizeof(uint32_t));
    ALLOC(lc.st.lambda, cap * sizeof(float));
    ALLOC(lc.st.mu, cap * sizeof(float));
    ALLOC(lc.st.density, cap * sizeof(float));
    ALLOC(lc.st.volume, cap * sizeof(float));
    ALLOC(lc.st.param, cap * sizeof(float));
    ALLOC(lc.bins.key, cap * sizeof(uint32_t));
    ALLOC(lc.bins.rank, cap * sizeof(uint32_t));
    ALLOC(lc.bins.perm, cap * sizeof(uint32_t));
    ALLOC(lc.bins.count, nodes * sizeof(uint32_t));
    ALLOC(lc.bins.start, (nodes + 1) * sizeof(uint32_t));
    ALLOC(lc.bins.block_sums, scan_block_count(ctx->node_count) * sizeof(uint32_t));
    ALLOC(lc.bins.affine, cap * sizeof(float4));
    ALLOC(lc.bins.aux, cap * sizeof(float4));
    ALLOC(lc.grid.v_solid, nodes * sizeof(float2));
    ALLOC(lc.grid.v_water, nodes * sizeof(float2));
    ALLOC(lc.grid.v_oil, nodes * sizeof(float2));
    ALLOC(lc.grid.liquid, nodes * sizeof(float2));
    const bool keep_fixed = c.p2g_mode == MPM_P2G_FIXED_POINT && (c.flags & MPM_FLAG_KEEP_FIXED_GRID);
    if (keep_fixed) ALLOC(lc.grid.fixed, nodes * 11 * sizeof(int32_t));
    ALLOC(ctx->d_params, sizeof(Params));
    ALLOC(ctx->d_ctl, sizeof(Control));
#undef ALLOC
    if ((e = hipHostMalloc((void**)&ctx->h_param_ring, sizeof(Params) * PARAM_RING, hipHostMallocDefault)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipHostMalloc(params)"));
    if ((e = hipHostMalloc((void**)&ctx->h_ctl, sizeof(Control), hipHostMallocDefault)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipHostMalloc(control)"));
    for (int k = 0; k < PARAM_RING; ++k)
        if ((e = hipEventCreateWithFlags(&ctx->ring_events[k], hipEventDisableTiming)) != hipSuccess)
            return bail(hip_fail(ctx, e, "hipEventCreate"));
    ctx->h_object_ids = new (std::nothrow) uint32_t[cap];
    if (!ctx->h_object_ids) return bail(fail(ctx, MPM_EDEVICE, "out of host memory"));

    // grid arrays start at zero so debug reads of absent phases are defined
    for (void* p : {(void*)lc.grid.v_solid, (void*)lc.grid.v_water, (void*)lc.grid.v_oil, (void*)lc.grid.liquid})
        if ((e = hipMemsetAsync(p, 0, nodes * sizeof(float2), ctx->stream)) != hipSuccess)
            return bail(hip_fail(ctx, e, "hipMemsetAsync(grid)"));
    if (keep_fixed && (e = hipMemsetAsync(lc.grid.fixed, 0, nodes * 11 * sizeof(int32_t), ctx->stream)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipMemsetAsync(fixed grid)"));
    if ((e = hipMemsetAsync(lc.bins.count, 0, nodes * sizeof(uint32_t), ctx->stream)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipMemsetAsync(count)"));
    if ((e = hipMemsetAsync(lc.bins.start, 0, (nodes + 1) * sizeof(uint32_t), ctx->stream)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipMemsetAsync(start)"));
    if ((e = hipMemsetAsync(lc.bins.key, 0xff, cap * sizeof(uint32_t), ctx->stream)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipMemsetAsync(key)"));
    Control init{};
    init.status[1] = 0xffffffffu;  // solver.ts:889-893
    *ctx->h_ctl = init;
    if ((e = hipMemcpyAsync(ctx->d_ctl, ctx->h_ctl, sizeof(Control), hipMemcpyHostToDevice, ctx->stream)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipMemcpyAsync(control)"));

    lc.stream = ctx->stream;
    lc.params = ctx->d_params;
    lc.ctl = ctx->d_ctl;
    lc.node_w = ctx->node_w;
    lc.node_h = ctx->node_h;
    lc.node_count = ctx->node_count;
    lc.fixed_point = c.p2g_mode == MPM_P2G_FIXED_POINT ? 1u : 0u;
    lc.keep_fixed = keep_fixed ? 1u : 0u;
    lc.scene_flags = 0u;
    lc.scan_blocks = (uint32_t)scan_block_count(ctx->node_count);
    ctx->cur = 0;
    if (n > 0) {
        int w = write_particles(ctx, packed20, object_ids, n, 0);
        if (w) return bail(w);
    }
    ctx->n = n;
    lc.n = n;
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return bail(hip_fail(ctx, e, "hipStreamSynchronize"));
    *out = ctx;
    return MPM_OK;
}

int mpm_step_batch(mpm_ctx* c, const void* params144, uint32_t step_count, uint32_t measure_speed) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!params144) return fail(c, MPM_EINVAL, "params must not be NULL");
    if (step_count < 1) return fail(c, MPM_EINVAL, "stepCount must be a positive safe integer.");
    Params p{};
    memcpy(&p, params144, MPM_PARAMS_BYTES);
    if (p.cells_x != c->cfg.cells_x || p.cells_y != c->cfg.cells_y || p.node_w != c->node_w || p.node_h != c->node_h)
        return fail(c, MPM_EINVAL, "params grid does not match the solver grid");
    if (p.node_count != c->node_count) return fail(c, MPM_EINVAL, "params nodeCount does not match the solver grid");
    if (p.particle_count != c->n)
        return fail(c, MPM_EINVAL, "params particleCount %u does not match the %u allocated particle slots",
                    p.particle_count, c->n);
    if (p.body_count != 0u) return fail(c, MPM_EINVAL, "bonded triangle bodies are not supported");
    if (!(p.dt > 0.0f) || !isfinite(p.dt)) return fail(c, MPM_EINVAL, "dt must be greater than zero");
    if (p.dx != c->cfg.cell_size) return fail(c, MPM_EINVAL, "params cellSize does not match the solver grid");
    if (!isfinite(p.gx) || !isfinite(p.gy)) return fail(c, MPM_EINVAL, "gravity must be finite");
    if (!isfinite(p.particle_mass) || !(p.particle_mass > 0.0f) || !isfinite(p.particle_volume) ||
        !(p.particle_volume > 0.0f) || !isfinite(p.grad_scale) || !(p.grad_scale > 0.0f))
        return fail(c, MPM_EINVAL, "params physical block is invalid");
    if (p.cfl_on >= 0.5f && (!isfinite(p.wave_bound) || !isfinite(p.vel_bound) || p.vel_bound < 0.0f))
        return fail(c, MPM_EINVAL, "CFL bounds must be finite and non-negative");
    p.drag_target_x = c->interaction[0];
    p.drag_target_y = c->interaction[1];
    p.drag_radius = c->interaction[2];
    p.drag_stiffness = c->interaction[3];
    p.drag_damping = c->interaction[4];
    HIP_TRY(c, hipSetDevice(c->device));
    const int slot = c->ring_next;
    c->ring_next = (c->ring_next + 1) % PARAM_RING;
    if (c->ring_used[slot]) HIP_TRY(c, hipEventSynchronize(c->ring_events[slot]));
    c->h_param_ring[slot] = p;
    HIP_TRY(c, hipMemcpyAsync(c->d_params, &c->h_param_ring[slot], sizeof(Params), hipMemcpyHostToDevice, c->stream));
    HIP_TRY(c, hipEventRecord(c->ring_events[slot], c->stream));
    c->ring_used[slot] = true;
    LaunchCtx lc = c->lc;
    lc.timing = c->times;
    const bool measure = measure_speed != 0u || p.cfl_on >= 0.5f;
    HIP_TRY(c, launch_batch(lc, c->cur, step_count, measure));
    c->cur = (c->cur + (int)(step_count & 1u)) & 1;
    return MPM_OK;
}

int mpm_read_compact(mpm_ctx* c, float* max_speed, uint32_t* completed, uint32_t status[8]) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    int s = sync(c);
    if (s) return s;
    HIP_TRY(c, hipMemcpy(c->h_ctl, c->d_ctl, sizeof(Control), hipMemcpyDeviceToHost));
    if (max_speed) *max_speed = u2f(c->h_ctl->max_speed_bits);
    if (completed) *completed = c->h_ctl->completed;
    if (status) memcpy(status, c->h_ctl->status, sizeof(uint32_t) * 8);
    return MPM_OK;
}

int mpm_read_status(mpm_ctx* c, uint32_t status[8]) { return mpm_read_compact(c, nullptr, nullptr, status); }

int mpm_clear_status(mpm_ctx* c) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    int s = sync(c);
    if (s) return s;
    uint32_t init[8] = {0u, 0xffffffffu, 0u, 0u, 0u, 0u, 0u, 0u};
    HIP_TRY(c, hipMemcpy(c->d_ctl->status, init, sizeof init, hipMemcpyHostToDevice));
    return MPM_OK;
}

int mpm_read_particles(mpm_ctx* c, const uint32_t* indices, uint32_t n, uint32_t* out20, uint32_t* object_ids) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (n > 0 && !out20) return fail(c, MPM_EINVAL, "output must not be NULL");
    if (!indices && n > c->n) return fail(c, MPM_EINVAL, "cannot read %u particles from %u slots", n, c->n);
    if (indices)
        for (uint32_t k = 0; k < n; ++k)
            if (indices[k] >= c->n) return fail(c, MPM_EINVAL, "particle index %u is out of range", indices[k]);
    int s = sync(c);
    if (s) return s;
    const uint32_t total = indices ? c->n : n;
    std::vector<float2> pos(total), vel(total);
    std::vector<float4> C(total), F(total);
    std::vector<float> pl(total), lambda(total), mu(total), density(total), volume(total), param(total);
    std::vector<uint32_t> model(total);
    const ParticleBuf& b = c->lc.buf[c->cur];
    if (total > 0) {
        HIP_TRY(c, hipMemcpy(pos.data(), b.pos, total * sizeof(float2), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(vel.data(), b.vel, total * sizeof(float2), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(C.data(), b.C, total * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(F.data(), b.F, total * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(pl.data(), b.plastic, total * sizeof(float), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(model.data(), c->lc.st.model, total * sizeof(uint32_t), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(lambda.data(), c->lc.st.lambda, total * sizeof(float), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(mu.data(), c->lc.st.mu, total * sizeof(float), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(density.data(), c->lc.st.density, total * sizeof(float), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(volume.data(), c->lc.st.volume, total * sizeof(float), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(param.data(), c->lc.st.param, total * sizeof(float), hipMemcpyDeviceToHost));
    }
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t i = indices ? indices[k] : k;
        uint32_t* r = out20 + 20ull * k;
        r[0] = f2u(pos[i].x); r[1] = f2u(pos[i].y); r[2] = f2u(vel[i].x); r[3] = f2u(vel[i].y);
        r[4] = f2u(C[i].x); r[5] = f2u(C[i].y); r[6] = f2u(C[i].z); r[7] = f2u(C[i].w);
        r[8] = f2u(F[i].x); r[9] = f2u(F[i].y); r[10] = f2u(F[i].z); r[11] = f2u(F[i].w);
        r[12] = f2u(pl[i]); r[13] = model[i]; r[14] = f2u(lambda[i]); r[15] = f2u(mu[i]);
        r[16] = f2u(density[i]); r[17] = f2u(volume[i]); r[18] = f2u(param[i]); r[19] = 0u;
        if (object_ids) object_ids[k] = c->h_object_ids[i];
    }
    return MPM_OK;
}

int mpm_append_particles(mpm_ctx* c, const uint32_t* packed20, const uint32_t* object_ids, uint32_t n) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (n == 0) return MPM_OK;
    if (!packed20) return fail(c, MPM_EINVAL, "packed particles must not be NULL");
    if ((uint64_t)c->n + n > c->cfg.capacity)
        return fail(c, MPM_ECAPACITY, "appending %u particles exceeds the capacity of %u", n, c->cfg.capacity);
    int v = validate_records(c, c->cfg, packed20, n);
    if (v) return v;
    int s = sync(c);
    if (s) return s;
    int w = write_particles(c, packed20, object_ids, n, c->n);
    if (w) return w;
    c->n += n;
    c->lc.n = c->n;
    return MPM_OK;
}

int mpm_set_interaction(mpm_ctx* c, const float p[8]) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!p) return fail(c, MPM_EINVAL, "interaction parameters must not be NULL");
    for (int k = 0; k < 8; ++k)
        if (!isfinite(p[k])) return fail(c, MPM_EINVAL, "interaction parameters must be finite");
    if (p[2] < 0.0f || p[3] < 0.0f || p[4] < 0.0f)
        return fail(c, MPM_EINVAL, "drag radius, stiffness and damping must be non-negative");
    memcpy(c->interaction, p, sizeof c->interaction);
    return MPM_OK;
}

int mpm_get_counts(mpm_ctx* c, uint32_t* particle_count, uint32_t* capacity, uint32_t* node_count) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (particle_count) *particle_count = c->n;
    if (capacity) *capacity = c->cfg.capacity;
    if (node_count) *node_count = c->node_count;
    return MPM_OK;
}

int mpm_read_fixed_grid(mpm_ctx* c, int32_t* out, size_t words) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!c->lc.grid.fixed) return fail(c, MPM_ESTATE, "fixed-point grid retention is not enabled on this solver");
    if (!out || words != (size_t)c->node_count * 11) return fail(c, MPM_EINVAL, "output must hold node_count*11 words");
    int s = sync(c);
    if (s) return s;
    HIP_TRY(c, hipMemcpy(out, c->lc.grid.fixed, words * sizeof(int32_t), hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_read_grid(mpm_ctx* c, float* out, size_t floats) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!out || floats != (size_t)c->node_count * 8) return fail(c, MPM_EINVAL, "output must hold node_count*8 floats");
    int s = sync(c);
    if (s) return s;
    std::vector<float2> a(c->node_count);
    const float2* src[4] = {c->lc.grid.v_solid, c->lc.grid.v_water, c->lc.grid.v_oil, c->lc.grid.liquid};
    for (int k = 0; k < 4; ++k) {
        HIP_TRY(c, hipMemcpy(a.data(), src[k], a.size() * sizeof(float2), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < c->node_count; ++i) {
            out[8ull * i + 2 * k] = a[i].x;
            out[8ull * i + 2 * k + 1] = a[i].y;
        }
    }
    return MPM_OK;
}

int mpm_read_bins(mpm_ctx* c, uint32_t* keys, uint32_t n, uint32_t* starts, size_t start_words) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (n > c->n) return fail(c, MPM_EINVAL, "cannot read %u keys from %u slots", n, c->n);
    if (starts && start_words != (size_t)c->node_count + 1)
        return fail(c, MPM_EINVAL, "starts must hold node_count+1 words");
    int s = sync(c);
    if (s) return s;
    if (keys && n) HIP_TRY(c, hipMemcpy(keys, c->lc.bins.key, n * sizeof(uint32_t), hipMemcpyDeviceToHost));
    if (starts) HIP_TRY(c, hipMemcpy(starts, c->lc.bins.start, start_words * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_set_timing(mpm_ctx* c, uint32_t enable) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    int s = sync(c);
    if (s) return s;
    if (enable) {
        if (!c->times) {
            c->times = new (std::nothrow) KernelTimes();
            if (!c->times) return fail(c, MPM_EDEVICE, "out of host memory");
        }
        c->times->reset();
    } else {
        delete c->times;
        c->times = nullptr;
    }
    return MPM_OK;
}

int mpm_read_timing(mpm_ctx* c, const char* kernel, double* mean_us, uint64_t* launches) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!kernel) return fail(c, MPM_EINVAL, "kernel name must not be NULL");
    if (!c->times) return fail(c, MPM_ESTATE, "timing is not enabled");
    int s = sync(c);
    if (s) return s;
    auto it = c->times->totals.find(kernel);
    if (it == c->times->totals.end() || it->second.second == 0) {
        if (mean_us) *mean_us = 0.0;
        if (launches) *launches = 0;
        return MPM_OK;
    }
    if (mean_us) *mean_us = 1000.0 * it->second.first / (double)it->second.second;
    if (launches) *launches = it->second.second;
    return MPM_OK;
}

void mpm_destroy(mpm_ctx* c) {
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c->times;
    c->mem.release();
    for (int k = 0; k < PARAM_RING; ++k)
        if (c->ring_events[k]) (void)hipEventDestroy(c->ring_events[k]);
    if (c->h_param_ring) (void)hipHostFree(c->h_param_ring);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    delete[] c->h_object_ids;
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char* mpm_last_error(mpm_ctx* c) { return c ? c->error.c_str() : g_create_error.c_str(); }

}  // extern "C"
