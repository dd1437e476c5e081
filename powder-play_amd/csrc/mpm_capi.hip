// C-ABI host layer: context lifetime, validation, uploads, readbacks.
// Replaces the WebGPU runtime below MpmMlsGpuSolver
// (src/lib/mpm/kernels/solver.ts:408-1230, :1407-1599, :2388-2731, :2950-2964).
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/mpm355.h"
#include "mpm_internal.h"
#include "mpm_timing.h"

using namespace mpm;

namespace {

constexpr int PARAM_RING = 16;
constexpr uint32_t MIN_BUCKET_CAP = 32;
constexpr uint32_t MAX_BUCKET_CAP = 1024;  // a bucket must fit one LDS chunk of the P2G tile
constexpr uint32_t MAX_PENDING = CONTROL_HISTORY;

thread_local std::string g_create_error = "no error";

uint32_t f2u(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
float u2f(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }

// mpm_tuning.tile_queue: 1 on, 2 off, 0 on when most tiles of the grid hold
// no particle (more than two grid nodes per particle) and there are more
// tiles than the device runs at once (four P2G tiles per compute unit); a
// smaller launch runs all its tiles in one round, so skipping the empty ones
// saves less than the compaction launch costs.
bool queue_wanted(uint32_t mode, uint64_t nodes, uint64_t particles, uint32_t p2g_tiles, int device) {
    if (mode != 0u) return mode == 1u;
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        return false;
    return nodes > 2u * particles && p2g_tiles >= 4u * (uint32_t)cus;
}

struct DeviceArrays {
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

struct Pending {
    Params params;
    uint32_t steps;
    bool measure;
    uint32_t seq;
    uint32_t done_before;  // microsteps of the caller's batch committed before a replay
};

// Host copy of the per-particle record fields that never change on the GPU.
struct HostStatic {
    std::vector<uint32_t> model, object_id, vertex;  // vertex: record word 19 (bonded vertex + 1, 0 = none)
    std::vector<float> lambda, mu, density, volume, param;
};

}  // namespace

struct mpm_ctx {
    mpm_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;
    DeviceArrays mem;       // home arrays, grid, control, params, materials
    DeviceArrays slot_mem;  // bucket slot buffers (reallocated when they grow)
    LaunchCtx lc{};
    Control* d_ctl = nullptr;
    Material* d_mats = nullptr;
    uint32_t mat_capacity = 0;
    SceneDesc* h_desc_ring = nullptr; // pinned ring: one descriptor (with its Params) per launch
    SceneDesc* d_scene = nullptr;     // device: this context as a one-scene step launch
    uint32_t* d_p2g_map = nullptr;    // device: tile -> scene (all 0; a one-scene launch does not read it)
    uint32_t* d_g2p_map = nullptr;
    uint32_t* d_queue = nullptr;      // device: tile queue counters and lists (sized for the most tiles)
    uint32_t queue_p2g = 0, queue_g2p = 0;  // list entries it holds for each kind
    uint32_t tile_queue = 0;          // mpm_tuning.tile_queue
    Shape shape = DEFAULT_SHAPE;      // requested launch shape (mpm_set_tuning)
    hipEvent_t ring_events[PARAM_RING] = {};
    bool ring_used[PARAM_RING] = {};
    int ring_next = 0;
    Control* h_ctl = nullptr;        // pinned readback
    HostStatic st;
    std::vector<Material> mats;
    std::map<std::tuple<uint32_t, uint32_t, uint32_t, uint32_t, uint32_t, uint32_t>, uint32_t> mat_index;
    float interaction[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    uint32_t n = 0;
    int cur = 0;                     // slot buffer holding the committed state
    uint32_t seq = 0;                // batch sequence number
    std::vector<Pending> pending;    // batches queued since the last settle
    uint32_t last_done_before = 0;   // done_before of the most recently launched batch
    std::string error = "no error";
    KernelTimes* times = nullptr;
    bool diag_on = false;
    bool shared_stream = false;        // the stream belongs to a batch
    struct mpm_batch* owner = nullptr; // the batch that owns this scene
    double* diag_saved = nullptr;  // diagnostics buffer while the reductions are off
    // config-2: bonded triangle bodies (from the records) and fracture bonds
    std::vector<uint32_t> body_start;  // particle index of each body's vertex 0
    std::vector<Bond> bonds;           // as set by mpm_set_fracture_bonds
    std::vector<uint32_t> aux_pid;     // particle of each aux entry
    DeviceArrays aux_mem;
    bool stepped = false;              // a step has been launched
    uint32_t removed = 0;              // particles removed by outlets (tombstones in their slots)
    uint64_t* d_small = nullptr;       // device: pick key and position, outlet counter
    float2* d_occ = nullptr;           // device: surface occupancy per node (first mpm_gather_occupancy)
};

namespace {

}  // namespace

// Independent scenes stepped together: every step launch covers the tiles
// of all scenes.  Each scene is a full context (own buffers, status, CFL
// bookkeeping and bucket growth) on the batch's stream.
struct mpm_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    std::vector<mpm_ctx*> scenes;
    std::vector<uint32_t> p2g_base, g2p_base;  // first tile of each scene
    uint32_t p2g_tiles = 0, g2p_tiles = 0;
    Shape shape = DEFAULT_SHAPE;      // requested launch shape (mpm_batch_set_tuning)
    Shape map_shape{};                // shape the tile maps were built for (zeros: none yet)
    uint32_t map_capacity = 0;        // entries the device tile maps hold
    SceneDesc* d_scenes = nullptr;   // device
    SceneDesc* h_scenes = nullptr;   // pinned staging
    uint32_t* d_p2g_map = nullptr;   // device: scene of every tile
    uint32_t* d_g2p_map = nullptr;
    uint32_t* d_queue = nullptr;     // device: tile queue counters and lists of the current tile maps
    uint32_t tile_queue = 0;         // mpm_tuning.tile_queue
    hipEvent_t staged = nullptr;     // the staging buffer's last copy
    bool staged_used = false;
    std::string error = "no error";
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

#define HIP_TRY(ctx, expr)                                     \
    do {                                                       \
        hipError_t _e = (expr);                                \
        if (_e != hipSuccess) return hip_fail(ctx, _e, #expr); \
    } while (0)

// validateMpmMlsGpuInput per-record checks at the float32 level
// (layout.ts:300-437): model code, finite state, positive volume/density,
// damage and plastic-volume ranges, domain bounds and a positive packed det.
// A removed tombstone (model 0xFFFFFFFF, left by an outlet and kept by a
// restored snapshot) is not checked: no kernel reads it.
int validate_records(mpm_ctx* c, const mpm_config& cfg, const uint32_t* rec, uint32_t n) {
    const float dom_w = (float)cfg.cells_x * cfg.cell_size;
    const float dom_h = (float)cfg.cells_y * cfg.cell_size;
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t* r = rec + 20ull * i;
        const uint32_t model = r[13];
        if (model == MODEL_REMOVED) {
            if (r[19] != 0u) return fail(c, MPM_EINVAL, "particles[%u] is removed but marked as a bonded vertex", i);
            continue;
        }
        if (model > 4u)
            return fail(c, MPM_EINVAL, "particles[%u].model code %u is not a supported material", i, model);
        if (r[19] > 3u)
            return fail(c, MPM_EINVAL, "particles[%u].bondedTriangleVertex must be 0, 1 or 2", i);
        for (int k = 0; k < 19; ++k) {
            if (k == 13) continue;
            if (!isfinite(u2f(r[k]))) return fail(c, MPM_EINVAL, "particles[%u] word %d must be finite", i, k);
        }
        const float x = u2f(r[0]), y = u2f(r[1]);
        if (x < 0.0f || x > dom_w || y < 0.0f || y > dom_h)
            return fail(c, MPM_EINVAL, "particles[%u].position is outside the float32 GPU grid domain", i);
        const float p1 = u2f(r[8]) * u2f(r[11]), p2 = u2f(r[9]) * u2f(r[10]);
        const float det = p1 - p2;
        if (!(det > 0.0f))
            return fail(c, MPM_EINVAL, "particles[%u].F must retain a finite positive determinant in float32", i);
        const float pv = u2f(r[12]), volume = u2f(r[17]), density = u2f(r[16]), param = u2f(r[18]);
        if (pv < 0.6f || pv > 20.0f)
            return fail(c, MPM_EINVAL, "particles[%u].plasticVolume must be between 0.6 and 20", i);
        if (!(volume > 0.0f)) return fail(c, MPM_EINVAL, "particles[%u].volume must be greater than zero", i);
        if ((model == 2u || model == 3u) && !(density > 0.0f))
            return fail(c, MPM_EINVAL, "particles[%u].material.density must be greater than zero", i);
        if ((model == 2u || model == 3u) && !(u2f(r[14]) > 0.0f))
            return fail(c, MPM_EINVAL, "particles[%u].material.lambda must be greater than zero", i);
        if (model == 0u && (param < 0.0f || param > 1.0f))
            return fail(c, MPM_EINVAL, "particles[%u].damage must be between zero and one", i);
        if (param < 0.0f) return fail(c, MPM_EINVAL, "particles[%u] material parameter must be non-negative", i);
    }
    return MPM_OK;
}

uint32_t material_row(mpm_ctx* c, const uint32_t* r) {
    auto key = std::make_tuple(r[13], r[14], r[15], r[16], r[17], r[18]);
    auto it = c->mat_index.find(key);
    if (it != c->mat_index.end()) return it->second;
    Material m{};
    m.model = r[13];
    m.lambda = u2f(r[14]);
    m.mu = u2f(r[15]);
    m.density = u2f(r[16]);
    m.volume = u2f(r[17]);
    m.param = u2f(r[18]);
    const bool liquid = m.model == MODEL_WATER || m.model == MODEL_OIL;
    m.mass = liquid ? m.density * m.volume : c->cfg.particle_mass;
    m.phase = m.model == MODEL_WATER ? 1u : (m.model == MODEL_OIL ? 2u : 0u);
    const uint32_t id = (uint32_t)c->mats.size();
    c->mats.push_back(m);
    c->mat_index.emplace(key, id);
    return id;
}

// Bucket of a record with the kernels' float32 stencil (shaders.ts:667-676).
bool host_bucket(const mpm_ctx* c, const uint32_t* r, uint32_t* bucket) {
    const float gx = u2f(r[0]) / c->cfg.cell_size, gy = u2f(r[1]) / c->cfg.cell_size;
    if (!isfinite(gx) || !isfinite(gy)) return false;
    const int bx = (int)floorf(gx - 0.5f), by = (int)floorf(gy - 0.5f);
    if (bx < -1 || by < -1 || bx + 2 > (int)c->cfg.cells_x + 1 || by + 2 > (int)c->cfg.cells_y + 1) return false;
    *bucket = (uint32_t)((by + 1) >> 2) * c->lc.bw + (uint32_t)((bx + 1) >> 2);
    return true;
}

int alloc_slots(mpm_ctx* c, uint32_t cap) {
    c->slot_mem.release();
    c->lc.cap = cap;
    const size_t slots = (size_t)c->lc.nb * cap;
    // the kernels index slots with 32-bit arithmetic
    if (slots > 0xffffffffull)
        return fail(c, MPM_EINVAL, "%zu bucket slots exceed the 2^32 - 1 the kernels address", slots);
    const bool general = scene_general(c->lc.scene_flags);
    hipError_t e;
    for (int b = 0; b < 2; ++b) {
        SlotBuf& s = c->lc.buf[b];
        s = SlotBuf{};
#define SALLOC(ptr, bytes)                                                                               \
    if ((e = c->slot_mem.alloc((void**)&(ptr), (bytes))) != hipSuccess)                                  \
        return fail(c, MPM_EDEVICE, "hipMalloc(%zu bytes) for bucket slots failed: %s", (size_t)(bytes), \
                    hipGetErrorString(e));
        SALLOC(s.pos, slots * sizeof(float2));
        SALLOC(s.vel, slots * sizeof(float2));
        SALLOC(s.C, slots * sizeof(float4));
        SALLOC(s.F, slots * sizeof(float4));
        SALLOC(s.pid, slots * sizeof(uint32_t));
        SALLOC(s.mat, slots * sizeof(uint32_t));
        SALLOC(s.cnt, (size_t)c->lc.nb * sizeof(uint32_t));
        if (general) {
            SALLOC(s.S, slots * sizeof(float4));
            SALLOC(s.plastic, slots * sizeof(float));
        }
#undef SALLOC
    }
    return MPM_OK;
}

int upload_materials(mpm_ctx* c) {
    if (c->mats.size() > c->mat_capacity) {
        uint32_t capm = std::max<uint32_t>(16u, c->mat_capacity);
        while (capm < c->mats.size()) capm *= 2;
        Material* d = nullptr;
        HIP_TRY(c, hipMalloc((void**)&d, capm * sizeof(Material)));
        if (c->d_mats) HIP_TRY(c, hipFree(c->d_mats));
        c->d_mats = d;
        c->mat_capacity = capm;
        c->lc.mats = d;
    }
    HIP_TRY(c, hipMemcpy(c->d_mats, c->mats.data(), c->mats.size() * sizeof(Material), hipMemcpyHostToDevice));
    c->lc.nmat = (uint32_t)c->mats.size();
    return MPM_OK;
}

// Write n records into the caller-order arrays at `offset`.
int write_home(mpm_ctx* c, const uint32_t* rec, const uint32_t* ids, uint32_t n, uint32_t offset) {
    std::vector<float2> pos(n), vel(n);
    std::vector<float4> C(n), F(n);
    std::vector<float> pl(n);
    std::vector<uint32_t> model(n), mat(n);
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t* r = rec + 20ull * i;
        pos[i] = make_float2(u2f(r[0]), u2f(r[1]));
        vel[i] = make_float2(u2f(r[2]), u2f(r[3]));
        C[i] = make_float4(u2f(r[4]), u2f(r[5]), u2f(r[6]), u2f(r[7]));
        F[i] = make_float4(u2f(r[8]), u2f(r[9]), u2f(r[10]), u2f(r[11]));
        pl[i] = u2f(r[12]);
        model[i] = r[13];
        mat[i] = r[13] == MODEL_REMOVED ? 0u : material_row(c, r);
        const uint32_t j = offset + i;
        c->st.model[j] = r[13];
        c->st.lambda[j] = u2f(r[14]);
        c->st.mu[j] = u2f(r[15]);
        c->st.density[j] = u2f(r[16]);
        c->st.volume[j] = u2f(r[17]);
        c->st.param[j] = u2f(r[18]);
        c->st.object_id[j] = ids ? ids[i] : 0u;
        c->st.vertex[j] = r[19];
        const uint32_t m = r[13];
        if (m == MODEL_REMOVED) {
            ++c->removed;
            continue;
        }
        c->lc.scene_flags |= m == 2u ? SCENE_WATER : (m == 3u ? SCENE_OIL : SCENE_SOLID);
        if ((m == 2u || m == 3u) && u2f(r[18]) > 0.0f) c->lc.scene_flags |= SCENE_TAIT;
    }
    const Home& h = c->lc.home;
    if (n) {
        HIP_TRY(c, hipMemcpy(h.pos + offset, pos.data(), n * sizeof(float2), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(h.vel + offset, vel.data(), n * sizeof(float2), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(h.C + offset, C.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(h.F + offset, F.data(), n * sizeof(float4), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(h.plastic + offset, pl.data(), n * sizeof(float), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(h.model + offset, model.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(h.mat + offset, mat.data(), n * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    return upload_materials(c);
}

// collectBondedBodyStarts (reference/solver.ts:1023-1062; layout.ts:452-495):
// triangles are three consecutive records with vertex codes 1, 2, 3 of one
// object and a solid material.
int collect_bodies(mpm_ctx* c, const uint32_t* rec, const uint32_t* ids, uint32_t n, std::vector<uint32_t>& starts) {
    starts.clear();
    for (uint32_t i = 0; i < n; ++i) {
        const uint32_t* r = rec + 20ull * i;
        if (r[19] == 0u) continue;
        if (r[19] != 1u)
            return fail(c, MPM_EINVAL, "particles[%u].bondedTriangleVertex must start each triangle at zero", i);
        if (i + 2 >= n) return fail(c, MPM_EINVAL, "bonded triangle at particle %u is incomplete", i);
        const uint32_t oid = ids ? ids[i] : 0u;
        for (uint32_t v = 0; v < 3; ++v) {
            const uint32_t* m = rec + 20ull * (i + v);
            if (m[19] != v + 1u || (ids ? ids[i + v] : 0u) != oid)
                return fail(c, MPM_EINVAL,
                            "bonded triangle at particle %u must contain contiguous vertices 0, 1, and 2 of one object", i);
            if (m[13] == MODEL_WATER || m[13] == MODEL_OIL)
                return fail(c, MPM_EINVAL, "bonded triangle vertices must use a solid material");
        }
        starts.push_back(i);
        i += 2;
    }
    if (starts.size() > 1024u)
        return fail(c, MPM_EINVAL, "at most 1024 bonded triangle bodies are supported per scene");
    return MPM_OK;
}

// Device tables of the aux particles (triangle vertices, then fracture
// endpoints in ascending index), the triangle field and the bonds.
int build_aux(mpm_ctx* c) {
    c->aux_mem.release();
    LaunchCtx& lc = c->lc;
    lc.aux = AuxData{};
    lc.scene_flags &= ~SCENE_AUX;
    const uint32_t nbodies = (uint32_t)c->body_start.size(), nbonds = (uint32_t)c->bonds.size();
    if (nbodies == 0 && nbonds == 0) return MPM_OK;
    const uint32_t capn = std::max<uint32_t>(c->cfg.capacity, 1u);
    std::vector<uint32_t> aux_of(capn, NO_AUX), pids;
    for (uint32_t b = 0; b < nbodies; ++b)
        for (uint32_t v = 0; v < 3; ++v) {
            aux_of[c->body_start[b] + v] = (uint32_t)pids.size();
            pids.push_back(c->body_start[b] + v);
        }
    const uint32_t frac0 = (uint32_t)pids.size();
    std::vector<uint32_t> ends;
    for (const Bond& bd : c->bonds) {
        ends.push_back(bd.first);
        ends.push_back(bd.second);
    }
    std::sort(ends.begin(), ends.end());
    ends.erase(std::unique(ends.begin(), ends.end()), ends.end());
    for (uint32_t e : ends) {
        aux_of[e] = (uint32_t)pids.size();
        pids.push_back(e);
    }
    const uint32_t naux = (uint32_t)pids.size();
    std::vector<uint32_t> off(naux + 1, 0u), inc(2 * nbonds);
    for (const Bond& bd : c->bonds) {
        off[aux_of[bd.first] + 1]++;
        off[aux_of[bd.second] + 1]++;
    }
    for (uint32_t a = 0; a < naux; ++a) off[a + 1] += off[a];
    std::vector<uint32_t> cur(off.begin(), off.end() - 1);
    for (uint32_t k = 0; k < nbonds; ++k) {
        inc[cur[aux_of[c->bonds[k].first]]++] = k;
        inc[cur[aux_of[c->bonds[k].second]]++] = k;
    }
    std::vector<float> damage(naux, 0.0f);
    for (uint32_t a = frac0; a < naux; ++a) damage[a] = c->st.param[pids[a]];
    std::vector<uint32_t> oid(std::max<uint32_t>(nbodies, 1u), 0u);
    for (uint32_t b = 0; b < nbodies; ++b) oid[b] = c->st.object_id[c->body_start[b]];
    AuxData A{};
    hipError_t e;
#define AALLOC(ptr, bytes)                                                                             \
    if ((e = c->aux_mem.alloc((void**)&(ptr), (bytes))) != hipSuccess)                                 \
        return fail(c, MPM_EDEVICE, "hipMalloc(%zu bytes) for the body tables failed: %s", (size_t)(bytes), \
                    hipGetErrorString(e));
    AALLOC(A.aux_of, capn * sizeof(uint32_t));
    AALLOC(A.aux_pid, naux * sizeof(uint32_t));
    AALLOC(A.slot[0], naux * sizeof(uint32_t));
    AALLOC(A.slot[1], naux * sizeof(uint32_t));
    AALLOC(A.damage[0], naux * sizeof(float));
    AALLOC(A.damage[1], naux * sizeof(float));
    AALLOC(A.body_oid, oid.size() * sizeof(uint32_t));
    AALLOC(A.field, (size_t)lc.node_count * sizeof(FieldNode));
    AALLOC(A.field_body, (size_t)lc.node_count * sizeof(uint32_t));
    AALLOC(A.bonds, std::max<uint32_t>(nbonds, 1u) * sizeof(Bond));
    AALLOC(A.bond_off, (naux + 1) * sizeof(uint32_t));
    AALLOC(A.bond_inc, std::max<uint32_t>(2 * nbonds, 1u) * sizeof(uint32_t));
#undef AALLOC
    HIP_TRY(c, hipMemcpy(A.aux_of, aux_of.data(), capn * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(A.aux_pid, pids.data(), naux * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(A.damage[0], damage.data(), naux * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(A.damage[1], damage.data(), naux * sizeof(float), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemcpy(A.body_oid, oid.data(), oid.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
    if (nbonds) {
        HIP_TRY(c, hipMemcpy(A.bonds, c->bonds.data(), nbonds * sizeof(Bond), hipMemcpyHostToDevice));
        HIP_TRY(c, hipMemcpy(A.bond_inc, inc.data(), 2 * nbonds * sizeof(uint32_t), hipMemcpyHostToDevice));
    }
    HIP_TRY(c, hipMemcpy(A.bond_off, off.data(), (naux + 1) * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIP_TRY(c, hipMemset(A.field_body, 0, (size_t)lc.node_count * sizeof(uint32_t)));
    A.nbodies = nbodies;
    A.naux = naux;
    A.nbonds = nbonds;
    A.frac_aux0 = frac0;
    lc.aux = A;
    lc.scene_flags |= SCENE_AUX | SCENE_GENERAL;
    c->aux_pid = pids;
    return MPM_OK;
}

// Bucket every live particle of the caller-order arrays into buffer cur,
// growing the bucket capacity until it fits.
int rebucket_all(mpm_ctx* c, uint32_t min_cap) {
    uint32_t cap = std::max(min_cap, c->lc.cap);
    for (;;) {
        if (cap > MAX_BUCKET_CAP)
            return fail(c, MPM_ECAPACITY, "particle density exceeds %u particles per 4x4-cell bucket",
                        MAX_BUCKET_CAP);
        if (cap != c->lc.cap || !c->lc.buf[0].pos) {
            int a = alloc_slots(c, cap);
            if (a) return a;
        }
        HIP_TRY(c, launch_clear_counts(c->lc, c->cur));
        // keep the status: a flag raised before the rebucket stays raised
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipMemcpy(c->h_ctl, c->d_ctl, sizeof(Control), hipMemcpyDeviceToHost));
        const uint32_t before = c->h_ctl->status[0];
        HIP_TRY(c, launch_bucket(c->lc, c->cur, 0, c->n));
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        HIP_TRY(c, hipMemcpy(c->h_ctl, c->d_ctl, sizeof(Control), hipMemcpyDeviceToHost));
        if (!(c->h_ctl->status[0] & ERR_INTERNAL_REBUCKET)) return MPM_OK;
        c->h_ctl->status[0] = before;
        HIP_TRY(c, hipMemcpy(&c->d_ctl->status[0], &before, sizeof(uint32_t), hipMemcpyHostToDevice));
        cap *= 2;
    }
}

int initial_cap(mpm_ctx* c, const uint32_t* rec, uint32_t n) {
    std::vector<uint32_t> occ(c->lc.nb, 0u);
    uint32_t mx = 0;
    for (uint32_t i = 0; i < n; ++i) {
        uint32_t b;
        if (rec[20ull * i + 13] == MODEL_REMOVED || !host_bucket(c, rec + 20ull * i, &b)) continue;
        mx = std::max(mx, ++occ[b]);
    }
    uint32_t cap = std::max(MIN_BUCKET_CAP, ((2 * mx + 15) / 16) * 16);
    return (int)std::min(cap, MAX_BUCKET_CAP);
}

// This context as one scene of a step launch with parameters p; the
// scene's tiles start at the given bases.
SceneDesc scene_desc(const mpm_ctx* c, const Params& p, uint32_t p2g_base, uint32_t g2p_base, uint32_t seq,
                     const Shape& shape) {
    const LaunchCtx& lc = c->lc;
    SceneDesc d{};
    d.P = p;
    d.ctl = lc.ctl;
    d.buf[0] = lc.buf[0];
    d.buf[1] = lc.buf[1];
    d.mats = lc.mats;
    d.grid = lc.grid;
    d.diag_grid = lc.diag ? lc.diag_grid : nullptr;
    d.diag_part = lc.diag ? lc.diag_part : nullptr;
    d.diag = lc.diag;
    d.nmat = lc.nmat;
    d.bw = lc.bw;
    d.bh = lc.bh;
    d.nb = lc.nb;
    d.cap = lc.cap;
    d.ptx = p2g_tiles_side(lc.node_w, shape.p2g_tile);
    d.pty = p2g_tiles_side(lc.node_h, shape.p2g_tile);
    d.gtx = (lc.bw + shape.g2p_tile - 1) / shape.g2p_tile;
    d.gty = (lc.bh + shape.g2p_tile - 1) / shape.g2p_tile;
    d.p2g_base = p2g_base;
    d.g2p_base = g2p_base;
    d.scene_flags = lc.scene_flags;
    d.keep_fixed = lc.keep_fixed;
    d.seq = seq;
    d.start_buf = (uint32_t)c->cur;
    d.aux = lc.aux;
    return d;
}

int settle(mpm_ctx* c);

// The default launch shape for `g2p_tiles8` 8x8-bucket G2P tiles on `device`.
Shape default_shape(uint32_t g2p_tiles8, int device) {
    int cus = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || cus <= 0)
        return DEFAULT_SHAPE;
    return g2p_tiles8 < (uint32_t)cus ? SMALL_GRID_SHAPE : DEFAULT_SHAPE;
}

// Shape a step of this context runs with: scenes with bodies or bonds run
// AUX_SHAPE's P2G tile and G2P threads with the requested G2P tile, and a
// P2G tile whose gather chunk is smaller than the bucket capacity falls back
// to the 28-node tile (a bucket must fit a chunk).
Shape effective_shape(const mpm_ctx* c, Shape s) {
    if (c->lc.scene_flags & SCENE_AUX) return {AUX_SHAPE.p2g_tile, s.g2p_tile == 4u ? 4u : 8u, AUX_SHAPE.g2p_threads};
    if (c->lc.cap > p2g_chunk_for(s.p2g_tile)) s.p2g_tile = 28u;
    return s;
}

int launch(mpm_ctx* c, const Params& p, uint32_t steps, bool measure, uint32_t done_before) {
    if (c->pending.size() >= MAX_PENDING) {
        // keep the replay list within the control block's history
        int s = settle(c);
        if (s) return s;
    }
    const int slot = c->ring_next;
    c->ring_next = (c->ring_next + 1) % PARAM_RING;
    if (c->ring_used[slot]) HIP_TRY(c, hipEventSynchronize(c->ring_events[slot]));
    const uint32_t seq = c->seq++;
    const Shape shape = effective_shape(c, c->shape);
    c->h_desc_ring[slot] = scene_desc(c, p, 0u, 0u, seq, shape);
    HIP_TRY(c, hipMemcpyAsync(c->d_scene, &c->h_desc_ring[slot], sizeof(SceneDesc), hipMemcpyHostToDevice,
                              c->stream));
    HIP_TRY(c, hipEventRecord(c->ring_events[slot], c->stream));
    c->ring_used[slot] = true;
    const SceneDesc& d = c->h_desc_ring[slot];
    StepBatch sb{};
    sb.stream = c->stream;
    sb.d_scenes = c->d_scene;
    sb.p2g_scene = c->d_p2g_map;
    sb.g2p_scene = c->d_g2p_map;
    sb.nscenes = 1;
    sb.p2g_tiles = p2g_tiles_for(c->lc.node_w, c->lc.node_h, shape.p2g_tile);
    sb.g2p_tiles = g2p_tiles_for(c->lc.bw, c->lc.bh, shape.g2p_tile);
    sb.max_slots = c->lc.nb * c->lc.cap;
    sb.max_p2g_tiles = sb.p2g_tiles;
    sb.shape = shape;
    sb.scene_flags = d.scene_flags;
    sb.fixed_point = c->lc.fixed_point;
    sb.diag = c->lc.diag != nullptr;
    sb.timing = c->times;
    sb.aux_bodies = c->lc.aux.nbodies;
    sb.aux_bonds = c->lc.aux.nbonds;
    sb.node_count = c->lc.node_count;
    if (!c->lc.grid.fixed && queue_wanted(c->tile_queue, c->lc.node_count, c->n, sb.p2g_tiles, c->device))
        sb.queue = {c->d_queue, {c->d_queue + 2 * QUEUE_SET_WORDS, c->d_queue + 2 * QUEUE_SET_WORDS + c->queue_p2g},
                    {sb.p2g_tiles, sb.g2p_tiles}};
    c->stepped = true;
    HIP_TRY(c, launch_steps(sb, steps, measure));
    c->pending.push_back({p, steps, measure, seq, done_before});
    c->last_done_before = done_before;
    c->cur = (c->cur + (int)(steps & 1u)) & 1;
    return MPM_OK;
}

// Wait for queued work; replay the microsteps a full bucket interrupted.
int settle(mpm_ctx* c) {
    HIP_TRY(c, hipSetDevice(c->device));
    for (;;) {
        HIP_TRY(c, hipStreamSynchronize(c->stream));
        if (c->times) c->times->resolve();
        HIP_TRY(c, hipMemcpy(c->h_ctl, c->d_ctl, sizeof(Control), hipMemcpyDeviceToHost));
        const uint32_t st = c->h_ctl->status[0];
        if (!(st & ERR_INTERNAL_REBUCKET)) {
            c->pending.clear();
            return MPM_OK;
        }
        // find the batch that ran out of slots; later ones did not run
        size_t k = 0;
        while (k < c->pending.size() &&
               c->h_ctl->history[c->pending[k].seq % CONTROL_HISTORY] >= c->pending[k].steps)
            ++k;
        std::vector<Pending> replay;
        if (k < c->pending.size()) {
            Pending first = c->pending[k];
            const uint32_t done = c->h_ctl->history[first.seq % CONTROL_HISTORY];
            first.steps -= done;
            first.done_before += done;
            replay.push_back(first);
            for (size_t q = k + 1; q < c->pending.size(); ++q) replay.push_back(c->pending[q]);
        }
        c->pending.clear();
        const uint32_t rest = st & ~ERR_INTERNAL_REBUCKET;
        HIP_TRY(c, hipMemcpy(&c->d_ctl->status[0], &rest, sizeof(uint32_t), hipMemcpyHostToDevice));
        // committed state -> caller-order arrays -> larger buckets
        HIP_TRY(c, launch_unbucket(c->lc, c->cur));
        int r = rebucket_all(c, c->lc.cap * 2);
        if (r) return r;
        if (rest != 0u) return MPM_OK;  // a real flag stops the replay
        for (const Pending& p : replay) {
            if (p.steps == 0) continue;
            r = launch(c, p.params, p.steps, p.measure, p.done_before);
            if (r) return r;
        }
    }
}

// Give a scene the FP64 stress cache (SCENE_GENERAL) after it has stepped:
// its committed state goes back to the caller-order arrays and is bucketed
// again into slot buffers that carry the cache.
int make_general(mpm_ctx* o) {
    int s = settle(o);
    if (s) return s;
    HIP_TRY(o, launch_unbucket(o->lc, o->cur));
    HIP_TRY(o, hipStreamSynchronize(o->stream));
    o->lc.scene_flags |= SCENE_GENERAL;
    int a = alloc_slots(o, o->lc.cap);
    if (a) return a;
    return rebucket_all(o, o->lc.cap);
}

}  // namespace

// RN(1/b) for the fixed-point kernel's division by particleMass and
// particleVolume (reciprocal plus one FMA correction, exact for divisors in
// [2^-20, 2^20]); 0 outside that range selects plain division.
static float reciprocal_for_division(float b) {
    if (!(b >= 0x1p-20f && b <= 0x1p20f)) return 0.0f;
    volatile float one = 1.0f;
    return one / b;
}

// Decode and check one 144-byte parameter block for context c (the checks
// stepBatch makes before encoding, solver.ts:1414-1441), and add the parts
// the host owns: drag uniform, diagnostics switch, exact reciprocals.
int prepare_params(mpm_ctx* c, const void* params144, Params& p) {
    if (!params144) return fail(c, MPM_EINVAL, "params must not be NULL");
    p = Params{};
    memcpy(&p, params144, MPM_PARAMS_BYTES);
    const LaunchCtx& lc = c->lc;
    if (p.cells_x != c->cfg.cells_x || p.cells_y != c->cfg.cells_y || p.node_w != lc.node_w || p.node_h != lc.node_h)
        return fail(c, MPM_EINVAL, "params grid does not match the solver grid");
    if (p.node_count != lc.node_count) return fail(c, MPM_EINVAL, "params nodeCount does not match the solver grid");
    if (p.particle_count != c->n)
        return fail(c, MPM_EINVAL, "params particleCount %u does not match the %u allocated particle slots",
                    p.particle_count, c->n);
    const uint32_t nbodies = (uint32_t)c->body_start.size();
    if (p.body_count != nbodies)
        return fail(c, MPM_EINVAL, "params bondedBodyCount %u does not match the %u bonded triangle bodies",
                    p.body_count, nbodies);
    if (nbodies) {
        if (!(p.bond_rest > 0.0f) || !isfinite(p.bond_rest))
            return fail(c, MPM_EINVAL, "bondRestLength is required for bonded triangle particles");
        if (!(p.bond_stiffness > 0.0f) || !(p.bond_stiffness <= 1.0f))
            return fail(c, MPM_EINVAL, "bondStiffness must be greater than zero and at most one");
        if (!(p.bond_ref_dt > 0.0f) || !isfinite(p.bond_ref_dt))
            return fail(c, MPM_EINVAL, "bondReferenceDt must be greater than zero");
        if (!(p.reach > 0.0f) || !isfinite(p.reach) || !(p.band > 0.0f))
            return fail(c, MPM_EINVAL, "params triangle contact band and reach are invalid");
    }
    if (p.fracture[1] != (uint32_t)c->bonds.size())
        return fail(c, MPM_EINVAL, "params fracture bond count %u does not match the %zu bonds", p.fracture[1],
                    c->bonds.size());
    if (!(p.dt > 0.0f) || !isfinite(p.dt)) return fail(c, MPM_EINVAL, "dt must be greater than zero");
    if (p.dx != c->cfg.cell_size) return fail(c, MPM_EINVAL, "params cellSize does not match the solver grid");
    if (p.particle_mass != c->cfg.particle_mass || p.particle_volume != c->cfg.particle_volume)
        return fail(c, MPM_EINVAL, "params particleMass/particleVolume do not match the solver configuration");
    if (!isfinite(p.gx) || !isfinite(p.gy)) return fail(c, MPM_EINVAL, "gravity must be finite");
    if (!isfinite(p.grad_scale) || !(p.grad_scale > 0.0f)) return fail(c, MPM_EINVAL, "params gradient scale is invalid");
    if (p.cfl_on >= 0.5f && (!isfinite(p.wave_bound) || !isfinite(p.vel_bound) || p.vel_bound < 0.0f))
        return fail(c, MPM_EINVAL, "CFL bounds must be finite and non-negative");
    p.diag_on = c->diag_on ? 1u : 0u;
    p.drag_pad[0] = p.drag_pad[1] = 0.0f;
    p.inv_mass_ref = reciprocal_for_division(p.particle_mass);
    p.inv_volume_ref = reciprocal_for_division(p.particle_volume);
    p.inv_dx = exact_reciprocal(p.dx);
    p.drag_target_x = c->interaction[0];
    p.drag_target_y = c->interaction[1];
    p.drag_radius = c->interaction[2];
    p.drag_stiffness = c->interaction[3];
    p.drag_damping = c->interaction[4];
    return MPM_OK;
}

// Release everything a context owns (its stream too unless a batch owns it).
static void destroy_ctx(mpm_ctx* c) {
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    delete c->times;
    c->slot_mem.release();
    c->mem.release();
    c->aux_mem.release();
    if (c->d_mats) (void)hipFree(c->d_mats);
    for (int k = 0; k < PARAM_RING; ++k)
        if (c->ring_events[k]) (void)hipEventDestroy(c->ring_events[k]);
    if (c->h_desc_ring) (void)hipHostFree(c->h_desc_ring);
    if (c->h_ctl) (void)hipHostFree(c->h_ctl);
    if (c->stream && !c->shared_stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

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
    std::vector<uint32_t> body_start;
    {
        int v = validate_records(nullptr, c, packed20, n);
        if (!v) v = collect_bodies(nullptr, packed20, object_ids, n, body_start);
        if (v) return v;
    }

    mpm_ctx* ctx = new (std::nothrow) mpm_ctx();
    if (!ctx) return fail(nullptr, MPM_EDEVICE, "out of host memory");
    ctx->cfg = c;
    ctx->cfg.capacity = capacity;
    auto bail = [&](int code) {
        g_create_error = ctx->error;
        mpm_destroy(ctx);
        return code;
    };
    LaunchCtx& lc = ctx->lc;
    lc.node_w = c.cells_x + 3;
    lc.node_h = c.cells_y + 3;
    lc.node_count = (uint32_t)nodes64;
    lc.cells_x = c.cells_x;
    lc.cells_y = c.cells_y;
    lc.cell_size = c.cell_size;
    lc.bw = (lc.node_w - 3) / BUCKET_CELLS + 1;  // stencil bases span storage cells [0, cells]
    lc.bh = (lc.node_h - 3) / BUCKET_CELLS + 1;
    lc.nb = lc.bw * lc.bh;
    try {
        HostStatic& s = ctx->st;
        for (auto* v : {&s.model, &s.object_id, &s.vertex}) v->assign(capacity ? capacity : 1, 0u);
        for (auto* v : {&s.lambda, &s.mu, &s.density, &s.volume, &s.param}) v->assign(capacity ? capacity : 1, 0.0f);
    } catch (...) {
        return bail(fail(ctx, MPM_EDEVICE, "out of host memory"));
    }
    hipError_t e;
    if (c.device >= 0 && (e = hipSetDevice(c.device)) != hipSuccess) return bail(hip_fail(ctx, e, "hipSetDevice"));
    if ((e = hipGetDevice(&ctx->device)) != hipSuccess) return bail(hip_fail(ctx, e, "hipGetDevice"));
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipStreamCreateWithFlags"));
    ctx->shape = default_shape(g2p_tiles_for(lc.bw, lc.bh, 8u), ctx->device);

    const size_t cap = capacity ? capacity : 1;
    const size_t nodes = lc.node_count;
    DeviceArrays& m = ctx->mem;
#define ALLOC(ptr, bytes)                                                                                    \
    do {                                                                                                     \
        if ((e = m.alloc((void**)&(ptr), (bytes))) != hipSuccess)                                            \
            return bail(fail(ctx, MPM_EDEVICE, "hipMalloc(%zu bytes) failed: %s", (size_t)(bytes),           \
                             hipGetErrorString(e)));                                                         \
    } while (0)
    ALLOC(lc.home.pos, cap * sizeof(float2));
    ALLOC(lc.home.vel, cap * sizeof(float2));
    ALLOC(lc.home.C, cap * sizeof(float4));
    ALLOC(lc.home.F, cap * sizeof(float4));
    ALLOC(lc.home.plastic, cap * sizeof(float));
    ALLOC(lc.home.model, cap * sizeof(uint32_t));
    ALLOC(lc.home.mat, cap * sizeof(uint32_t));
    ALLOC(lc.grid.v_solid, nodes * sizeof(float2));
    ALLOC(lc.grid.v_water, nodes * sizeof(float2));
    ALLOC(lc.grid.v_oil, nodes * sizeof(float2));
    ALLOC(lc.grid.liquid, nodes * sizeof(float2));
    const bool keep_fixed = c.p2g_mode == MPM_P2G_FIXED_POINT && (c.flags & MPM_FLAG_KEEP_FIXED_GRID);
    if (keep_fixed) ALLOC(lc.grid.fixed, nodes * 11 * sizeof(int32_t));
    ALLOC(ctx->d_ctl, sizeof(Control));
    ALLOC(ctx->d_scene, sizeof(SceneDesc));
    ALLOC(ctx->d_small, 4 * sizeof(uint64_t));
    // tile maps sized for the most tiles any shape makes
    const uint32_t max_p2g = p2g_tiles_for(lc.node_w, lc.node_h, MIN_P2G_TILE);
    const uint32_t max_g2p = g2p_tiles_for(lc.bw, lc.bh, 4u);
    ALLOC(ctx->d_p2g_map, sizeof(uint32_t) * max_p2g);
    ALLOC(ctx->d_g2p_map, sizeof(uint32_t) * max_g2p);
    ALLOC(ctx->d_queue, sizeof(uint32_t) * (2 * QUEUE_SET_WORDS + max_p2g + max_g2p));
    ctx->queue_p2g = max_p2g;
    ctx->queue_g2p = max_g2p;
#undef ALLOC
    if ((e = hipMemset(ctx->d_p2g_map, 0, sizeof(uint32_t) * max_p2g)) != hipSuccess ||
        (e = hipMemset(ctx->d_g2p_map, 0, sizeof(uint32_t) * max_g2p)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipMemset(tile maps)"));
    if ((e = hipHostMalloc((void**)&ctx->h_desc_ring, sizeof(SceneDesc) * PARAM_RING, hipHostMallocDefault)) !=
        hipSuccess)
        return bail(hip_fail(ctx, e, "hipHostMalloc(descriptors)"));
    if ((e = hipHostMalloc((void**)&ctx->h_ctl, sizeof(Control), hipHostMallocDefault)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipHostMalloc(control)"));
    for (int k = 0; k < PARAM_RING; ++k)
        if ((e = hipEventCreateWithFlags(&ctx->ring_events[k], hipEventDisableTiming)) != hipSuccess)
            return bail(hip_fail(ctx, e, "hipEventCreate"));
    for (void* p : {(void*)lc.grid.v_solid, (void*)lc.grid.v_water, (void*)lc.grid.v_oil, (void*)lc.grid.liquid})
        if ((e = hipMemset(p, 0, nodes * sizeof(float2))) != hipSuccess) return bail(hip_fail(ctx, e, "hipMemset"));
    if (keep_fixed && (e = hipMemset(lc.grid.fixed, 0, nodes * 11 * sizeof(int32_t))) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipMemset(fixed grid)"));
    Control init{};
    init.status[1] = 0xffffffffu;  // solver.ts:889-893
    init.deferred_index[0] = init.deferred_index[1] = 0xffffffffu;
    init.fail_step = 0xffffffffu;
    *ctx->h_ctl = init;
    if ((e = hipMemcpy(ctx->d_ctl, ctx->h_ctl, sizeof(Control), hipMemcpyHostToDevice)) != hipSuccess)
        return bail(hip_fail(ctx, e, "hipMemcpy(control)"));

    lc.stream = ctx->stream;
    lc.ctl = ctx->d_ctl;
    lc.fixed_point = c.p2g_mode == MPM_P2G_FIXED_POINT ? 1u : 0u;
    lc.keep_fixed = keep_fixed ? 1u : 0u;
    lc.scene_flags = 0u;
    ctx->cur = 0;
    int w = write_home(ctx, packed20, object_ids, n, 0);
    if (w) return bail(w);
    ctx->n = n;
    lc.n = n;
    const uint32_t cap0 = c.bucket_slots ? std::min<uint32_t>(c.bucket_slots, MAX_BUCKET_CAP)
                                         : (uint32_t)initial_cap(ctx, packed20, n);
    {
        ctx->body_start = body_start;
        int b = build_aux(ctx);
        if (b) return bail(b);
    }
    int a = alloc_slots(ctx, cap0);
    if (a) return bail(a);
    int r = rebucket_all(ctx, cap0);
    if (r) return bail(r);
    *out = ctx;
    return MPM_OK;
}


int mpm_step_batch(mpm_ctx* c, const void* params144, uint32_t step_count, uint32_t measure_speed) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (step_count < 1) return fail(c, MPM_EINVAL, "stepCount must be a positive safe integer.");
    Params p{};
    const int v = prepare_params(c, params144, p);
    if (v) return v;
    HIP_TRY(c, hipSetDevice(c->device));
    return launch(c, p, step_count, measure_speed != 0u || p.cfl_on >= 0.5f, 0u);
}

int mpm_read_compact(mpm_ctx* c, float* max_speed, uint32_t* completed, uint32_t status[8]) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    int s = settle(c);
    if (s) return s;
    if (max_speed) *max_speed = u2f(c->h_ctl->max_speed_bits);
    if (completed) *completed = c->h_ctl->completed + c->last_done_before;
    if (status) memcpy(status, c->h_ctl->status, sizeof(uint32_t) * 8);
    return MPM_OK;
}

int mpm_read_status(mpm_ctx* c, uint32_t status[8]) { return mpm_read_compact(c, nullptr, nullptr, status); }

int mpm_clear_status(mpm_ctx* c) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    int s = settle(c);
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
    int s = settle(c);
    if (s) return s;
    HIP_TRY(c, launch_unbucket(c->lc, c->cur));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const uint32_t total = indices ? c->n : n;
    std::vector<float2> pos(total), vel(total);
    std::vector<float4> C(total), F(total);
    std::vector<float> pl(total);
    const Home& h = c->lc.home;
    if (total > 0) {
        HIP_TRY(c, hipMemcpy(pos.data(), h.pos, total * sizeof(float2), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(vel.data(), h.vel, total * sizeof(float2), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(C.data(), h.C, total * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(F.data(), h.F, total * sizeof(float4), hipMemcpyDeviceToHost));
        HIP_TRY(c, hipMemcpy(pl.data(), h.plastic, total * sizeof(float), hipMemcpyDeviceToHost));
    }
    const HostStatic& st = c->st;
    std::vector<float> damage;  // fracture endpoints: the committed damage
    if (c->lc.aux.naux) {
        damage.resize(c->lc.aux.naux);
        HIP_TRY(c, hipMemcpy(damage.data(), c->lc.aux.damage[c->cur], damage.size() * sizeof(float),
                             hipMemcpyDeviceToHost));
    }
    std::vector<float> param_now;
    if (!damage.empty()) {
        param_now.assign(st.param.begin(), st.param.begin() + c->n);
        for (uint32_t a = c->lc.aux.frac_aux0; a < c->lc.aux.naux; ++a) param_now[c->aux_pid[a]] = damage[a];
    }
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t i = indices ? indices[k] : k;
        uint32_t* r = out20 + 20ull * k;
        r[0] = f2u(pos[i].x); r[1] = f2u(pos[i].y); r[2] = f2u(vel[i].x); r[3] = f2u(vel[i].y);
        r[4] = f2u(C[i].x); r[5] = f2u(C[i].y); r[6] = f2u(C[i].z); r[7] = f2u(C[i].w);
        r[8] = f2u(F[i].x); r[9] = f2u(F[i].y); r[10] = f2u(F[i].z); r[11] = f2u(F[i].w);
        r[12] = f2u(pl[i]); r[13] = st.model[i]; r[14] = f2u(st.lambda[i]); r[15] = f2u(st.mu[i]);
        r[16] = f2u(st.density[i]); r[17] = f2u(st.volume[i]);
        r[18] = f2u(param_now.empty() ? st.param[i] : param_now[i]);
        r[19] = st.vertex[i];
        if (object_ids) object_ids[k] = st.object_id[i];
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
    for (uint32_t i = 0; i < n; ++i)
        if (packed20[20ull * i + 19] != 0u)
            return fail(c, MPM_EINVAL, "appending bonded triangle vertices is not supported");
    int s = settle(c);
    if (s) return s;
    // committed state -> caller-order arrays, append, bucket everything again
    HIP_TRY(c, launch_unbucket(c->lc, c->cur));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    const bool was_general = scene_general(c->lc.scene_flags);
    int w = write_home(c, packed20, object_ids, n, c->n);
    if (w) return w;
    c->n += n;
    c->lc.n = c->n;
    if (scene_general(c->lc.scene_flags) != was_general) {
        int a = alloc_slots(c, c->lc.cap);
        if (a) return a;
    }
    int r = rebucket_all(c, c->lc.cap);
    if (r || !c->owner || was_general || !scene_general(c->lc.scene_flags)) return r;
    // One kernel variant serves a batch's launch, so when this scene now
    // needs the stress cache every scene of its batch carries it, as
    // mpm_batch_create arranges for scenes that need it from the start.
    for (mpm_ctx* o : c->owner->scenes) {
        if (o == c || scene_general(o->lc.scene_flags)) continue;
        r = make_general(o);
        if (r) return fail(c, r, "batch scene upgrade: %s", o->error.c_str());
    }
    return MPM_OK;
}

// validateMpmSolidBondGraph (reference/solid-bonds.ts:11-66) on the packed
// 24-byte bonds (layout.ts:527-548), then the device tables.
int mpm_set_fracture_bonds(mpm_ctx* c, const void* bonds24, uint32_t count) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (c->owner) return fail(c, MPM_EINVAL, "fracture bonds cannot be set on a scene of a batch");
    if (count > 0 && !bonds24) return fail(c, MPM_EINVAL, "bonds must not be NULL");
    if (c->stepped) return fail(c, MPM_ESTATE, "fracture bonds must be set before the first step");
    std::vector<Bond> bonds(count);
    if (count) memcpy(bonds.data(), bonds24, count * sizeof(Bond));
    std::vector<std::pair<uint32_t, uint32_t>> pairs;
    for (uint32_t k = 0; k < count; ++k) {
        const Bond& b = bonds[k];
        if (b.first >= c->n || b.second >= c->n)
            return fail(c, MPM_EINVAL, "fractureBonds[%u] endpoints must identify existing particles", k);
        if (b.first == b.second) return fail(c, MPM_EINVAL, "fractureBonds[%u] must connect two distinct particles", k);
        if (!(b.rest > 0.0f) || !isfinite(b.rest))
            return fail(c, MPM_EINVAL, "fractureBonds[%u].restLength must be positive and finite", k);
        if (!(b.failure > 0.0f) || !isfinite(b.failure))
            return fail(c, MPM_EINVAL, "fractureBonds[%u].failureStrain must be positive and finite", k);
        if (!(b.damage >= 0.0f && b.damage <= 1.0f))
            return fail(c, MPM_EINVAL, "fractureBonds[%u].damage must be between zero and one", k);
        if (b.broken > 1u) return fail(c, MPM_EINVAL, "fractureBonds[%u].broken must be 0 or 1", k);
        if (c->st.object_id[b.first] != c->st.object_id[b.second])
            return fail(c, MPM_EINVAL, "fractureBonds[%u] must connect particles of one object", k);
        if (c->st.vertex[b.first] != 0u || c->st.vertex[b.second] != 0u)
            return fail(c, MPM_EINVAL, "fractureBonds[%u] cannot reuse bonded triangle vertices", k);
        if (c->st.model[b.first] != MODEL_NEO_HOOKEAN || c->st.model[b.second] != MODEL_NEO_HOOKEAN)
            return fail(c, MPM_EINVAL, "fractureBonds[%u] endpoints must be neo-hookean solids", k);
        pairs.emplace_back(std::min(b.first, b.second), std::max(b.first, b.second));
    }
    std::vector<std::pair<uint32_t, uint32_t>> sorted = pairs;
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
        return fail(c, MPM_EINVAL, "fractureBonds duplicates an existing particle pair");
    int st = settle(c);
    if (st) return st;
    c->bonds = bonds;
    int a = build_aux(c);
    if (a) return a;
    if (!scene_general(c->lc.scene_flags) || !c->lc.buf[0].S) {
        a = alloc_slots(c, c->lc.cap);
        if (a) return a;
    }
    return rebucket_all(c, c->lc.cap);
}

int mpm_read_fracture_bonds(mpm_ctx* c, void* out24, uint32_t count) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (count != (uint32_t)c->bonds.size())
        return fail(c, MPM_EINVAL, "the context holds %zu fracture bonds, not %u", c->bonds.size(), count);
    if (count == 0) return MPM_OK;
    if (!out24) return fail(c, MPM_EINVAL, "output must not be NULL");
    int s = settle(c);
    if (s) return s;
    HIP_TRY(c, hipMemcpy(out24, c->lc.aux.bonds, count * sizeof(Bond), hipMemcpyDeviceToHost));
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

// removeParticlesInOutlet (solver.ts:2280-2386; shaders.ts:3041-3065).
int mpm_remove_in_outlet(mpm_ctx* c, const float bounds[4], uint32_t* removed) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!bounds) return fail(c, MPM_EINVAL, "outlet bounds must not be NULL");
    for (int k = 0; k < 4; ++k)
        if (!isfinite(bounds[k])) return fail(c, MPM_EINVAL, "outlet bounds must be finite");
    if (bounds[0] > bounds[2] || bounds[1] > bounds[3])
        return fail(c, MPM_EINVAL, "outlet minimum coordinates must not exceed the maximum.");
    if (removed) *removed = 0;
    if (c->n == 0) return MPM_OK;
    int s = settle(c);
    if (s) return s;
    uint32_t* d_count = reinterpret_cast<uint32_t*>(c->d_small + 2);
    HIP_TRY(c, launch_unbucket(c->lc, c->cur));
    HIP_TRY(c, hipMemsetAsync(d_count, 0, sizeof(uint32_t), c->stream));
    HIP_TRY(c, launch_outlet(c->lc, make_float4(bounds[0], bounds[1], bounds[2], bounds[3]), d_count));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    uint32_t count = 0;
    HIP_TRY(c, hipMemcpy(&count, d_count, sizeof count, hipMemcpyDeviceToHost));
    if (count) {
        HIP_TRY(c, hipMemcpy(c->st.model.data(), c->lc.home.model, (size_t)c->n * sizeof(uint32_t),
                             hipMemcpyDeviceToHost));
        c->removed += count;
        int r = rebucket_all(c, c->lc.cap);
        if (r) return r;
    }
    if (removed) *removed = count;
    return MPM_OK;
}

// pick (solver.ts:2733-2823; pickParticle, shaders.ts:3082-3118).
int mpm_pick(mpm_ctx* c, float x, float y, float radius, uint32_t out[4]) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!out) return fail(c, MPM_EINVAL, "pick result must not be NULL");
    if (!isfinite(x) || !isfinite(y) || !isfinite(radius) || radius < 0.0f)
        return fail(c, MPM_EINVAL, "pick point and radius must be finite, the radius non-negative");
    out[0] = 0xffffffffu;
    out[1] = out[2] = out[3] = 0u;
    int s = settle(c);
    if (s) return s;
    unsigned long long* d_best = reinterpret_cast<unsigned long long*>(c->d_small);
    float2* d_where = reinterpret_cast<float2*>(c->d_small + 1);
    HIP_TRY(c, hipMemsetAsync(d_best, 0xff, sizeof(unsigned long long), c->stream));
    HIP_TRY(c, launch_pick(c->lc, c->cur, x, y, radius * radius, d_best, d_where));
    uint64_t h[2] = {0, 0};
    HIP_TRY(c, hipMemcpyAsync(h, c->d_small, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    if (h[0] == ~0ull) return MPM_OK;
    const uint32_t index = (uint32_t)h[0];
    if (index >= c->n) return fail(c, MPM_EDEVICE, "pick returned a particle index outside the active range");
    out[0] = index;
    out[1] = (uint32_t)h[1];
    out[2] = (uint32_t)(h[1] >> 32);
    out[3] = c->st.object_id[index];
    return MPM_OK;
}

// clone (solver.ts:2852-2948; cloneObject, shaders.ts:3120-3147).
int mpm_clone(mpm_ctx* c, uint32_t start, uint32_t count, float offset_x, float offset_y, uint32_t object_id) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (count == 0 || start >= c->n || count > c->n - start)
        return fail(c, MPM_EINVAL, "clone range [%u, %u + %u) must lie inside the %u particle slots", start, start,
                    count, c->n);
    if (!isfinite(offset_x) || !isfinite(offset_y)) return fail(c, MPM_EINVAL, "clone offset must be finite");
    if (c->removed) return fail(c, MPM_ESTATE, "objects cannot be cloned after outlet removal");
    for (const Bond& b : c->bonds)
        if ((b.first >= start && b.first < start + count) || (b.second >= start && b.second < start + count))
            return fail(c, MPM_ESTATE, "objects with fracture bonds cannot be cloned");
    if ((uint64_t)c->n + count > c->cfg.capacity)
        return fail(c, MPM_ECAPACITY, "cloning %u particles exceeds the capacity of %u", count, c->cfg.capacity);
    std::vector<uint32_t> new_bodies;
    for (uint32_t b : c->body_start) {
        const bool first_in = b >= start && b < start + count, last_in = b + 2 >= start && b + 2 < start + count;
        if (first_in != last_in) return fail(c, MPM_EINVAL, "the clone range must hold whole bonded triangles");
        if (first_in) new_bodies.push_back(b - start + c->n);
    }
    if (c->body_start.size() + new_bodies.size() > 1024u)
        return fail(c, MPM_ECAPACITY, "at most 1024 bonded triangle bodies are supported per scene");
    int s = settle(c);
    if (s) return s;
    HIP_TRY(c, launch_unbucket(c->lc, c->cur));
    const float dom_w = (float)c->cfg.cells_x * c->cfg.cell_size, dom_h = (float)c->cfg.cells_y * c->cfg.cell_size;
    HIP_TRY(c, launch_clone(c->lc, start, count, c->n, offset_x, offset_y, dom_w, dom_h));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HostStatic& st = c->st;
    for (uint32_t k = 0; k < count; ++k) {
        const uint32_t from = start + k, to = c->n + k;
        st.model[to] = st.model[from];
        st.lambda[to] = st.lambda[from];
        st.mu[to] = st.mu[from];
        st.density[to] = st.density[from];
        st.volume[to] = st.volume[from];
        st.param[to] = st.param[from];
        st.vertex[to] = st.vertex[from];
        st.object_id[to] = object_id;
    }
    c->n += count;
    c->lc.n = c->n;
    if (!new_bodies.empty()) {
        c->body_start.insert(c->body_start.end(), new_bodies.begin(), new_bodies.end());
        int b = build_aux(c);
        if (b) return b;
    }
    return rebucket_all(c, c->lc.cap);
}

}  // extern "C"

namespace {

// The surface occupancy gather of the committed state into dst (node_count
// float2 of device memory), then copied to the caller's out.  Every node of
// the grid is the interior of exactly one tile, so each launch writes all of
// dst and nothing in it needs clearing first.
int run_occupancy(mpm_ctx* c, float2* dst, float* out, size_t floats, KernelTimes* times) {
    // the grid of this context as a parameter block (no step: dt 0, no gravity)
    Params p{};
    p.cells_x = c->cfg.cells_x;
    p.cells_y = c->cfg.cells_y;
    p.node_w = c->lc.node_w;
    p.node_h = c->lc.node_h;
    p.particle_count = c->n;
    p.node_count = c->lc.node_count;
    p.dx = c->cfg.cell_size;
    p.inv_dx = exact_reciprocal(p.dx);
    p.particle_mass = c->cfg.particle_mass;
    p.particle_volume = c->cfg.particle_volume;
    p.grad_scale = 4.0f / (p.dx * p.dx);
    p.dom_w = (float)c->cfg.cells_x * c->cfg.cell_size;
    p.dom_h = (float)c->cfg.cells_y * c->cfg.cell_size;
    const Shape shape = {28u, 8u, 256u};
    SceneDesc d = scene_desc(c, p, 0u, 0u, c->seq, shape);
    d.grid.occupancy = dst;
    d.diag_grid = nullptr;
    HIP_TRY(c, hipMemcpy(c->d_scene, &d, sizeof d, hipMemcpyHostToDevice));
    HIP_TRY(c, launch_occupancy(c->stream, c->d_scene, p2g_tiles_for(c->lc.node_w, c->lc.node_h, 28u),
                                scene_general(c->lc.scene_flags), times));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(out, dst, floats * sizeof(float), hipMemcpyDeviceToHost));
    return MPM_OK;
}

// Scratch of the allocation-free occupancy gather: one float2 per grid node,
// MPM_STATIC_SCRATCH_NODES of them (8 MiB).  Each GPU the library runs on
// holds its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in; the device address comes from hipGetSymbolAddress,
// once per device, and only that address is handed to kernels and copies.
static __device__ __align__(256) float2 scratch[MPM_STATIC_SCRATCH_NODES];

constexpr int STATIC_SCRATCH_DEVICES = 64;

struct StaticScratch {
    std::mutex lock;        // held for a whole call: calls on one GPU share the array
    float2* ptr = nullptr;  // device address of this device's copy
};
StaticScratch g_static_scratch[STATIC_SCRATCH_DEVICES];

}  // namespace

extern "C" {

// gatherSurfaceOccupancy (shaders.ts:710-806) of the committed state.
int mpm_gather_occupancy(mpm_ctx* c, float* out, size_t floats) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!out || floats != (size_t)c->lc.node_count * 2)
        return fail(c, MPM_EINVAL, "output must hold node_count*2 floats");
    int s = settle(c);
    if (s) return s;
    if (c->h_ctl->status[0] != 0u)
        return fail(c, MPM_ESTATE, "the numerical status is set: clear it before gathering the surface occupancy");
    if (!c->d_occ) {
        hipError_t e = c->mem.alloc((void**)&c->d_occ, (size_t)c->lc.node_count * sizeof(float2));
        if (e != hipSuccess) return hip_fail(c, e, "hipMalloc(occupancy)");
    }
    return run_occupancy(c, c->d_occ, out, floats, c->times);
}

// mpm_gather_occupancy without any allocation: the per-node scratch is this
// device's copy of the static array.  Queued steps are waited for, but a
// batch that ran out of bucket slots (whose replay reallocates the buckets)
// is refused instead of replayed.  Kernel timing is not recorded.
int mpm_gather_occupancy_static(mpm_ctx* c, float* out, size_t floats) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!out || floats != (size_t)c->lc.node_count * 2)
        return fail(c, MPM_EINVAL, "output must hold node_count*2 floats");
    if (c->lc.node_count > MPM_STATIC_SCRATCH_NODES)
        return fail(c, MPM_EINVAL, "the grid has %u nodes; the static scratch holds at most %u", c->lc.node_count,
                    (uint32_t)MPM_STATIC_SCRATCH_NODES);
    if (c->device < 0 || c->device >= STATIC_SCRATCH_DEVICES)
        return fail(c, MPM_EDEVICE, "device %d is outside the %d devices the static scratch serves", c->device,
                    STATIC_SCRATCH_DEVICES);
    HIP_TRY(c, hipSetDevice(c->device));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    HIP_TRY(c, hipMemcpy(c->h_ctl, c->d_ctl, sizeof(Control), hipMemcpyDeviceToHost));
    if (c->h_ctl->status[0] & ERR_INTERNAL_REBUCKET)
        return fail(c, MPM_ESTATE, "queued steps ran out of bucket slots and must be replayed with larger buckets, "
                                   "which allocates: read the status or particles first");
    c->pending.clear();
    if (c->h_ctl->status[0] != 0u)
        return fail(c, MPM_ESTATE, "the numerical status is set: clear it before gathering the surface occupancy");
    StaticScratch& st = g_static_scratch[c->device];
    std::lock_guard<std::mutex> hold(st.lock);
    if (!st.ptr) {
        void* ptr = nullptr;
        HIP_TRY(c, hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail(c, MPM_EDEVICE, "hipGetSymbolAddress(scratch) returned no address");
        st.ptr = static_cast<float2*>(ptr);
    }
    return run_occupancy(c, st.ptr, out, floats, nullptr);
}

int mpm_get_counts(mpm_ctx* c, uint32_t* particle_count, uint32_t* capacity, uint32_t* node_count) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (particle_count) *particle_count = c->n;
    if (capacity) *capacity = c->cfg.capacity;
    if (node_count) *node_count = c->lc.node_count;
    return MPM_OK;
}

int mpm_read_fixed_grid(mpm_ctx* c, int32_t* out, size_t words) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!c->lc.grid.fixed) return fail(c, MPM_ESTATE, "fixed-point grid retention is not enabled on this solver");
    if (!out || words != (size_t)c->lc.node_count * 11) return fail(c, MPM_EINVAL, "output must hold node_count*11 words");
    int s = settle(c);
    if (s) return s;
    HIP_TRY(c, hipMemcpy(out, c->lc.grid.fixed, words * sizeof(int32_t), hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_read_grid(mpm_ctx* c, float* out, size_t floats) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!out || floats != (size_t)c->lc.node_count * 8) return fail(c, MPM_EINVAL, "output must hold node_count*8 floats");
    int s = settle(c);
    if (s) return s;
    std::vector<float2> a(c->lc.node_count);
    const float2* src[4] = {c->lc.grid.v_solid, c->lc.grid.v_water, c->lc.grid.v_oil, c->lc.grid.liquid};
    for (int k = 0; k < 4; ++k) {
        HIP_TRY(c, hipMemcpy(a.data(), src[k], a.size() * sizeof(float2), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < c->lc.node_count; ++i) {
            out[8ull * i + 2 * k] = a[i].x;
            out[8ull * i + 2 * k + 1] = a[i].y;
        }
    }
    return MPM_OK;
}

int mpm_read_bins(mpm_ctx* c, uint32_t* keys, uint32_t n, uint32_t* bucket_counts, size_t count_words) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (n > c->n) return fail(c, MPM_EINVAL, "cannot read %u keys from %u slots", n, c->n);
    if (bucket_counts && count_words != (size_t)c->lc.nb)
        return fail(c, MPM_EINVAL, "bucket_counts must hold %u words", c->lc.nb);
    int s = settle(c);
    if (s) return s;
    HIP_TRY(c, launch_unbucket(c->lc, c->cur));
    if (keys && n) {
        uint32_t* d = nullptr;
        HIP_TRY(c, hipMalloc((void**)&d, (size_t)c->n * sizeof(uint32_t)));
        hipError_t e = launch_keys(c->lc, d);
        if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
        if (e == hipSuccess) e = hipMemcpy(keys, d, n * sizeof(uint32_t), hipMemcpyDeviceToHost);
        (void)hipFree(d);
        if (e != hipSuccess) return hip_fail(c, e, "keys");
    }
    if (bucket_counts)
        HIP_TRY(c, hipMemcpy(bucket_counts, c->lc.buf[c->cur].cnt, count_words * sizeof(uint32_t),
                             hipMemcpyDeviceToHost));
    return MPM_OK;
}

int mpm_get_buckets(mpm_ctx* c, uint32_t* buckets_x, uint32_t* buckets_y, uint32_t* slots_per_bucket) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (buckets_x) *buckets_x = c->lc.bw;
    if (buckets_y) *buckets_y = c->lc.bh;
    if (slots_per_bucket) *slots_per_bucket = c->lc.cap;
    return MPM_OK;
}

}  // extern "C"

namespace {

Shape batch_shape(const mpm_batch* b);

bool tuning_valid(const mpm_tuning* t, std::string& why) {
    const Shape s = {t->p2g_tile, t->g2p_tile, t->g2p_threads};
    if (!shape_supported(s)) {
        why = "unsupported launch shape: p2g_tile " + std::to_string(t->p2g_tile) + ", g2p_tile " +
              std::to_string(t->g2p_tile) + ", g2p_threads " + std::to_string(t->g2p_threads);
        return false;
    }
    if (t->tile_queue > 2u) {
        why = "tile_queue must be 0 (automatic), 1 (on) or 2 (off)";
        return false;
    }
    const uint32_t cap = t->bucket_capacity;
    if (cap != 0u && (cap % 16u != 0u || cap < MIN_BUCKET_CAP || cap > MAX_BUCKET_CAP)) {
        why = "bucket_capacity must be 0 or a multiple of 16 in [" + std::to_string(MIN_BUCKET_CAP) + ", " +
              std::to_string(MAX_BUCKET_CAP) + "]";
        return false;
    }
    return true;
}

// Re-bucket the committed state with `cap` slots per bucket (doubling while a
// bucket overflows).
int set_bucket_capacity(mpm_ctx* c, uint32_t cap) {
    if (cap == 0u || cap == c->lc.cap) return MPM_OK;
    int s = settle(c);
    if (s) return s;
    HIP_TRY(c, launch_unbucket(c->lc, c->cur));
    HIP_TRY(c, hipStreamSynchronize(c->stream));
    int a = alloc_slots(c, cap);
    if (a) return a;
    return rebucket_all(c, cap);
}

void tuning_of(const Shape& s, uint32_t cap, uint32_t queue, mpm_tuning* out) {
    out->p2g_tile = s.p2g_tile;
    out->g2p_tile = s.g2p_tile;
    out->g2p_threads = s.g2p_threads;
    out->bucket_capacity = cap;
    out->tile_queue = queue;
}

}  // namespace

extern "C" {

int mpm_set_tuning(mpm_ctx* c, const mpm_tuning* t) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!t) return fail(c, MPM_EINVAL, "tuning must not be NULL");
    std::string why;
    if (!tuning_valid(t, why)) return fail(c, MPM_EINVAL, "%s", why.c_str());
    const Shape s = {t->p2g_tile, t->g2p_tile, t->g2p_threads};
    if (c->owner) {
        const Shape& bs = c->owner->shape;
        if (bs.p2g_tile != s.p2g_tile || bs.g2p_tile != s.g2p_tile || bs.g2p_threads != s.g2p_threads)
            return fail(c, MPM_EINVAL, "the scenes of a batch share its launch shape: use mpm_batch_set_tuning");
    }
    if (hipSetDevice(c->device) != hipSuccess) return fail(c, MPM_EDEVICE, "hipSetDevice failed");
    const int r = set_bucket_capacity(c, t->bucket_capacity);
    if (r) return r;
    c->shape = s;
    c->tile_queue = t->tile_queue;
    return MPM_OK;
}

int mpm_get_tuning(mpm_ctx* c, mpm_tuning* out) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    if (!out) return fail(c, MPM_EINVAL, "out must not be NULL");
    tuning_of(c->owner ? batch_shape(c->owner) : effective_shape(c, c->shape), c->lc.cap,
              c->owner ? c->owner->tile_queue : c->tile_queue, out);
    return MPM_OK;
}

int mpm_device_arch(int32_t device, char* arch, uint32_t arch_bytes, uint32_t* compute_units) {
    hipDeviceProp_t prop;
    const hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) {
        g_create_error = std::string("hipGetDeviceProperties: ") + hipGetErrorString(e);
        return MPM_EDEVICE;
    }
    if (arch && arch_bytes) {
        strncpy(arch, prop.gcnArchName, arch_bytes - 1);
        arch[arch_bytes - 1] = '\0';
    }
    if (compute_units) *compute_units = (uint32_t)prop.multiProcessorCount;
    return MPM_OK;
}

int mpm_set_diagnostics(mpm_ctx* c, uint32_t enable) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    int s = settle(c);
    if (s) return s;
    if (enable && !c->lc.diag) {
        const uint32_t tiles = p2g_tiles_for(c->lc.node_w, c->lc.node_h, MIN_P2G_TILE);  // any shape
        hipError_t e;
        if ((e = c->mem.alloc((void**)&c->lc.diag_grid, sizeof(double) * 3 * tiles)) != hipSuccess ||
            (e = c->mem.alloc((void**)&c->lc.diag_part, sizeof(double) * 3 * DIAG_PART_BLOCKS)) != hipSuccess ||
            (e = c->mem.alloc((void**)&c->lc.diag, sizeof(double) * 9)) != hipSuccess)
            return hip_fail(c, e, "hipMalloc(diagnostics)");
        HIP_TRY(c, hipMemset(c->lc.diag_grid, 0, sizeof(double) * 3 * tiles));
        HIP_TRY(c, hipMemset(c->lc.diag, 0, sizeof(double) * 9));
    }
    c->diag_on = enable != 0u;
    if (!c->diag_on) {
        // keep the buffers; launches skip the reductions when the pointer is null
        c->diag_saved = c->lc.diag;
        c->lc.diag = nullptr;
    } else if (!c->lc.diag) {
        c->lc.diag = c->diag_saved;
    }
    return MPM_OK;
}

int mpm_read_diagnostics(mpm_ctx* c, double out[9]) {
    if (!c || !out) return fail(c, MPM_EINVAL, "context and output must not be NULL");
    if (!c->lc.diag && !c->diag_saved) return fail(c, MPM_ESTATE, "diagnostics were never enabled");
    int s = settle(c);
    if (s) return s;
    double d[9];
    HIP_TRY(c, hipMemcpy(d, c->lc.diag ? c->lc.diag : c->diag_saved, sizeof d, hipMemcpyDeviceToHost));
    // device order: mass (input, grid, output), momentum x (...), momentum y (...)
    const int order[9] = {0, 1, 2, 3, 6, 4, 7, 5, 8};
    for (int k = 0; k < 9; ++k) out[k] = d[order[k]];
    return MPM_OK;
}

int mpm_set_timing(mpm_ctx* c, uint32_t enable) {
    if (!c) return fail(nullptr, MPM_EINVAL, "context must not be NULL");
    int s = settle(c);
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
    int s = settle(c);
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
    if (!c || c->owner) return;  // scenes of a batch go with mpm_batch_destroy
    destroy_ctx(c);
}


const char* mpm_last_error(mpm_ctx* c) { return c ? c->error.c_str() : g_create_error.c_str(); }


// ------------------------------------------------------------------ batches

static int batch_fail(mpm_batch* b, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (b) b->error = buf; else g_create_error = buf;
    return code;
}

void mpm_batch_destroy(mpm_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (mpm_ctx* c : b->scenes) {
        if (!c) continue;
        c->owner = nullptr;
        destroy_ctx(c);
    }
    if (b->d_scenes) (void)hipFree(b->d_scenes);
    if (b->d_p2g_map) (void)hipFree(b->d_p2g_map);
    if (b->d_g2p_map) (void)hipFree(b->d_g2p_map);
    if (b->d_queue) (void)hipFree(b->d_queue);
    if (b->h_scenes) (void)hipHostFree(b->h_scenes);
    if (b->staged) (void)hipEventDestroy(b->staged);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

}  // extern "C"

namespace {

// The shape every scene of the batch can run (one kernel variant serves the
// launch).
Shape batch_shape(const mpm_batch* b) {
    Shape s = b->shape;
    for (const mpm_ctx* c : b->scenes) s = effective_shape(c, s);
    return s;
}

// Tiles of the step launches for a shape, scene after scene, and the device
// tile -> scene maps; rebuilt only when the shape changes.  The caller has
// made sure no launch of the batch still reads the old maps.
int batch_maps(mpm_batch* b, const Shape& shape) {
    if (b->map_shape.p2g_tile == shape.p2g_tile && b->map_shape.g2p_tile == shape.g2p_tile &&
        b->map_shape.g2p_threads == shape.g2p_threads)
        return MPM_OK;
    std::vector<uint32_t> pmap, gmap;
    b->p2g_base.clear();
    b->g2p_base.clear();
    b->p2g_tiles = b->g2p_tiles = 0;
    for (uint32_t i = 0; i < (uint32_t)b->scenes.size(); ++i) {
        const mpm_ctx* c = b->scenes[i];
        const uint32_t pt = p2g_tiles_for(c->lc.node_w, c->lc.node_h, shape.p2g_tile);
        const uint32_t gt = g2p_tiles_for(c->lc.bw, c->lc.bh, shape.g2p_tile);
        b->p2g_base.push_back(b->p2g_tiles);
        b->g2p_base.push_back(b->g2p_tiles);
        pmap.insert(pmap.end(), pt, i);
        gmap.insert(gmap.end(), gt, i);
        b->p2g_tiles += pt;
        b->g2p_tiles += gt;
    }
    hipError_t e;
    const uint32_t need = (uint32_t)std::max(pmap.size(), gmap.size());
    if (need > b->map_capacity) {
        if (b->d_p2g_map) (void)hipFree(b->d_p2g_map);
        if (b->d_g2p_map) (void)hipFree(b->d_g2p_map);
        b->d_p2g_map = b->d_g2p_map = nullptr;
        b->map_capacity = 0;
        if ((e = hipMalloc((void**)&b->d_p2g_map, sizeof(uint32_t) * need)) != hipSuccess ||
            (e = hipMalloc((void**)&b->d_g2p_map, sizeof(uint32_t) * need)) != hipSuccess)
            return batch_fail(b, MPM_EDEVICE, "batch tile maps: %s", hipGetErrorString(e));
        b->map_capacity = need;
    }
    if (b->d_queue) (void)hipFree(b->d_queue);
    b->d_queue = nullptr;
    if ((e = hipMalloc((void**)&b->d_queue, sizeof(uint32_t) * (2 * QUEUE_SET_WORDS + b->p2g_tiles + b->g2p_tiles))) !=
        hipSuccess)
        return batch_fail(b, MPM_EDEVICE, "batch tile queue: %s", hipGetErrorString(e));
    if ((e = hipMemcpy(b->d_p2g_map, pmap.data(), sizeof(uint32_t) * pmap.size(), hipMemcpyHostToDevice)) !=
            hipSuccess ||
        (e = hipMemcpy(b->d_g2p_map, gmap.data(), sizeof(uint32_t) * gmap.size(), hipMemcpyHostToDevice)) !=
            hipSuccess)
        return batch_fail(b, MPM_EDEVICE, "batch tile maps: %s", hipGetErrorString(e));
    b->map_shape = shape;
    return MPM_OK;
}

}  // namespace

extern "C" {

int mpm_batch_create(const mpm_config* configs, const uint32_t* const* packed20, const uint32_t* counts,
                     uint32_t nscenes, mpm_batch** out) {
    if (!out) return batch_fail(nullptr, MPM_EINVAL, "out must not be NULL");
    *out = nullptr;
    if (!configs || !packed20 || !counts) return batch_fail(nullptr, MPM_EINVAL, "configs, packed and counts must not be NULL");
    if (nscenes < 1 || nscenes > 65536) return batch_fail(nullptr, MPM_EINVAL, "scene count must be in [1, 65536]");
    for (uint32_t i = 1; i < nscenes; ++i) {
        if (configs[i].device != configs[0].device)
            return batch_fail(nullptr, MPM_EINVAL, "scene %u: every scene of a batch must use the same device", i);
        if (configs[i].p2g_mode != configs[0].p2g_mode)
            return batch_fail(nullptr, MPM_EINVAL, "scene %u: every scene of a batch must use the same P2G mode", i);
    }
    mpm_batch* b = new (std::nothrow) mpm_batch();
    if (!b) return batch_fail(nullptr, MPM_EDEVICE, "out of host memory");
    auto bail = [&](int code) {
        const std::string msg = b->error;
        mpm_batch_destroy(b);
        g_create_error = msg;
        return code;
    };
    b->scenes.reserve(nscenes);
    for (uint32_t i = 0; i < nscenes; ++i) {
        mpm_ctx* c = nullptr;
        const int r = mpm_create(&configs[i], packed20[i], nullptr, counts[i], &c);
        if (r != MPM_OK) {
            b->error = std::string("scene ") + std::to_string(i) + ": " + mpm_last_error(nullptr);
            return bail(r);
        }
        c->owner = b;
        b->scenes.push_back(c);
        if (c->lc.aux.naux) {
            b->error = std::string("scene ") + std::to_string(i) + ": bonded triangle bodies cannot be batched";
            return bail(MPM_EINVAL);
        }
    }
    hipError_t e;
    b->device = b->scenes[0]->device;
    {
        uint32_t tiles8 = 0;
        for (const mpm_ctx* c : b->scenes) tiles8 += g2p_tiles_for(c->lc.bw, c->lc.bh, 8u);
        b->shape = default_shape(tiles8, b->device);
    }
    if ((e = hipSetDevice(b->device)) != hipSuccess ||
        (e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreateWithFlags(&b->staged, hipEventDisableTiming)) != hipSuccess) {
        batch_fail(b, MPM_EDEVICE, "batch stream: %s", hipGetErrorString(e));
        return bail(MPM_EDEVICE);
    }
    // every scene moves to the batch's stream
    uint32_t flags = 0;
    for (mpm_ctx* c : b->scenes) {
        (void)hipStreamSynchronize(c->stream);
        (void)hipStreamDestroy(c->stream);
        c->stream = b->stream;
        c->lc.stream = b->stream;
        c->shared_stream = true;
        flags |= c->lc.scene_flags;
    }
    // one kernel variant serves the batch: when one scene needs the stress
    // cache, every scene carries it
    if (scene_general(flags)) {
        for (mpm_ctx* c : b->scenes) {
            if (scene_general(c->lc.scene_flags)) continue;
            c->lc.scene_flags |= SCENE_GENERAL;
            int r = alloc_slots(c, c->lc.cap);
            if (!r) r = rebucket_all(c, c->lc.cap);
            if (r) {
                b->error = c->error;
                return bail(r);
            }
        }
    }
    if ((e = hipMalloc((void**)&b->d_scenes, sizeof(SceneDesc) * nscenes)) != hipSuccess ||
        (e = hipHostMalloc((void**)&b->h_scenes, sizeof(SceneDesc) * nscenes, hipHostMallocDefault)) != hipSuccess) {
        batch_fail(b, MPM_EDEVICE, "batch buffers: %s", hipGetErrorString(e));
        return bail(MPM_EDEVICE);
    }
    const int m = batch_maps(b, batch_shape(b));
    if (m) return bail(m);
    *out = b;
    return MPM_OK;
}

mpm_ctx* mpm_batch_scene(mpm_batch* b, uint32_t scene) {
    if (!b || scene >= b->scenes.size()) return nullptr;
    return b->scenes[scene];
}

uint32_t mpm_batch_count(mpm_batch* b) { return b ? (uint32_t)b->scenes.size() : 0u; }

int mpm_batch_step(mpm_batch* b, const void* params, uint32_t step_count, uint32_t measure_speed) {
    if (!b) return batch_fail(nullptr, MPM_EINVAL, "batch must not be NULL");
    if (!params) return batch_fail(b, MPM_EINVAL, "params must not be NULL");
    if (step_count < 1) return batch_fail(b, MPM_EINVAL, "stepCount must be a positive safe integer.");
    const uint32_t n = (uint32_t)b->scenes.size();
    std::vector<Params> ps(n);
    bool measure = measure_speed != 0u;
    for (uint32_t i = 0; i < n; ++i) {
        const int r = prepare_params(b->scenes[i], (const uint8_t*)params + (size_t)MPM_PARAMS_BYTES * i, ps[i]);
        if (r) return batch_fail(b, r, "scene %u: %s", i, b->scenes[i]->error.c_str());
        measure = measure || ps[i].cfl_on >= 0.5f;
    }
    if (hipSetDevice(b->device) != hipSuccess) return batch_fail(b, MPM_EDEVICE, "hipSetDevice failed");
    for (mpm_ctx* c : b->scenes) {
        if (c->pending.size() >= MAX_PENDING) {
            const int r = settle(c);
            if (r) return batch_fail(b, r, "%s", c->error.c_str());
        }
    }
    if (b->staged_used && hipEventSynchronize(b->staged) != hipSuccess)
        return batch_fail(b, MPM_EDEVICE, "hipEventSynchronize failed");
    const Shape shape = batch_shape(b);
    if (!(shape.p2g_tile == b->map_shape.p2g_tile && shape.g2p_tile == b->map_shape.g2p_tile &&
          shape.g2p_threads == b->map_shape.g2p_threads)) {
        // launches already queued read the old maps: let them finish first
        if (hipStreamSynchronize(b->stream) != hipSuccess)
            return batch_fail(b, MPM_EDEVICE, "hipStreamSynchronize failed");
        const int m = batch_maps(b, shape);
        if (m) return m;
    }
    StepBatch sb{};
    sb.stream = b->stream;
    sb.d_scenes = b->d_scenes;
    sb.p2g_scene = b->d_p2g_map;
    sb.g2p_scene = b->d_g2p_map;
    sb.nscenes = n;
    sb.p2g_tiles = b->p2g_tiles;
    sb.g2p_tiles = b->g2p_tiles;
    sb.fixed_point = b->scenes[0]->lc.fixed_point;
    sb.shape = shape;
    for (uint32_t i = 0; i < n; ++i) {
        mpm_ctx* c = b->scenes[i];
        const uint32_t seq = c->seq++;
        b->h_scenes[i] = scene_desc(c, ps[i], b->p2g_base[i], b->g2p_base[i], seq, shape);
        c->pending.push_back({ps[i], step_count, measure, seq, 0u});
        c->last_done_before = 0u;
        c->cur = (c->cur + (int)(step_count & 1u)) & 1;
        sb.max_slots = std::max(sb.max_slots, c->lc.nb * c->lc.cap);
        sb.max_p2g_tiles = std::max(sb.max_p2g_tiles, p2g_tiles_for(c->lc.node_w, c->lc.node_h, shape.p2g_tile));
        sb.scene_flags |= c->lc.scene_flags;
        sb.diag = sb.diag || c->lc.diag != nullptr;
    }
    uint64_t nodes = 0, particles = 0;
    bool fixed_grid = false;
    for (const mpm_ctx* c : b->scenes) {
        nodes += c->lc.node_count;
        particles += c->n;
        fixed_grid = fixed_grid || c->lc.grid.fixed != nullptr;
    }
    if (!fixed_grid && queue_wanted(b->tile_queue, nodes, particles, b->p2g_tiles, b->device))
        sb.queue = {b->d_queue, {b->d_queue + 2 * QUEUE_SET_WORDS, b->d_queue + 2 * QUEUE_SET_WORDS + b->p2g_tiles},
                    {b->p2g_tiles, b->g2p_tiles}};
    hipError_t e;
    if ((e = hipMemcpyAsync(b->d_scenes, b->h_scenes, sizeof(SceneDesc) * n, hipMemcpyHostToDevice, b->stream)) !=
            hipSuccess ||
        (e = hipEventRecord(b->staged, b->stream)) != hipSuccess)
        return batch_fail(b, MPM_EDEVICE, "staging the batch: %s", hipGetErrorString(e));
    b->staged_used = true;
    if ((e = launch_steps(sb, step_count, measure)) != hipSuccess)
        return batch_fail(b, MPM_EDEVICE, "batch launch: %s", hipGetErrorString(e));
    return MPM_OK;
}

int mpm_batch_set_tuning(mpm_batch* b, const mpm_tuning* t) {
    if (!b) return batch_fail(nullptr, MPM_EINVAL, "batch must not be NULL");
    if (!t) return batch_fail(b, MPM_EINVAL, "tuning must not be NULL");
    std::string why;
    if (!tuning_valid(t, why)) return batch_fail(b, MPM_EINVAL, "%s", why.c_str());
    if (hipSetDevice(b->device) != hipSuccess) return batch_fail(b, MPM_EDEVICE, "hipSetDevice failed");
    for (uint32_t i = 0; i < (uint32_t)b->scenes.size(); ++i) {
        const int r = set_bucket_capacity(b->scenes[i], t->bucket_capacity);
        if (r) return batch_fail(b, r, "scene %u: %s", i, b->scenes[i]->error.c_str());
    }
    b->shape = {t->p2g_tile, t->g2p_tile, t->g2p_threads};
    b->tile_queue = t->tile_queue;
    return MPM_OK;
}

int mpm_batch_get_tuning(mpm_batch* b, mpm_tuning* out) {
    if (!b) return batch_fail(nullptr, MPM_EINVAL, "batch must not be NULL");
    if (!out) return batch_fail(b, MPM_EINVAL, "out must not be NULL");
    uint32_t cap = 0;
    for (const mpm_ctx* c : b->scenes) cap = std::max(cap, c->lc.cap);
    tuning_of(batch_shape(b), cap, b->tile_queue, out);
    return MPM_OK;
}

const char* mpm_batch_last_error(mpm_batch* b) { return b ? b->error.c_str() : g_create_error.c_str(); }

}  // extern "C"
