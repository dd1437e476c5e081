// Internal types shared by the gfx950 kernels (mpm_kernels.hip) and the C-ABI
// host layer (mpm_capi.hip).  Not part of the public interface.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace mpm {

// Error bits of the persistent status word 0 (shaders.ts:157-172).
enum : uint32_t {
    ERR_STENCIL = 1u,
    ERR_DEFORMATION = 2u,
    ERR_P2G = 4u,
    ERR_GRID = 8u,
    ERR_G2P = 16u,
    ERR_SUBSTEP_LIMIT = 32u,
    ERR_PROJECTED_DEFORMATION = 64u,
    ERR_INVALID_MATERIAL = 128u,
    ERR_NONFINITE_PARTICLE = 256u,
    ERR_NONFINITE_ACCUMULATION = 1024u,
    ERR_NONFINITE_NODE = 2048u,
    ERR_TRIANGLE_CONTACT = 4096u,
    ERR_TRIANGLE_BOND = 8192u,
    ERR_CFL_BOUND = 16384u,
    ERR_FRACTURE_BOND = 32768u,
    ERR_FIXED_POINT = 65536u,
    // Not a reference flag: a bucket ran out of slots.  The step rolls back
    // like any failure and the host grows the buckets and replays the
    // remaining microsteps; callers never see this bit.
    ERR_INTERNAL_REBUCKET = 0x80000000u,
};

enum : uint32_t {
    MODEL_NEO_HOOKEAN = 0u,
    MODEL_SNOW = 1u,
    MODEL_WATER = 2u,
    MODEL_OIL = 3u,
    MODEL_POWDER = 4u,
    MODEL_REMOVED = 0xffffffffu,
};

constexpr uint32_t INVALID_KEY = 0xffffffffu;
constexpr int BUCKET_CELLS = 4;  // a bucket holds the particles of 4x4 stencil-base cells

// Decoded step-parameter block: the 144 bytes of packMpmMlsGpuStepParameters
// (layout.ts:896-997) followed by the 32-byte drag interaction uniform
// (solver.ts:1389-1401).  Lives in device memory so a captured launch
// sequence stays valid when the parameters change.
struct Params {
    uint32_t cells_x, cells_y, node_w, node_h;                 // grid
    uint32_t particle_count, node_count, pair_count, body_count; // counts
    float dx, particle_mass, particle_volume, dt;              // physical
    float lambda, mu, stress_scale, grad_scale;                // material
    float gx, gy, dom_w, dom_h;                                // gravityAndDomain
    float bond_rest, bond_stiffness, band, reach;              // bonding
    float wave_bound, vel_bound, cfl_on, bond_ref_dt;          // cflBounds
    uint32_t fracture[4];                                      // fracture
    float wall_friction;                                       // wall
    float inv_mass_ref, inv_volume_ref;  // RN(1/particleMass), RN(1/particleVolume) or 0 (host-set)
    float inv_dx;  // 1/dx when dx is a power of two (x/dx == x*inv_dx exactly), else 0 (host-set)
    float drag_target_x, drag_target_y, drag_radius, drag_stiffness; // interaction
    float drag_damping;
    uint32_t diag_on;  // FP64 mass/momentum reductions of the batch (host-set)
    float drag_pad[2];
};
static_assert(sizeof(Params) == 176, "Params must mirror the 144+32 byte blocks");

// 1/x when x is a power of two whose reciprocal is a normal float (then
// y / x and y * (1/x) are the same correctly rounded value for every y), else
// 0: the kernels divide by x.
__host__ __device__ inline float exact_reciprocal(float x) {
    uint32_t u;
    __builtin_memcpy(&u, &x, 4);
    const uint32_t e = (u >> 23) & 0xffu;
    if ((u & 0x807fffffu) != 0u || e < 2u || e > 252u) return 0.0f;
    return 1.0f / x;
}

// Control words: status (shaders.ts:157-172) plus batch bookkeeping.
struct Control {
    uint32_t status[8];
    uint32_t completed;       // committed microsteps of the current batch
    uint32_t max_speed_bits;  // f32 bits of the measured maximum speed
    uint32_t deferred[2];       // per slot buffer: stress-cache failures of the state it holds
    uint32_t deferred_index[2]; // lowest particle index among them
    uint32_t fail_step;         // first microstep of the batch that raised a flag
    uint32_t pad;
    uint32_t history[64];       // committed microsteps per batch, by sequence number
};
constexpr uint32_t CONTROL_HISTORY = 64;

// Pointer fields of the structures the kernels read out of a scene
// descriptor in device memory.  In the device pass they carry the global
// address space: the compiler cannot infer it for a pointer loaded from
// memory, and a generic pointer compiles to flat memory instructions, which
// take 64-bit addresses in vector registers and are counted together with
// the LDS operations (every wait for an LDS result then also waits for the
// memory loads in flight).  Both passes see 64-bit pointers, so the layouts
// agree; kernel parameters stay generic, so kernel names agree too.  Only
// the kernels' translation unit asks for it (MPM_GLOBAL_POINTERS): the device
// pass also checks host code, which fills these fields from host pointers.
#if defined(__HIP_DEVICE_COMPILE__) && defined(MPM_GLOBAL_POINTERS)
#define MPM_GLOBAL __attribute__((address_space(1)))
#else
#define MPM_GLOBAL
#endif
template <class T>
using dptr = T MPM_GLOBAL*;

// One material row: particles reference it by index (material id).  Rows are
// the distinct (model, lambda, mu, density, volume, param) of the scene.
struct Material {
    uint32_t model;
    float lambda, mu, density, volume, param;
    float mass;     // density*volume for liquids, particleMass otherwise
    uint32_t phase; // 0 solid, 1 water, 2 oil
};
static_assert(sizeof(Material) == 32, "Material row is 32 bytes");

// Particle state stored by bucket: slot = bucket * cap + k, k < cnt[bucket].
struct SlotBuf {
    dptr<float2> pos;
    dptr<float2> vel;
    dptr<float4> C;
    dptr<float4> F;
    dptr<float4> S;        // P*F^T stress cache (scenes with solids or Tait liquids)
    dptr<float> plastic;   // snow plastic volume (scenes with solids)
    dptr<uint32_t> pid;    // particle index in the caller's order
    dptr<uint32_t> mat;    // material row
    dptr<uint32_t> cnt;    // per bucket
};

// Particle arrays in the caller's index order (upload, readback, removed).
struct Home {
    dptr<float2> pos;
    dptr<float2> vel;
    dptr<float4> C;
    dptr<float4> F;
    dptr<float> plastic;
    dptr<uint32_t> model;
    dptr<uint32_t> mat;
};

struct Grid {
    dptr<float2> v_solid;
    dptr<float2> v_water;
    dptr<float2> v_oil;
    dptr<float2> liquid;     // (water+oil mass, water+oil volume)
    dptr<int32_t> fixed;     // node_count*11 or nullptr
    dptr<float2> occupancy;  // (water, oil) surface occupancy per node, written by the occupancy gather only
};

// Scene content flags, host-derived from the particle models.  SCENE_AUX:
// the scene has bonded triangle bodies or fracture bonds, whose particles are
// reached by index through the aux table below.
enum : uint32_t { SCENE_SOLID = 1u, SCENE_WATER = 2u, SCENE_OIL = 4u, SCENE_TAIT = 8u, SCENE_GENERAL = 16u,
                  SCENE_AUX = 32u };

// Config-2 bodies and bonds.  Triangle field of one node (buildTriangleField,
// shaders.ts:1324-1436): signed distance, outward normal and barycentric
// weights of the nearest point of the nearest triangle within the band, and
// the impulse the fluid contact puts on that body at this node.
struct FieldNode {
    float dist, nx, ny, b0, b1, b2, ix, iy;
};
// A fracture bond in the reference's 24-byte layout (layout.ts:527-548).
struct Bond {
    uint32_t first, second;
    float rest, failure, damage;
    uint32_t broken;
};
static_assert(sizeof(Bond) == 24, "fracture bond record is 24 bytes");
// Particles the contact and fracture kernels address by index ("aux"): the
// vertices of body b are aux 3b, 3b+1, 3b+2; then every fracture-bond
// endpoint in ascending particle index.  Both slot buffers keep where each
// aux particle sits; the bucket and G2P kernels write it.
struct AuxData {
    dptr<uint32_t> aux_of;      // per particle index: aux index or NO_AUX
    dptr<uint32_t> aux_pid;     // per aux: particle index
    dptr<uint32_t> slot[2];     // per aux and slot buffer: slot
    dptr<float> damage[2];      // per aux and slot buffer: neo-Hookean damage (fracture endpoints)
    dptr<uint32_t> body_oid;    // per body: object id
    dptr<FieldNode> field;      // per node
    dptr<uint32_t> field_body;  // per node: body index + 1, 0 = no triangle within the band
    dptr<Bond> bonds;
    dptr<uint32_t> bond_off;    // per aux + 1: CSR offsets into bond_inc
    dptr<uint32_t> bond_inc;    // incident bond indices
    uint32_t nbodies, naux, nbonds, frac_aux0;  // frac_aux0 = 3 * nbodies
};
constexpr uint32_t NO_AUX = 0xffffffffu;
// Scenes with solids or Tait liquids carry the FP64 stress cache; SCENE_GENERAL
// asks for it regardless (a scene stepped together with such scenes).
__host__ __device__ inline bool scene_general(uint32_t flags) {
    return (flags & (SCENE_SOLID | SCENE_TAIT | SCENE_GENERAL)) != 0u;
}

struct KernelTimes;  // host-side timing (mpm_timing.h)

struct LaunchCtx {
    hipStream_t stream;
    Control* ctl;          // device
    SlotBuf buf[2];
    Home home;
    const Material* mats;  // device
    uint32_t nmat;         // material rows
    Grid grid;
    uint32_t n;            // active particle slots (caller indices)
    uint32_t node_w, node_h, node_count;
    uint32_t cells_x, cells_y;
    float cell_size;
    uint32_t bw, bh, nb, cap;  // buckets: bw x bh, nb total, cap slots each
    uint32_t fixed_point;  // 0 = float gather, 1 = int32 fixed point
    uint32_t keep_fixed;
    uint32_t scene_flags;
    AuxData aux;           // bodies / fracture (zeros without SCENE_AUX)
    KernelTimes* timing;   // nullptr when timing is off
    // FP64 diagnostics (nullptr when off): per-tile grid partials of the last
    // P2G, per-block particle partials, and the nine reported sums
    double* diag_grid;
    double* diag_part;
    double* diag;
};
constexpr uint32_t DIAG_PART_BLOCKS = 1024;  // blocks of the particle-sum kernel

// Launch shape of the step kernels: the P2G tile side in grid nodes, the
// G2P tile side in buckets and the threads of a G2P block (particles per
// wavefront = particles of a G2P tile / its waves).  Every supported shape is
// compiled; a step's results do not depend on the shape.
struct Shape {
    uint32_t p2g_tile, g2p_tile, g2p_threads;
};
constexpr Shape DEFAULT_SHAPE = {20u, 8u, 256u};
// Default for a grid whose 8x8-bucket G2P tiles are fewer than the device's
// compute units: every G2P block is then alone on its CU and latency-bound,
// and a thread's particles run one after another, so the 4x4-bucket tile (a
// quarter of the particles a block) finishes sooner (the 1,600-particle Tait
// tank: 10.8 us against 24.9 us per G2P launch on MI355X).
constexpr Shape SMALL_GRID_SHAPE = {20u, 4u, 256u};
// Scenes with bonded bodies or fracture bonds run the largest P2G tile, whose
// gather chunk holds any bucket, and 256-thread G2P blocks of either tile
// side (the G2P tile of the requested shape; 8 here).
constexpr Shape AUX_SHAPE = {28u, 8u, 256u};
constexpr uint32_t MIN_P2G_TILE = 12u;  // the most P2G tiles a grid can have
// Particles the float-gather P2G stages per chunk, for each tile side: the
// particles of one bucket must fit one chunk.
constexpr uint32_t p2g_chunk_for(uint32_t tile) { return tile == 28u ? 1408u : (tile == 20u ? 960u : 448u); }
inline bool shape_supported(const Shape& s) {
    const bool p = s.p2g_tile == 28u || s.p2g_tile == 20u || s.p2g_tile == 12u;
    const bool g = (s.g2p_tile == 8u && (s.g2p_threads == 128u || s.g2p_threads == 256u || s.g2p_threads == 512u)) ||
                   (s.g2p_tile == 4u && (s.g2p_threads == 128u || s.g2p_threads == 256u));
    return p && g;
}
// P2G tiles a side: tile t covers interior nodes [tile t - 1, tile t + tile - 2]
// (its region then starts on a bucket boundary), so n nodes need
// ceil((n + 1) / tile) tiles.
inline uint32_t p2g_tiles_side(uint32_t nodes, uint32_t tile) { return (nodes + tile) / tile; }
inline uint32_t p2g_tiles_for(uint32_t node_w, uint32_t node_h, uint32_t tile) {
    return p2g_tiles_side(node_w, tile) * p2g_tiles_side(node_h, tile);
}
inline uint32_t g2p_tiles_for(uint32_t bw, uint32_t bh, uint32_t tile) {
    return ((bw + tile - 1) / tile) * ((bh + tile - 1) / tile);
}

// One scene of a step launch, in device memory.  The step kernels of a
// launch cover the tiles of every scene; a block finds its scene through a
// tile -> scene map and reads everything scene-specific from here.
struct SceneDesc {
    Params P;           // this scene's step parameters
    dptr<Control> ctl;
    SlotBuf buf[2];
    dptr<const Material> mats;
    Grid grid;
    dptr<double> diag_grid;  // per-tile grid sums of the last P2G (diagnostics on), else nullptr
    dptr<double> diag_part;  // particle partial sums
    dptr<double> diag;       // the nine reported sums
    uint32_t nmat, bw, bh, nb, cap;
    uint32_t ptx, pty, gtx, gty;  // P2G and G2P tile grids
    uint32_t p2g_base, g2p_base;  // first tile of this scene in the P2G / G2P launches
    uint32_t scene_flags, keep_fixed;
    AuxData aux;
    uint32_t seq;                 // batch sequence number of this scene
    uint32_t start_buf;           // slot buffer holding the state the batch starts from
};

// Active-tile queues of a step launch.  When `sets` is set, a compaction
// kernel lists each microstep the P2G (kind 0) and G2P (kind 1) tiles that
// hold particles, and the step kernels run as a fixed number of blocks that
// share out those lists; otherwise every tile gets a block.  A list holds
// the active tiles of an XCD's range of the launch (the ranges of xcd_tile)
// from that range's start.  Two sets of counters, fill[kind][xcd] (tiles
// listed), alternate by microstep parity.
constexpr uint32_t QUEUE_XCDS = 8;
constexpr uint32_t QUEUE_SET_WORDS = 2 * QUEUE_XCDS;
struct TileQueue {
    uint32_t* sets;      // device, 2 x QUEUE_SET_WORDS; nullptr: one block per tile
    uint32_t* list[2];   // device: tiles[0] and tiles[1] entries
    uint32_t tiles[2];   // tiles of the P2G and G2P launches
};
__host__ __device__ inline uint32_t xcd_range_start(uint32_t tiles, uint32_t x) {
    const uint32_t per = tiles / QUEUE_XCDS, extra = tiles % QUEUE_XCDS;
    return x * per + (x < extra ? x : extra);
}
__host__ __device__ inline uint32_t xcd_of_tile(uint32_t tiles, uint32_t t) {
    const uint32_t per = tiles / QUEUE_XCDS, extra = tiles % QUEUE_XCDS, big = extra * (per + 1u);
    return t < big ? t / (per + 1u) : extra + (t - big) / per;  // per > 0 whenever t >= big
}

// Host description of one step launch over `nscenes` scenes that share a
// stream (each starts from its own slot buffer).
struct StepBatch {
    hipStream_t stream;
    const SceneDesc* d_scenes;   // device, nscenes entries (each carries its Params)
    const uint32_t* p2g_scene;   // device: scene of every P2G tile
    const uint32_t* g2p_scene;   // device: scene of every G2P tile
    uint32_t nscenes;
    uint32_t p2g_tiles, g2p_tiles;
    uint32_t max_slots;          // largest nb * cap of the scenes
    uint32_t max_p2g_tiles;      // largest per-scene P2G tile count
    uint32_t scene_flags;        // union of the scenes' flags (kernel variant)
    uint32_t fixed_point;
    uint32_t aux_bodies, aux_bonds;  // SCENE_AUX (one scene): body and bond counts
    uint32_t node_count;             // SCENE_AUX: the scene's grid nodes
    Shape shape;                     // launch shape (28-node P2G, 256-thread G2P for SCENE_AUX)
    bool diag;                   // some scene has diagnostics on
    KernelTimes* timing;
    TileQueue queue;             // sets == nullptr: one block per tile
};

hipError_t launch_steps(const StepBatch& sb, uint32_t steps, bool measure);
// Bucket particles [first, first+count) of the home arrays into buffer b
// (counts must already hold the particles bucketed so far).
hipError_t launch_bucket(const LaunchCtx& lc, int b, uint32_t first, uint32_t count);
hipError_t launch_clear_counts(const LaunchCtx& lc, int b);
// Write buffer b's particles back to the home arrays.
hipError_t launch_unbucket(const LaunchCtx& lc, int b);
// Stencil-base keys of the home positions (debug/parity).
hipError_t launch_keys(const LaunchCtx& lc, uint32_t* keys);
// Interaction operations on the caller-order arrays (home) or the committed
// slot buffer b.  outlet: water and oil particles inside the closed box
// (min x, min y, max x, max y) become removed, *removed counts them.  pick:
// *best receives the smallest (float32 bits of the squared distance << 32 |
// particle index) over the particles within sqrt(r2) of (px, py), and *where
// the position of that particle.  clone: home particles [src, src + count)
// are copied to [dst, dst + count), translated and clamped to the domain.
hipError_t launch_outlet(const LaunchCtx& lc, float4 box, uint32_t* removed);
// Surface occupancy of the committed state (gatherSurfaceOccupancy,
// shaders.ts:710-806) by the float gather over 28-node tiles: d_scene holds
// one descriptor whose Params carry the grid, grid.occupancy the output.
hipError_t launch_occupancy(hipStream_t stream, const SceneDesc* d_scene, uint32_t tiles, bool general,
                            KernelTimes* timing);
hipError_t launch_pick(const LaunchCtx& lc, int b, float px, float py, float r2, unsigned long long* best,
                       float2* where);
hipError_t launch_clone(const LaunchCtx& lc, uint32_t src, uint32_t count, uint32_t dst, float ox, float oy,
                        float dom_w, float dom_h);

}  // namespace mpm
