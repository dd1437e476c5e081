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
    float wall_friction, wall_pad[3];                          // wall
    float drag_target_x, drag_target_y, drag_radius, drag_stiffness; // interaction
    float drag_damping, drag_pad[3];
};
static_assert(sizeof(Params) == 176, "Params must mirror the 144+32 byte blocks");

// Control words: status (shaders.ts:157-172) plus batch bookkeeping.
struct Control {
    uint32_t status[8];
    uint32_t completed;       // committed microsteps of the current batch
    uint32_t max_speed_bits;  // f32 bits of the measured maximum speed
    uint32_t deferred;        // prepare failures found by G2P for the next microstep
    uint32_t deferred_index;  // lowest particle index among them
    uint32_t pad[4];
};

// Dynamic particle state of one ping-pong buffer (SoA).
struct ParticleBuf {
    float2* pos;
    float2* vel;
    float4* C;
    float4* F;
    float* plastic;
};

// Particle fields that do not change within a step.
struct ParticleStatic {
    uint32_t* model;
    float* lambda;
    float* mu;
    float* density;
    float* volume;
    float* param;
};

struct Bins {
    uint32_t* key;    // per particle: gridIndex(base) or INVALID_KEY
    uint32_t* rank;   // per particle: slot inside its bin (arrival order)
    uint32_t* count;  // per node; zero between uses
    uint32_t* start;  // node_count + 1
    uint32_t* perm;   // particle indices grouped by bin
    uint32_t* block_sums;
    float4* affine;   // per particle: A = m*C + (-dt*V*4/dx^2) P F^T (preparedAffines)
    float4* aux;      // per particle: (mass, volume, det F, phase bits)
};

struct Grid {
    float2* v_solid;
    float2* v_water;
    float2* v_oil;
    float2* liquid;     // (water+oil mass, water+oil volume)
    int32_t* fixed;     // node_count*11 or nullptr
};

// Scene content flags, host-derived from the particle models.
enum : uint32_t { SCENE_SOLID = 1u, SCENE_WATER = 2u, SCENE_OIL = 4u, SCENE_TAIT = 8u };
// Scenes with solids or Tait liquids need the FP64 stress path.
inline bool scene_general(uint32_t flags) { return (flags & (SCENE_SOLID | SCENE_TAIT)) != 0u; }

struct KernelTimes;  // host-side timing (mpm_capi.hip)

// Launchers (mpm_kernels.hip).  All work is enqueued on `stream`.
struct LaunchCtx {
    hipStream_t stream;
    const Params* params;  // device
    Control* ctl;          // device
    ParticleBuf buf[2];
    ParticleStatic st;
    Bins bins;
    Grid grid;
    uint32_t n;            // active particle slots
    uint32_t node_w, node_h, node_count;
    uint32_t fixed_point;  // 0 = float gather, 1 = int32 fixed point
    uint32_t keep_fixed;
    uint32_t scene_flags;
    uint32_t scan_blocks;
    KernelTimes* timing;   // nullptr when timing is off
};

hipError_t launch_batch(const LaunchCtx& lc, int start_buf, uint32_t steps, bool measure);
size_t scan_block_count(uint32_t node_count);

}  // namespace mpm
