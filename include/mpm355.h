/*
 * mpm355 — MI355X (gfx950) 2D MLS-MPM step path behind a plain C ABI.
 *
 * Drop-in boundary for the WebGPU runtime below `MpmMlsGpuSolver`
 * (src/lib/mpm/kernels/solver.ts:106-144).  Every entry point takes plain
 * pointers and sizes; no GPU or framework types cross the boundary.  Wire
 * formats are the reference's own:
 *   - particle record: 20 x 32-bit words (layout.ts:385-437)
 *       [0]x [1]y [2]vx [3]vy [4..7]C(m00,m01,m10,m11) [8..11]F(m00,m01,m10,m11)
 *       [12]plasticVolume [13]model(u32) [14]lambda [15]mu [16]density
 *       [17]volume [18]param (Tait exponent / powder friction / NH damage)
 *       [19]bondedTriangleVertex+1 (u32, 0 = none)
 *   - step parameters: the 144-byte block of packMpmMlsGpuStepParameters
 *       (layout.ts:896-997)
 *   - numerical status: 8 x u32 (shaders.ts:157-172, solver.ts:889-893)
 *   - compact readback: f32 maxSpeed | u32 completed | 8 x u32 status
 *       (solver.ts:2396-2441)
 *
 * Conventions (solver.ts:1414-1441, kernels/solver.spec.ts:1324-1337):
 * arguments are validated before any device work and an invalid call returns
 * MPM_EINVAL with no state change; numerical failures inside a step are not
 * error returns — they set persistent status words and roll the particles
 * back to the last committed microstep.  Calls on one context must be
 * serialized by the caller (the reference allows one batch in flight).
 */
#ifndef MPM355_H
#define MPM355_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MPM_OK 0
#define MPM_EINVAL 1    /* invalid argument: maps to RangeError/TypeError */
#define MPM_EDEVICE 2   /* HIP runtime failure: maps to Error with cause */
#define MPM_ESTATE 3    /* call not valid in the context's current state */
#define MPM_ECAPACITY 4 /* particle capacity exceeded */

#define MPM_P2G_GATHER 0u      /* float32 P2G (createMpmMlsGpuSolver) */
#define MPM_P2G_FIXED_POINT 1u /* int32 2^24 fixed-point P2G
                                  (createExperimentalMpmMlsAtomicGpuSolver) */

#define MPM_FLAG_KEEP_FIXED_GRID 1u /* keep the last step's 11-word int32
                                       accumulators readable through
                                       mpm_read_fixed_grid */

#define MPM_PARAMS_BYTES 144
#define MPM_STATUS_WORDS 8
#define MPM_PARTICLE_WORDS 20

typedef struct mpm_ctx mpm_ctx;

/* Solver configuration: the GPU-relevant part of MpmMlsGpuConfig
 * (layout.ts:17-19) after float32 packing (layout.ts:242-270). */
typedef struct mpm_config {
    uint32_t cells_x, cells_y;
    float cell_size;
    float particle_mass;
    float particle_volume;
    float wall_friction;
    uint32_t capacity;   /* particle slots; >= initial count */
    uint32_t p2g_mode;   /* MPM_P2G_GATHER or MPM_P2G_FIXED_POINT */
    uint32_t flags;      /* MPM_FLAG_* */
    int32_t device;      /* HIP device ordinal; -1 = current */
    uint32_t bucket_slots; /* initial slots per 4x4-cell bucket; 0 = twice the
                              densest bucket of the initial state (>= 32).
                              Buckets grow on demand either way. */
} mpm_config;

/* Context creation and upload of the initial packed particles.
 * Replaces createGpuSolver/createRuntime (solver.ts:408-469, :605-1230).
 * packed20: n records of 20 words; object_ids: n words (may be NULL).
 * Bonded triangle bodies are three consecutive records with word 19 = 1, 2, 3
 * (bondedTriangleVertex + 1) of one object and a solid material
 * (collectBondedBodyStarts, reference/solver.ts:1023-1062); their contact
 * field, fluid contact, edge constraints and pair contacts run in every
 * microstep (at most 1024 bodies; such scenes cannot be batched and cannot
 * take appended bonded vertices). */
int mpm_create(const mpm_config* cfg, const uint32_t* packed20, const uint32_t* object_ids,
               uint32_t n, mpm_ctx** out);

/* Enqueue `step_count` microsteps with one parameter block.
 * Replaces stepBatch encode+submit (solver.ts:1407-1599): per microstep
 * prepare+bin -> P2G -> grid update -> G2P (+CFL guard when params enable it)
 * -> commit/rollback; the maximum particle speed of the final state is
 * measured when measure_speed != 0 (or the CFL guard is on).  Asynchronous:
 * returns once the work is queued. */
int mpm_step_batch(mpm_ctx* ctx, const void* params144, uint32_t step_count, uint32_t measure_speed);

/* Wait for queued work and read the 40-byte compact snapshot.
 * Replaces measureMaximumParticleSpeed's copy + mapAsync (solver.ts:2388-2512). */
int mpm_read_compact(mpm_ctx* ctx, float* max_speed, uint32_t* completed, uint32_t status[8]);

/* checkNumericalStatus (solver.ts:2514-2580) */
int mpm_read_status(mpm_ctx* ctx, uint32_t status[8]);

/* writeBuffer(stateError, zeros) (solver.ts:1878-1882): resets status to
 * {0, 0xFFFFFFFF, 0, 0, 0, 0, 0, 0}. */
int mpm_clear_status(mpm_ctx* ctx);

/* readParticles (solver.ts:2594-2731): indices == NULL reads slots [0, n).
 * out20 receives n records of 20 words, object_ids n words (may be NULL). */
int mpm_read_particles(mpm_ctx* ctx, const uint32_t* indices, uint32_t n, uint32_t* out20,
                       uint32_t* object_ids);

/* appendParticles device writes (solver.ts:2216-2242): appends n records
 * after the active slots; MPM_ECAPACITY if they do not fit. */
int mpm_append_particles(mpm_ctx* ctx, const uint32_t* packed20, const uint32_t* object_ids, uint32_t n);

/* Fracture bonds (config.fractureBonds, validated as validateMpmSolidBondGraph,
 * reference/solid-bonds.ts:11-66) in the reference's 24-byte records
 * (layout.ts:527-548): u32 first, u32 second, f32 restLength,
 * f32 failureStrain, f32 damage (1 when broken), u32 broken.  Set once, before
 * the first step.  Each microstep then updates the bonds' damage from the
 * tensile strain and raises the endpoints' damage (record word 18), as
 * updateFractureBonds / applyFractureParticleDamage (shaders.ts:2957-3039).
 * mpm_read_fracture_bonds returns the committed bond state. */
int mpm_set_fracture_bonds(mpm_ctx* ctx, const void* bonds24, uint32_t count);
int mpm_read_fracture_bonds(mpm_ctx* ctx, void* out24, uint32_t count);

/* Drag interaction uniform (solver.ts:1389-1401):
 * p[0..3] = target x, target y, radius (0 = off), stiffness; p[4] = damping. */
int mpm_set_interaction(mpm_ctx* ctx, const float p[8]);

/* removeParticlesInOutlet (solver.ts:2280-2386; removeParticlesInOutlet,
 * shaders.ts:3041-3065): every water and oil particle whose position lies in
 * the closed box bounds = {minX, minY, maxX, maxY} becomes a removed
 * tombstone (record word 13 = 0xFFFFFFFF; its slot stays allocated and
 * mpm_read_particles still returns it).  *removed (may be NULL) receives the
 * count.  MPM_EINVAL for non-finite bounds or min > max. */
int mpm_remove_in_outlet(mpm_ctx* ctx, const float bounds[4], uint32_t* removed);

/* pick (solver.ts:2733-2823; pickParticle, shaders.ts:3082-3118), as a
 * parallel arg-min over the committed particles: the non-removed particle
 * whose float32 squared distance (x - px)^2 + (y - py)^2 is at most
 * radius^2 and smallest, the lowest index on a tie (the reference's serial
 * scan order).  out = {index or 0xFFFFFFFF, f32 bits of its x, of its y,
 * its object id}: the reference's 16-byte pick result. */
int mpm_pick(mpm_ctx* ctx, float x, float y, float radius, uint32_t out[4]);

/* clone (solver.ts:2852-2948; cloneObject, shaders.ts:3120-3147): copies the
 * particles [start, start + count) after the active slots with their
 * positions translated by the offset and clamped to the domain, and object
 * id `object_id`; bonded triangle bodies inside the range are cloned as
 * bodies.  MPM_ECAPACITY when the copies do not fit; MPM_ESTATE (the
 * reference returns false) after an outlet removal or when the range holds
 * fracture-bond endpoints; MPM_EINVAL for a bad range or offset. */
int mpm_clone(mpm_ctx* ctx, uint32_t start, uint32_t count, float offset_x, float offset_y, uint32_t object_id);

/* gatherSurfaceOccupancy (shaders.ts:710-806; CPU form
 * reference/occupancy.ts:70-140): the material-separated surface field of
 * the committed state, out[2 * node] = (water, oil) = the display area
 * 2 V det F of the liquid particles, B-spline weighted onto the node and
 * divided by the cell area (node_count * 2 floats, gridIndex order).  The
 * sums follow the P2G gather's fixed order (the reference's follow its
 * linked lists).  MPM_ESTATE while the numerical status is set. */
int mpm_gather_occupancy(mpm_ctx* ctx, float* out, size_t floats);

/* Grid nodes the allocation-free occupancy gather can serve (8 MiB of
 * float2 scratch per GPU). */
#define MPM_STATIC_SCRATCH_NODES 1048576u

/* No counterpart in the reference.  mpm_gather_occupancy for callers that
 * cannot allocate device memory (tight render loops, hosts that forbid
 * hipMalloc): the same output, bit for bit, with the per-node scratch in a
 * static device array of MPM_STATIC_SCRATCH_NODES float2, one copy per GPU,
 * instead of a buffer the library allocates.  It calls no hipMalloc or
 * hipFree.  MPM_EINVAL for a grid of more than MPM_STATIC_SCRATCH_NODES
 * nodes; MPM_ESTATE while the numerical status is set or while queued steps
 * await a bucket reallocation (any readback settles those first).  Calls on
 * one GPU share its array; the library serialises them with a per-device
 * lock, and each call rewrites every node it returns.  Kernel timing
 * (mpm_set_timing) does not record this gather. */
int mpm_gather_occupancy_static(mpm_ctx* ctx, float* out, size_t floats);

/* Number of active particle slots and the capacity. */
int mpm_get_counts(mpm_ctx* ctx, uint32_t* particle_count, uint32_t* capacity, uint32_t* node_count);

/* Debug/parity: the int32 fixed-point accumulators of the last committed P2G
 * (node_count*11 words, order [solid m,qx,qy | water m,qx,qy,V | oil m,qx,qy,V]).
 * Requires MPM_P2G_FIXED_POINT and MPM_FLAG_KEEP_FIXED_GRID. */
int mpm_read_fixed_grid(mpm_ctx* ctx, int32_t* out, size_t words);

/* Debug/parity: per-node grid state of the last step (node_count*8 floats:
 * solid v(2), water v(2), oil v(2), liquid mass, liquid volume). */
int mpm_read_grid(mpm_ctx* ctx, float* out, size_t floats);

/* Debug/parity: stencil-base keys of the committed particles
 * (gridIndex(base), shaders.ts:518-520; 0xFFFFFFFF for removed particles) and
 * the particle count of every 4x4-cell bucket (bucket_counts may be NULL;
 * count_words must equal the bucket count of mpm_get_buckets). */
int mpm_read_bins(mpm_ctx* ctx, uint32_t* keys, uint32_t n, uint32_t* bucket_counts, size_t count_words);

/* Bucket grid (buckets_x * buckets_y buckets of 4x4 stencil-base cells) and
 * the current slots per bucket (grows on demand). */
int mpm_get_buckets(mpm_ctx* ctx, uint32_t* buckets_x, uint32_t* buckets_y, uint32_t* slots_per_bucket);

/* FP64 mass and momentum of the batch (MpmStepDiagnostics,
 * reference/solver.ts:96-103, :481-492, :2505-2526).  When enabled, every
 * mpm_step_batch also sums, in FP64 and in a fixed order, the particle mass
 * and momentum of the state it starts from and of the state it commits, and
 * the grid mass and momentum of its last P2G.  out[9] = inputMass, gridMass,
 * outputMass, inputMomentum x, y, gridMomentum x, y, outputMomentum x, y of
 * the last batch.  No counterpart call in the reference, whose CPU step
 * returns these per step. */
int mpm_set_diagnostics(mpm_ctx* ctx, uint32_t enable);
int mpm_read_diagnostics(mpm_ctx* ctx, double out[9]);

/* Batched scenes: many independent scenes advanced by the same kernels, one
 * launch per microstep over the tiles of every scene.  The reference has no
 * counterpart (one solver per scene, solver.ts:390-406); this serves callers
 * that step many small scenes, whose single-scene steps would be bound by
 * launch latency.  Each scene keeps its own status words, rollback and
 * bucket growth, exactly as a context from mpm_create.
 *
 * mpm_batch_create: scenes [0, nscenes) from configs[i], packed20[i] and
 *   counts[i] (same device and P2G mode for all).
 * mpm_batch_step: params holds nscenes consecutive 144-byte blocks; every
 *   scene advances step_count microsteps with its own block.
 * mpm_batch_scene: the context of one scene, for the per-scene calls
 *   (mpm_read_compact, mpm_read_particles, mpm_read_status,
 *   mpm_clear_status, diagnostics...).  It belongs to the batch:
 *   mpm_destroy ignores it and mpm_batch_destroy releases it. */
typedef struct mpm_batch mpm_batch;
int mpm_batch_create(const mpm_config* configs, const uint32_t* const* packed20, const uint32_t* counts,
                     uint32_t nscenes, mpm_batch** out);
int mpm_batch_step(mpm_batch* batch, const void* params, uint32_t step_count, uint32_t measure_speed);
mpm_ctx* mpm_batch_scene(mpm_batch* batch, uint32_t scene);
uint32_t mpm_batch_count(mpm_batch* batch);
void mpm_batch_destroy(mpm_batch* batch);
const char* mpm_batch_last_error(mpm_batch* batch);

/* ---------------- kernel tuning ----------------
 * No counterpart in the reference (its WGSL kernels have fixed workgroup
 * sizes).  The launch shape of the step kernels and the bucket capacity,
 * for the autotuner (powder-play_amd/autotune.py), which times candidates on
 * a caller's scene and caches the fastest per (GPU architecture, scene
 * class).  A step's results do not depend on the tuning: fixed-point P2G
 * sums are integers, the float gather sums each node in a tile-independent
 * order (the one exception, a region bucket row holding more particles than
 * one gather chunk, needs ~8 particles per cell), and G2P is per particle.
 *
 * mpm_tuning fields:
 *   p2g_tile         P2G tile side in grid nodes: 28, 20 (default) or 12.
 *   g2p_tile         G2P tile side in 4x4-cell buckets: 8 (default) or 4
 *                    (default for a grid, or a batch's grids together,
 *                    with fewer 8-bucket tiles than the device has compute
 *                    units: those launches are latency-bound, and the
 *                    smaller tile gives each thread fewer particles).
 *   g2p_threads      threads of a G2P block: 128, 256 (default) or 512 for
 *                    8-bucket tiles; 128 or 256 for 4-bucket tiles.  Particles
 *                    per wavefront = the tile's particles / (g2p_threads/64).
 *   bucket_capacity  slots per bucket: 0 keeps the current capacity, else a
 *                    multiple of 16 in [32, 1024]; it still grows on demand.
 *   tile_queue       0 (automatic), 1 (on) or 2 (off).  On: a compaction
 *                    kernel lists each microstep the tiles that hold
 *                    particles, and P2G and G2P run as the blocks the device
 *                    holds at once, taking tiles from those lists, so the
 *                    empty tiles of a sparse grid cost no blocks.
 *                    Automatic: on when the grid nodes (of a batch, all of
 *                    them) exceed twice the particles and there are at
 *                    least four P2G tiles per compute unit.  Never used for
 *                    a retained fixed-point grid or bonded bodies.
 * mpm_set_tuning: a shape that is not compiled, or a bad capacity, returns
 *   MPM_EINVAL with no state change; the scene of a batch takes the batch's
 *   shape (mpm_batch_set_tuning) and returns MPM_EINVAL for another shape.
 * mpm_get_tuning: the shape the next step runs and the current capacity.
 *   Scenes with bonded triangle bodies or fracture bonds run 28-node P2G
 *   tiles and 256-thread G2P blocks, with the requested G2P tile side; a P2G
 *   tile whose gather chunk (448 particles for 12, 960 for 20) is below the
 *   bucket capacity runs the 28-node tile.
 * mpm_device_arch: the architecture name (e.g. "gfx950:sramecc+:xnack-")
 *   and compute-unit count of a device, the autotuner's cache key. */
typedef struct {
    uint32_t p2g_tile;
    uint32_t g2p_tile;
    uint32_t g2p_threads;
    uint32_t bucket_capacity;
    uint32_t tile_queue;
} mpm_tuning;
int mpm_set_tuning(mpm_ctx* ctx, const mpm_tuning* tuning);
int mpm_get_tuning(mpm_ctx* ctx, mpm_tuning* out);
int mpm_batch_set_tuning(mpm_batch* batch, const mpm_tuning* tuning);
int mpm_batch_get_tuning(mpm_batch* batch, mpm_tuning* out);
int mpm_device_arch(int32_t device, char* arch, uint32_t arch_bytes, uint32_t* compute_units);

/* Mean kernel time in microseconds of the named kernel over the last
 * mpm_step_batch calls made while timing was enabled (HIP events on the
 * context's stream). enable != 0 turns timing on and resets counters. */
int mpm_set_timing(mpm_ctx* ctx, uint32_t enable);
int mpm_read_timing(mpm_ctx* ctx, const char* kernel, double* mean_us, uint64_t* launches);

/* ---------------- differentiable rollout ----------------
 * No counterpart in the reference (it has no differentiation code).  A
 * fixed-corotated elastic solid — the snow stress of constitutiveStressTimesFt
 * (reference/solver.ts:2171-2203) at hardening factor 1 without the plastic
 * projection — stepped `steps` times by the solid branch of stepMpm
 * (reference/solver.ts:396-824) in float32 with an FP64 polar/stress, and the
 * hand-written reverse of that rollout, for callers that optimise initial
 * conditions or material parameters.
 *
 * State: MPM_DIFF_STATE_FLOATS floats per particle,
 *   x, y, vx, vy, C00, C01, C10, C11, F00, F01, F10, F11.
 * Every pointer argument may be host or device memory (copies use
 * hipMemcpyDefault); the calls return after the results are in place.
 *
 * mpm_diff_create: segment_steps is the checkpoint interval (0 = 20); the
 *   reverse pass recomputes one segment at a time from its checkpoint.
 * mpm_diff_forward: runs the rollout from state0 with per-particle lambda and
 *   mu, keeps the checkpoints, writes the final state (state_final may be
 *   NULL).
 * mpm_diff_backward: given dL/d(final state), writes dL/d(state0) and
 *   dL/dlambda, dL/dmu (any output may be NULL); needs a successful forward.
 *   The derivative through a confined axis (confineAxis) is that of the
 *   clamped branch: zero.
 * Numerical anomalies (status bits below, except CONFINED which is
 * informational) make forward/backward return MPM_ESTATE; the bits stay
 * readable through mpm_diff_read_status.
 * mpm_diff_read_timing: device time of the last forward and backward kernel
 *   sequences (HIP events on the rollout's stream, copies excluded). */
#define MPM_DIFF_STATE_FLOATS 12
#define MPM_DIFF_STENCIL 1u     /* a stencil left the grid (that particle was skipped) */
#define MPM_DIFF_DEFORMATION 2u /* det F <= 0 */
#define MPM_DIFF_NONFINITE 4u   /* non-finite velocity */
#define MPM_DIFF_SUBSTEPS 8u    /* more than 64 deformation substeps */
#define MPM_DIFF_CONFINED 16u   /* a particle was clamped to a wall */

#define MPM_DIFF_NO_GRAPHS 1u     /* config flag: launch kernel by kernel instead of
                                     replaying captured HIP graphs */

typedef struct mpm_diff mpm_diff;
typedef struct mpm_diff_config {
    uint32_t particle_count, cells_x, cells_y, steps, segment_steps;
    float cell_size, dt, particle_mass, particle_volume;
    float gravity_x, gravity_y, wall_friction;
    int32_t device; /* HIP device ordinal; -1 = current */
    uint32_t flags; /* MPM_DIFF_* config flags */
} mpm_diff_config;

int mpm_diff_create(const mpm_diff_config* cfg, mpm_diff** out);
int mpm_diff_forward(mpm_diff* d, const float* state0, const float* lambda, const float* mu, float* state_final);
int mpm_diff_backward(mpm_diff* d, const float* bar_final, float* bar_state0, float* grad_lambda, float* grad_mu);
int mpm_diff_read_status(mpm_diff* d, uint32_t* status);
int mpm_diff_read_timing(mpm_diff* d, double* forward_ms, double* backward_ms);
int mpm_diff_get_layout(mpm_diff* d, uint32_t* segment_steps, uint32_t* checkpoints, size_t* device_bytes);
void mpm_diff_destroy(mpm_diff* d);
/* Message of the last failed call on d (of the last failed mpm_diff_create
 * on this thread when d is NULL). Never NULL. */
const char* mpm_diff_last_error(mpm_diff* d);

/* dispose (solver.ts:2950-2964). Safe on NULL; ignores the scenes of a batch. */
void mpm_destroy(mpm_ctx* ctx);

/* Message of the last failed call on ctx (or of the last failed mpm_create
 * when ctx is NULL). Never NULL. */
const char* mpm_last_error(mpm_ctx* ctx);

#ifdef __cplusplus
}
#endif
#endif
