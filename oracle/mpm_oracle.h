/*
 * CPU oracle for the 2D MLS-MPM step path — TEST INFRASTRUCTURE ONLY.
 *
 * This library is the checker, never the product: only tests/, the smoke
 * check and bench.py's cpu_baseline leg load it.  It restates, in plain C:
 *
 *   1. orc_step_f64    — the FP64 reference step `stepMpm`
 *                        (src/lib/mpm/reference/solver.ts:396-824) for the
 *                        five constitutive models, two liquid phases,
 *                        water/oil grid contact, wall cushion and friction.
 *   2. orc_prepare_f32 / orc_fixed_p2g_f32 — the float32 per-particle
 *                        stress/affine cache and the fixed-point (2^24) int32
 *                        P2G scatter of the WGSL kernels
 *                        (src/lib/mpm/kernels/shaders.ts:808-1070), written
 *                        as an explicitly ordered f32 op sequence so the HIP
 *                        integer path can be compared bit for bit.
 *   3. orc_bin        — stable counting sort of particles by stencil base
 *                        node (shaders.ts:657-693 binning, made deterministic).
 *
 * Parity pinning (what fixes each function independently of this code):
 *   orc_step_f64:   closed forms / invariants from reference/solver.spec.ts
 *                   (weights partition of unity, rest state, uniform
 *                   translation, wall clipping, snow/NH/powder/liquid clamps,
 *                   mass and momentum conservation, water–oil contact
 *                   momentum); see tests/test_oracle.py.
 *   orc_fixed_p2g:  hand-computed integer contributions for 1–3 particles,
 *                   order-independence, and decode == FP64 P2G within the
 *                   2^-24 quantum; see tests/test_oracle.py.
 *   orc_bin:        brute-force O(N * cells) scan on small inputs.
 *   orc_step_bodies_f64: bonded triangle bodies and fracture bonds (config 2),
 *                   pinned by closed forms of the triangle distance field,
 *                   the SAT pair contact, the PBD edge solve and the bond
 *                   damage law, and by momentum bookkeeping; see
 *                   tests/test_oracle_bodies.py.  orc_step_f64 itself still
 *                   rejects bonded vertices (ORC_EUNSUPPORTED).
 */
#ifndef MPM_ORACLE_H
#define MPM_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    ORC_MODEL_NEO_HOOKEAN = 0,
    ORC_MODEL_SNOW = 1,
    ORC_MODEL_WATER = 2,
    ORC_MODEL_OIL = 3,
    ORC_MODEL_POWDER = 4,
    ORC_MODEL_REMOVED = -1
};

enum {
    ORC_OK = 0,
    ORC_EINVALID = -1,      /* input validation failed (RangeError/TypeError) */
    ORC_ENUMERIC = -2,      /* a computed quantity is non-finite or out of range */
    ORC_EDETERMINANT = -3,  /* non-positive deformation determinant */
    ORC_ESUBSTEP = -4,      /* deformation update exceeds 64 substeps */
    ORC_EUNSUPPORTED = -5
};

/* One particle in double precision. `param` carries the model parameter the
 * same way the 20-word GPU record does (layout.ts:385-437): Tait exponent for
 * liquids (0 = linear tension-free EOS), friction coefficient for powder,
 * damage for Neo-Hookean; unused for snow. */
typedef struct {
    double x, y, vx, vy;
    double c00, c01, c10, c11;
    double f00, f01, f10, f11;
    double plastic_volume;
    double lambda, mu, density, volume, param;
    int32_t model;
    int32_t bonded_vertex; /* 0 = none; nonzero is unsupported here */
} orc_particle;

typedef struct {
    double dt;
    int32_t cells_x, cells_y;
    double cell_size;
    double particle_mass, particle_volume;
    double gravity_x, gravity_y;
    double wall_friction;
    /* Drag interaction (the uniform of writeInteractionParameters,
     * kernels/solver.ts:1389-1401; applied by finishGridVelocity,
     * kernels/shaders.ts:1748-1777): target point, radius (0 = off),
     * spring stiffness and damping.  The CPU reference step has no drag. */
    double drag_x, drag_y, drag_radius, drag_stiffness, drag_damping;
} orc_config;

/* inputMass, gridMass, outputMass, inputMomentum(x,y), gridMomentum(x,y),
 * outputMomentum(x,y) — MpmStepDiagnostics (reference/solver.ts:96-103). */
typedef struct {
    double input_mass, grid_mass, output_mass;
    double input_px, input_py, grid_px, grid_py, output_px, output_py;
} orc_diagnostics;

/* FP64 explicit MLS-MPM step. Returns ORC_OK or a negative error code;
 * `out` is only meaningful on ORC_OK. */
int orc_step_f64(const orc_particle* in, int32_t n, const orc_config* cfg,
                 orc_particle* out, orc_diagnostics* diag);

/* Config-2 extras of the step (solver.ts:80-94): bonded triangle bodies use
 * bond_rest_length, bond_stiffness (reference default 0.35) and
 * bond_reference_dt (reference default: dt). */
typedef struct {
    double bond_rest_length, bond_stiffness, bond_reference_dt;
} orc_body_config;

/* A fracture bond (MpmSolidBond, solver.ts:61-67); damage and broken are
 * updated in place by orc_step_bodies_f64. */
typedef struct {
    int32_t first, second;
    double rest_length, failure_strain, damage;
    int32_t broken, pad;
} orc_bond;

/* orc_step_f64 plus bonded triangle bodies and fracture bonds (stepMpm with
 * buildTriangleField, solveFluidTriangleGridContact,
 * projectParticleOutsideTriangles, applyTriangleContactImpulses,
 * enforceBondedTriangleConstraints, resolveBondedTriangleContacts and
 * updateMpmSolidBondDamage, solver.ts:1023-1874, solid-bonds.ts:68-108).
 * Particles with bonded_vertex = vertex + 1 form triangles of three
 * consecutive particles of one object; object_ids has one entry per particle. */
int orc_step_bodies_f64(const orc_particle* in, int32_t n, const uint32_t* object_ids, const orc_config* cfg,
                        const orc_body_config* bodies, orc_bond* bonds, int32_t nbonds, orc_particle* out,
                        orc_diagnostics* diag);

/* The two geometric primitives of the contact code, exposed for pinning:
 * triangleDistanceSample (solver.ts:1190-1291): out6 = signed distance,
 * normal x, y, barycentric of the closest point (3); and trianglePairContact
 * (:1804-1874): returns 1 with out3 = normal x, y (from b towards a),
 * penetration, or 0 without contact. */
int orc_triangle_sample(double px, double py, const double* tri6, double* out6);
int orc_triangle_pair_contact(const double* a6, const double* b6, double skin, double* out3);

/* Run `steps` FP64 steps in place (buffer `p`); returns the first error. */
int orc_run_f64(orc_particle* p, int32_t n, const orc_config* cfg, int32_t steps);

/* orc_run_f64 on `threads` OpenMP threads: the all-core CPU baseline of the
 * benchmark, never a checker.  P2G sums follow bands of grid rows instead of
 * the particle order, so results equal orc_run_f64's to rounding
 * (tests/test_oracle.py compares them). */
int orc_run_f64_parallel(orc_particle* p, int32_t n, const orc_config* cfg, int32_t steps, int32_t threads);

/* FP64 grid after P2G + grid update of one step (for component checks).
 * node_out has 17 doubles per node (nodeCount = (cx+3)*(cy+3)):
 * [mass, qx, qy, vx, vy, wm, wqx, wqy, wvx, wvy, wV, om, oqx, oqy, ovx, ovy, oV]. */
int orc_grid_f64(const orc_particle* in, int32_t n, const orc_config* cfg,
                 double* node_out);

/* ---------------- float32 GPU-semantics integer path ---------------- */

/* The 144-byte step-parameter block as floats/ints, only the fields the
 * integer path reads (layout.ts:896-997). */
typedef struct {
    uint32_t cells_x, cells_y, node_width, node_height;
    uint32_t particle_count, node_count;
    float cell_size, particle_mass, particle_volume, dt;
    float gradient_scale;
} orc_params_f32;

/* Per-particle f32 affine A = m*C + (-dt*V*4/dx^2) * P*F^T
 * (shaders.ts:808-963; liquid EOS in f32, solid stresses in FP64 rounded to
 * f32). packed20: 20-word AoS records. Writes affine (4 floats/particle).
 * Returns the status flag bits the WGSL kernel would raise (0 on success). */
uint32_t orc_prepare_f32(const uint32_t* packed20, uint32_t n,
                         const orc_params_f32* prm, float* affine_out);

/* Fixed-point P2G: 11 int32 words per node
 * [solid m,qx,qy | water m,qx,qy,V | oil m,qx,qy,V], scale 2^24, normalised
 * by particleMass / particleVolume (shaders.ts:971-1043).  `grid_out` must
 * hold node_count*11 words and is overwritten.  Overflow is flagged (65536)
 * when a single contribution exceeds 2147483520 or a word's exact sum leaves
 * the int32 range.  Returns status flags. */
uint32_t orc_fixed_p2g_f32(const uint32_t* packed20, uint32_t n,
                           const orc_params_f32* prm, int32_t* grid_out);

/* Decode of the int32 words into f32 node mass/momentum/volume exactly as
 * gatherAtomicGrid (shaders.ts:1045-1070): f32(word) / 2^24 * scale.
 * node_out: 11 floats per node, same order as the words. */
void orc_decode_fixed_f32(const int32_t* grid, uint32_t node_count,
                          const orc_params_f32* prm, float* node_out);

/* Stable counting sort by stencil base node gridIndex(base)
 * (shaders.ts:518-520). key_out[i] = node key or 0xFFFFFFFF (removed /
 * outside the halo); start_out has node_count+1 entries; perm_out lists
 * particle indices grouped by key, ascending index within a key. Returns the
 * stencil error flag (1) if any particle's stencil leaves the halo. */
uint32_t orc_bin(const uint32_t* packed20, uint32_t n, const orc_params_f32* prm,
                 uint32_t* key_out, uint32_t* start_out, uint32_t* perm_out);

#ifdef __cplusplus
}
#endif
#endif
