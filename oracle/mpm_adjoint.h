/*
 * FP64 differentiable rollout of the 2D MLS-MPM step for a fixed-corotated
 * elastic solid, with its hand-written reverse pass — TEST INFRASTRUCTURE
 * ONLY.  It is the checker of the GPU adjoint; the product never loads it.
 *
 * The forward step is the reference step `stepMpm`
 * (src/lib/mpm/reference/solver.ts:396-824) specialised to one solid phase:
 *   stencil       updateParticleStencil        solver.ts:2017-2034
 *   stress        snow branch of constitutiveStressTimesFt with hardening
 *                 factor 1 (plastic volume 1)   solver.ts:2171-2203
 *                 P F^T = 2 mu (F - R) F^T + lambda J (J - 1) I
 *   P2G           solver.ts:418-479
 *   grid update   updateGridVelocity + wall cushion / tangential friction
 *                                               solver.ts:1507-1524, 2542-2559
 *   G2P           solver.ts:610-773 (no plastic projection: the solid stays
 *                 elastic, which is what "fixed-corotated" means here)
 *   F update      updateDeformationGradient, with its substeps
 *                                               solver.ts:2561-2592
 *   confinement   confineAxis                   solver.ts:2528-2540
 * R is the polar rotation of F, written through the angle
 * theta = atan2(F10 - F01, F00 + F11); for det F > 0 it equals the U V^T of
 * decompose2x2 (solver.ts:2425-2463) and, unlike an SVD, its derivative has no
 * 1/(s1 - s2) term, so repeated singular values need no special case.
 *
 * Particle state: 12 doubles per particle,
 *   x, y, vx, vy, C00, C01, C10, C11, F00, F01, F10, F11.
 * Per-particle lambda and mu are parameters; mass and volume come from the
 * configuration (solids use particleMass, solver.ts:2130-2146).
 *
 * Parity pinning (what fixes these functions independently of this code):
 *   orc_diff_step:     equals orc_step_f64 for a snow particle with plastic
 *                      volume 1 whose singular values stay inside the snow
 *                      yield band (no projection) — tests/test_adjoint.py.
 *   orc_diff_backward: central finite differences of orc_diff_forward, and
 *                      the free-flight closed form (momentum is conserved
 *                      exactly by P2G/G2P, so d com(T)/d v0_p = T dt / N and
 *                      d com(T)/d lambda = d com(T)/d mu = 0 away from walls).
 */
#ifndef MPM_ADJOINT_H
#define MPM_ADJOINT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_DIFF_STATE 12

typedef struct {
    double dt, cell_size, particle_mass, particle_volume;
    double gravity_x, gravity_y, wall_friction;
    int32_t cells_x, cells_y, steps;
} orc_diff_config;

/* One forward step; out may not alias in.  Returns ORC_OK or a negative
 * oracle error code (mpm_oracle.h). */
int orc_diff_step(const orc_diff_config* c, int32_t n, const double* in, const double* lambda,
                  const double* mu, double* out);

/* c->steps forward steps from state0; writes the final state. */
int orc_diff_forward(const orc_diff_config* c, int32_t n, const double* state0, const double* lambda,
                     const double* mu, double* state_final);

/* Reverse pass of orc_diff_forward: given dL/d(final state) (12 per
 * particle), writes dL/d(state0) (12 per particle), dL/dlambda and dL/dmu
 * (one per particle).  `confined` (optional, int64) receives how many
 * particle-axis confinements the forward applied; their derivative is the
 * derivative of the clamped branch (zero). */
int orc_diff_backward(const orc_diff_config* c, int32_t n, const double* state0, const double* lambda,
                      const double* mu, const double* bar_final, double* bar_state0, double* grad_lambda,
                      double* grad_mu, int64_t* confined);

#ifdef __cplusplus
}
#endif
#endif
