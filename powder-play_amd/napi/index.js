'use strict';
// JavaScript face of the mpm355 Node-API addon: the native runtime beneath a
// TypeScript MpmMlsGpuSolver (src/lib/mpm/kernels/solver.ts:106-144).  The
// caller keeps the reference's host code (validateMpmMlsGpuInput,
// packMpmMlsGpuStepParameters, the CFL planning and the adaptive loop) and
// hands this class the packed words it would hand WebGPU.
const path = require('path');

const addon = require(path.join(__dirname, '..', 'mpm355.node'));

class MpmNativeRuntime {
	// config: { cellsX, cellsY, cellSize, particleMass, particleVolume,
	//           wallFriction?, capacity?, device?, bucketSlots?, keepFixedGrid? }
	// packed: Uint32Array of 20-word records (layout.ts:385-437)
	constructor(config, packed, objectIds, atomicP2g) {
		this.ctx = addon.create(config, packed, objectIds || null, Boolean(atomicP2g));
	}

	// stepBatch encode + submit (solver.ts:1407-1599); params: the 144-byte block
	stepBatch(params, stepCount, measureSpeed) {
		addon.stepBatch(this.ctx, params, stepCount, Boolean(measureSpeed));
	}

	// { maxSpeed, completed, status } once the queued work is done (solver.ts:2388-2512)
	readCompact() {
		return addon.readCompact(this.ctx);
	}

	clearStatus() {
		addon.clearStatus(this.ctx);
	}

	counts() {
		return addon.counts(this.ctx);
	}

	readParticles(indices) {
		return addon.readParticles(this.ctx, indices || null);
	}

	appendParticles(packed, objectIds) {
		addon.appendParticles(this.ctx, packed, objectIds || null);
	}

	setInteraction(uniform) {
		addon.setInteraction(this.ctx, uniform);
	}

	// removeParticlesInOutlet (solver.ts:2280-2386): Float32Array [minX, minY, maxX, maxY] -> count
	removeParticlesInOutlet(bounds) {
		return addon.removeParticlesInOutlet(this.ctx, bounds);
	}

	// gatherSurfaceOccupancy (shaders.ts:710-806): Float32Array (water, oil) per node
	gatherOccupancy() {
		return addon.gatherOccupancy(this.ctx);
	}

	// pick (solver.ts:2733-2823): Uint32Array [index | 0xFFFFFFFF, x bits, y bits, objectId]
	pick(x, y, radius) {
		return addon.pick(this.ctx, x, y, radius);
	}

	// clone (solver.ts:2852-2948): false where the reference refuses
	clone(start, count, offsetX, offsetY, objectId) {
		return addon.clone(this.ctx, start, count, offsetX, offsetY, objectId);
	}

	setDiagnostics(enable) {
		addon.setDiagnostics(this.ctx, Boolean(enable));
	}

	readDiagnostics() {
		return addon.readDiagnostics(this.ctx);
	}

	// config.fractureBonds as 6-word records (solver.ts:929-935), and their state
	setFractureBonds(records) {
		addon.setFractureBonds(this.ctx, records);
	}

	readFractureBonds(count) {
		return addon.readFractureBonds(this.ctx, count);
	}

	// launch shape and bucket capacity ({ p2gTile, g2pTile, g2pThreads,
	// bucketCapacity? }); results do not depend on them
	setTuning(tuning) {
		addon.setTuning(this.ctx, tuning);
	}

	getTuning() {
		return addon.getTuning(this.ctx);
	}
}

module.exports = { addon, MpmNativeRuntime };
