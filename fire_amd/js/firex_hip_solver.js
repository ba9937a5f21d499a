'use strict';
/**
 * FireXHipSolver — the compute half of the reference's FireXEngine
 * (src/sim/firex/FireXEngine.ts:370) on an MI355X, over the Node-API addon
 * firex_napi.node and libfirex_hip.so.
 *
 * It keeps the reference's solver API: construct with a scene and grid,
 * setParameter (FireXEngine.ts:937-1032), initialize (:506-547), reset
 * (:843-852), renderOfflineFrame(dt, present = false, signal) (:643-697),
 * frame(t, dt) (:629-641), getGridInfo (:478-490), dispose (:1034-1062),
 * plus field readback (new surface).  Presentation (present = true) is not
 * part of this solver: rendering stays on the reference path.
 *
 * CommonJS, no dependencies; runs on Node >= 12 (N-API 8).
 */
const path = require('path');

const native = require(path.join(__dirname, '..', 'firex_napi.node'));

const FIELD_COMPONENTS = {
  flow: 4, species: 4, flowPredicted: 4, pressure: 1, correction: 1, divergence: 1, reaction: 1, liquid: 1,
  packedFlow: 4, packedChemistry: 4, packedMedia: 4,
};
// per-particle fields (records per SPH droplet, not per cell)
const PARTICLE_COMPONENTS = { particles: 8, particleDensity: 4 };
const REDUCTIONS = {
  sumFuel: 0, sumOxygen: 1, sumProducts: 2, sumVapor: 3, sumTemperature: 4, sumReaction: 5,
  residualL2: 6, flowDivergenceL2: 7,
};

function abortError(message) {
  const error = new Error(message || 'The Fire-X offline render was aborted.');
  error.name = 'AbortError';
  return error;
}

class FireXHipSolver {
  /**
   * @param {object} options  tier ('low'..'cinematic' | 'hd' | 'qhd' | 'uhd'),
   *   dims, pressureIterations, correctionIterations, timeStep,
   *   maximumSubsteps, speciesStorage ('f32' | 'f16' | 'bf16'),
   *   heavyMode, plan ('reference' | 'fused'), device, memoryBudgetBytes,
   *   particleCount (SPH droplets; the tier's count by default),
   *   scene (a { key: value } map applied with setParameter).
   */
  constructor(options = {}) {
    this.options = Object.assign({ tier: 'balanced', plan: 'fused' }, options);
    this.backend = 'HIP · gfx950 thermochemical grid';
    this.handle = null;
    this.disposed = false;
    this.pending = [];
    if (options.scene) {
      for (const key of Object.keys(options.scene)) this.setParameter(key, options.scene[key]);
    }
  }

  async initialize() {
    if (this.handle) return;
    if (this.disposed) throw new Error('The Fire-X engine has already been disposed.');
    this.handle = native.create(this.options);
    for (const [name, value] of this.pending) native.setParameter(this.handle, name, value);
    this.pending = [];
    native.reset(this.handle);  // the reset state follows the scene's oxygenRate
  }

  setParameter(name, value) {
    if (this.handle) native.setParameter(this.handle, name, value);
    else this.pending.push([name, value]);
  }

  getParameter(name) {
    return native.getParameter(this.requireHandle(), name);
  }

  setEmitters(emitters, jitter = 0, velocityNoise = 0, seed = 0) {
    native.setEmitters(this.requireHandle(), emitters, jitter, velocityNoise, seed);
  }

  reset() {
    if (this.handle) native.reset(this.handle);
  }

  /** Live frame: fixed-step clock, returns the host milliseconds spent. */
  frame(timeSeconds, deltaSeconds) {
    void timeSeconds;
    if (!this.handle || this.disposed) return 0;
    const start = Date.now();
    native.frame(this.handle, deltaSeconds);
    return Date.now() - start;
  }

  /** step(dt): count fixed steps of dt seconds, enqueued on the device. */
  step(dt, count = 1) {
    native.step(this.requireHandle(), dt, count);
  }

  async renderOfflineFrame(deltaSeconds, present = false, signal) {
    if (!Number.isFinite(deltaSeconds) || deltaSeconds < 0) {
      throw new Error('Offline frame delta must be finite and non-negative.');
    }
    if (present) throw new Error('Presentation stays on the reference renderer; call with present = false.');
    if (signal && signal.aborted) throw abortError();
    const flag = new Int32Array(1);
    if (signal && typeof signal.addEventListener === 'function') {
      signal.addEventListener('abort', () => { flag[0] = 1; }, { once: true });
    }
    native.renderOfflineFrame(this.requireHandle(), deltaSeconds, flag);
  }

  sync() {
    native.sync(this.requireHandle());
  }

  /** Copies a field to a new (or the given) Float32Array / Float64Array / Uint32Array. */
  readField(name, out) {
    if (!(name in FIELD_COMPONENTS) && !(name in PARTICLE_COMPONENTS) && name !== 'scene' && name !== 'frame') {
      throw new RangeError(`Unknown Fire-X field: ${name}`);
    }
    const layout = native.fieldLayout(this.requireHandle(), name);
    const length = layout.records * layout.components;
    const target = out || (name === 'liquid' ? new Uint32Array(length) : new Float32Array(length));
    return native.readField(this.handle, name, target);
  }

  writeField(name, data) {
    native.writeField(this.requireHandle(), name, data);
  }

  /** The cinematic scene pass's volume ray marcher (qualityShaders.ts:374-575): packs the render
   *  volumes and renders width x height rgba16f pixels, read back as readField('scene'). */
  renderVolume(width, height, raySteps = 0) {
    native.renderVolume(this.requireHandle(), { width, height, raySteps });
  }

  /** A presented frame (ray march, bloom and filmic composite with the cinematic presentation): a new
   *  Uint8Array of width x height rgba8 pixels, row 0 at the top. */
  renderFrame(width, height, raySteps = 0) {
    native.renderFrame(this.requireHandle(), { width, height, raySteps });
    return this.readFieldRaw('frame');
  }

  /** Stored format of a field: {kind: 'f32'|'f16'|'bf16'|'u32'|'u8', components, records, storedBytes}. */
  fieldLayout(name) {
    return native.fieldLayout(this.requireHandle(), name);
  }

  /** The field's stored bytes, verbatim (a new Uint8Array, or the given one). */
  readFieldRaw(name, out) {
    const layout = native.fieldLayout(this.requireHandle(), name);
    return native.readFieldRaw(this.handle, name, out || new Uint8Array(layout.storedBytes));
  }

  /** 64-bit checksum of the field's stored bytes, computed on the device, as 16 hex digits. */
  fieldChecksum(name) {
    return native.fieldChecksum(this.requireHandle(), name);
  }

  /** packFields (qualityShaders.ts:37-73): fill the three render volumes
   *  (readField('packedFlow' | 'packedChemistry' | 'packedMedia')). */
  packFields() {
    native.packFields(this.requireHandle());
  }

  reduce(name) {
    if (!(name in REDUCTIONS)) throw new RangeError(`Unknown Fire-X reduction: ${name}`);
    return native.reduce(this.requireHandle(), REDUCTIONS[name]);
  }

  getGridInfo() {
    const info = native.getInfo(this.requireHandle());
    const domain = [
      native.getParameter(this.handle, 'domainWidth'),
      native.getParameter(this.handle, 'domainHeight'),
      native.getParameter(this.handle, 'domainDepth'),
    ];
    const estimate = native.estimateMemory(info.dimensions, 0);
    return {
      dimensions: info.dimensions,
      domainMeters: domain,
      cellSizeMeters: domain.map((d, i) => d / info.dimensions[i]),
      cellCount: info.cellCount,
      estimatedBytes: estimate.totalBytes,
      deviceBytes: info.deviceBytes,
      simulationTime: info.simulationTime,
      frameIndex: info.frameIndex,
      custom: this.options.dims !== undefined,
    };
  }

  dispose() {
    if (this.disposed) return;
    this.disposed = true;
    if (this.handle) native.destroy(this.handle);
    this.handle = null;
  }

  requireHandle() {
    if (this.disposed) throw new Error('The Fire-X engine has already been disposed.');
    if (!this.handle) throw new Error('The Fire-X solver is not initialized.');
    return this.handle;
  }
}

/** estimateFireXGridMemory, src/sim/firex/gridConfiguration.ts:272-291. */
function estimateFireXGridMemory(dimensions, particleCount = 0) {
  return native.estimateMemory(dimensions, particleCount);
}

/** splitOfflineSimulationDelta, src/sim/offlineTimeline.ts:5-20. */
function splitOfflineSimulationDelta(deltaSeconds, maximumStepSeconds) {
  return native.splitOfflineDelta(deltaSeconds, maximumStepSeconds);
}

module.exports = {
  FireXHipSolver, estimateFireXGridMemory, splitOfflineSimulationDelta, native, REDUCTIONS, FIELD_COMPONENTS,
};
