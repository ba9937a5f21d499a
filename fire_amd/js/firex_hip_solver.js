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
 * the configuration setters setGridMemoryBudget (:929-935), setQuality
 * (:863-893), setGridDimensions (:895-927), setOfflineRenderTier (:799-831)
 * and setOfflineOpticalDetailTarget (:834-841), serialised like the
 * reference's resourceMutationQueue (:1882-1889) and rebuilding the device
 * buffers transactionally (firex_reconfigure), plus field readback (new
 * surface).  Presentation (present = true) is not part of this solver:
 * rendering stays on the reference path.
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

// Tier names (FIREX_QUALITY and FIREX_OFFLINE_QUALITY, FireXEngine.ts:187-277)
// and the offline tiers' minimum transition guards (minimumMemoryGuardBytes).
const QUALITIES = ['low', 'balanced', 'high', 'maximum', 'cinematic'];
const OFFLINE_TIERS = ['hd', 'qhd', 'uhd'];
const MIB = 1024 * 1024;
const OFFLINE_MINIMUM_GUARD = { hd: 512 * MIB, qhd: 1024 * MIB, uhd: 2048 * MIB };
// gridConfiguration.ts:111, 118-128: the default budget and the largest guard.
const DEFAULT_GRID_MEMORY_BUDGET_BYTES = 512 * MIB;
const MAX_GRID_MEMORY_BUDGET_BYTES = 32 * 1024 * MIB;

/** formatFireXBytes, gridConfiguration.ts:372-399. */
function formatBytes(bytes) {
  const decimal = (value, digits) => value.toFixed(digits).replace(/\.0+$|(\.[0-9]*?)0+$/, '$1');
  if (!Number.isFinite(bytes) || bytes < 0) return '\u2014';
  const mebibytes = bytes / MIB;
  if (mebibytes < 1) return `${decimal(bytes / 1024, 1)} KiB`;
  const gibibytes = mebibytes / 1024;
  if (gibibytes >= 1024) {
    const tebibytes = gibibytes / 1024;
    return `${decimal(tebibytes, tebibytes >= 10 ? 1 : 2)} TiB`;
  }
  if (gibibytes >= 1) return `${decimal(gibibytes, gibibytes >= 10 ? 1 : 2)} GiB`;
  return `${decimal(mebibytes, mebibytes >= 10 ? 1 : 2)} MiB`;
}

function sameDimensions(a, b) {
  return Boolean(a && b) && a.length === 3 && b.length === 3 && a.every((value, index) => value === b[index]);
}

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
    // The engine's configuration state (FireXEngine.ts:405-421).  A live
    // quality tier takes options.dims as its grid override (exact); an
    // offline tier takes them as an exact grid (an addition: the reference's
    // offline override is a per-axis minimum, resolveFireXOfflineDimensions).
    const tier = this.options.tier;
    if (!QUALITIES.includes(tier) && !OFFLINE_TIERS.includes(tier)) {
      throw new RangeError(`Unknown Fire-X tier "${tier}".`);
    }
    this.quality = QUALITIES.includes(tier) ? tier : 'cinematic';
    this.offlineRenderTier = OFFLINE_TIERS.includes(tier) ? tier : null;
    this.gridOverride = QUALITIES.includes(tier) && options.dims ? [...options.dims] : null;
    this.exactDimensions = OFFLINE_TIERS.includes(tier) && options.dims ? [...options.dims] : null;
    this.gridMemoryBudgetBytes = DEFAULT_GRID_MEMORY_BUDGET_BYTES;
    this.offlineOpticalDetailTarget = null;
    this.rebuilding = false;
    this.resourceMutationQueue = Promise.resolve();
    if (options.scene) {
      for (const key of Object.keys(options.scene)) this.setParameter(key, options.scene[key]);
    }
  }

  async initialize() {
    if (this.handle) return;
    if (this.disposed) throw new Error('The Fire-X engine has already been disposed.');
    this.handle = native.create(this.nativeOptions());
    for (const [name, value] of this.pending) native.setParameter(this.handle, name, value);
    this.pending = [];
    native.reset(this.handle);  // the reset state follows the scene's oxygenRate
  }

  // ---- configuration (FireXEngine.ts:799-935) ------------------------------

  /** setGridMemoryBudget: the transactional-rebuild guard, clamped to [256 MiB, 32 GiB]. */
  setGridMemoryBudget(bytes) {
    if (!Number.isFinite(bytes) || bytes <= 0) return;
    this.gridMemoryBudgetBytes = Math.max(256 * MIB, Math.min(MAX_GRID_MEMORY_BUDGET_BYTES, Math.floor(bytes)));
  }

  /** setQuality: a live quality tier; rebuilds when its grid or droplet count differs. */
  setQuality(quality) {
    return this.enqueueResourceMutation(async () => {
      if (this.offlineRenderTier !== null) {
        throw new Error('Compute quality cannot change while an exact offline solver bundle is active.');
      }
      if (!QUALITIES.includes(quality)) throw new RangeError(`Unknown Fire-X quality "${quality}".`);
      if (quality === this.quality) return;
      const previous = this.quality;
      this.quality = quality;
      let rebuilt = false;
      this.rebuilding = true;
      try {
        if (this.handle) rebuilt = this.applyConfiguration(this.gridMemoryBudgetBytes);
      } catch (error) {
        this.quality = previous;
        throw error;
      } finally {
        this.rebuilding = false;
      }
      if (!rebuilt && this.handle) this.reset();
    });
  }

  /** setGridDimensions: the live grid override (null: the quality's grid). */
  setGridDimensions(dimensions) {
    const next = dimensions ? [dimensions[0], dimensions[1], dimensions[2]] : null;
    return this.enqueueResourceMutation(async () => {
      if (this.offlineRenderTier !== null) {
        throw new Error('Grid dimensions cannot change while an exact offline solver bundle is active.');
      }
      const previous = this.gridOverride;
      if (previous === next || sameDimensions(previous, next)) return;
      const target = next || this.tierConfiguration(this.quality).dimensions;
      if (!this.handle || sameDimensions(native.getInfo(this.handle).dimensions, target)) {
        this.gridOverride = next;
        return;
      }
      this.rebuilding = true;
      try {
        this.gridOverride = next;
        this.applyConfiguration(this.gridMemoryBudgetBytes);
      } catch (error) {
        this.gridOverride = previous;
        throw error;
      } finally {
        this.rebuilding = false;
      }
    });
  }

  /** setOfflineRenderTier: an exact offline solver bundle (null: back to the live quality). */
  setOfflineRenderTier(tier) {
    return this.enqueueResourceMutation(async () => {
      if (tier === this.offlineRenderTier) return;
      if (tier !== null && !OFFLINE_TIERS.includes(tier)) throw new RangeError(`Unknown offline tier "${tier}".`);
      const previousTier = this.offlineRenderTier;
      const previousExact = this.exactDimensions;
      const transitionGuard = Math.max(
        this.gridMemoryBudgetBytes,
        previousTier ? OFFLINE_MINIMUM_GUARD[previousTier] : 0,
        tier ? OFFLINE_MINIMUM_GUARD[tier] : 0,
      );
      this.rebuilding = true;
      try {
        this.offlineRenderTier = tier;
        this.exactDimensions = null;
        if (this.handle) this.applyConfiguration(transitionGuard);
      } catch (error) {
        this.offlineRenderTier = previousTier;
        this.exactDimensions = previousExact;
        throw error;
      } finally {
        this.rebuilding = false;
      }
    });
  }

  /** setOfflineOpticalDetailTarget: the offline render-frequency target, snapped to 0 | 256 | 512 | 1024. */
  setOfflineOpticalDetailTarget(target) {
    this.offlineOpticalDetailTarget = target === null
      ? null
      : target >= 768 ? 1024 : target >= 384 ? 512 : target >= 192 ? 256 : 0;
    if (this.offlineOpticalDetailTarget !== null) this.setParameter('opticalDetailTarget', this.offlineOpticalDetailTarget);
  }

  /** The grid the next build uses (effectiveGridDimensions, FireXEngine.ts:2800-2805). */
  effectiveGridDimensions() {
    if (this.exactDimensions) return [...this.exactDimensions];
    if (this.offlineRenderTier !== null) return native.resolveOfflineDims(this.offlineRenderTier, this.gridOverride);
    return this.gridOverride ? [...this.gridOverride] : this.tierConfiguration(this.quality).dimensions;
  }

  tierConfiguration(tier) {
    return native.configFromTier(tier);
  }

  /** Options for the native solver: the active tier (activeConfiguration, :2807-2811) on the effective grid. */
  nativeOptions() {
    const options = Object.assign({}, this.options, {
      tier: this.offlineRenderTier !== null ? this.offlineRenderTier : this.quality,
      dims: this.effectiveGridDimensions(),
    });
    delete options.scene;
    return options;
  }

  /** Applies the configuration to the live solver; true when the device buffers were rebuilt. */
  applyConfiguration(guardBytes) {
    return native.reconfigure(this.handle, this.nativeOptions(), guardBytes);
  }

  enqueueResourceMutation(operation) {
    const queued = this.resourceMutationQueue.then(operation);
    this.resourceMutationQueue = queued.then(() => undefined, () => undefined);
    return queued;
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
    if (!this.handle || this.disposed || this.rebuilding) return 0;
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
    if (this.disposed) throw new Error('The Fire-X engine has already been disposed.');
    const info = this.handle
      ? native.getInfo(this.handle)
      : { dimensions: this.effectiveGridDimensions(), cellCount: 0, deviceBytes: 0, simulationTime: 0, frameIndex: 0 };
    if (!this.handle) info.cellCount = info.dimensions[0] * info.dimensions[1] * info.dimensions[2];
    const parameter = (name, fallback) => (this.handle ? native.getParameter(this.handle, name) : fallback);
    const domain = [parameter('domainWidth', 1.15), parameter('domainHeight', 1.35), parameter('domainDepth', 1.15)];
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
      custom: this.gridOverride !== null || this.exactDimensions !== null,
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

/** resolveFireXOfflineDimensions, FireXEngine.ts:279-290. */
function resolveFireXOfflineDimensions(tier, override) {
  return native.resolveOfflineDims(tier, override || null);
}

module.exports = {
  FireXHipSolver, estimateFireXGridMemory, splitOfflineSimulationDelta, resolveFireXOfflineDimensions, formatBytes,
  native, REDUCTIONS, FIELD_COMPONENTS,
};
