// Type declarations for firex_hip_solver.js (the TypeScript view of the
// MI355X Fire-X solver, drop-in for FireXEngine's compute path,
// src/sim/firex/FireXEngine.ts:370).

export type FireXFieldName =
  | 'flow' | 'species' | 'flowPredicted' | 'pressure' | 'correction' | 'divergence' | 'reaction' | 'liquid'
  | 'packedFlow' | 'packedChemistry' | 'packedMedia'
  /** per SPH droplet: 8 (positionAge, velocityMass) and 4 (density, near density, pressures) values */
  | 'particles' | 'particleDensity'
  /** the last renderVolume image: width x height pixels of 4 values (rgba16f), row 0 at the top */
  | 'scene'
  /** the last renderFrame: width x height rgba8 pixels */
  | 'frame';

export type FireXReductionName =
  | 'sumFuel' | 'sumOxygen' | 'sumProducts' | 'sumVapor' | 'sumTemperature' | 'sumReaction'
  | 'residualL2' | 'flowDivergenceL2';

export interface FireXHipSolverOptions {
  tier?: 'low' | 'balanced' | 'high' | 'maximum' | 'cinematic' | 'hd' | 'qhd' | 'uhd';
  dims?: readonly [number, number, number];
  pressureIterations?: number;
  correctionIterations?: number;
  timeStep?: number;
  maximumSubsteps?: number;
  speciesStorage?: 'f32' | 'f16' | 'bf16';
  heavyMode?: 'auto' | 'off' | 'f16' | 'f32';
  plan?: 'reference' | 'fused';
  cinematic?: boolean;
  device?: number;
  memoryBudgetBytes?: number;
  jacobiSweepsPerLaunch?: number;
  /** 1: time candidate Jacobi plans on the device at creation and keep the fastest (cached per shape). */
  jacobiAutotune?: number;
  /** 1: replay everything after simulate as a captured hipGraph. */
  captureGraphs?: number;
  /** SPH droplets of the resource set (384-4096 by tier, FireXEngine.ts:187-277). */
  particleCount?: number;
  /** Ray-march samples per pixel (raySteps by tier, 80-384). */
  raySteps?: number;
  scene?: Readonly<Record<string, number | boolean>>;
}

export interface FireXHipGridInfo {
  dimensions: [number, number, number];
  domainMeters: [number, number, number];
  cellSizeMeters: [number, number, number];
  cellCount: number;
  estimatedBytes: number;
  deviceBytes: number;
  simulationTime: number;
  frameIndex: number;
  custom: boolean;
}

export interface FireXFieldLayout {
  kind: 'f32' | 'f16' | 'bf16' | 'u32' | 'u8';
  kindCode: number;
  components: number;
  records: number;
  storedBytes: number;
}

export declare class FireXHipSolver {
  readonly backend: string;
  constructor(options?: FireXHipSolverOptions);
  initialize(): Promise<void>;
  setParameter(name: string, value: number | boolean): void;
  getParameter(name: string): number;
  setEmitters(
    emitters: ReadonlyArray<readonly [number, number, number, number, number]>,
    jitter?: number, velocityNoise?: number, seed?: number,
  ): void;
  reset(): void;
  frame(timeSeconds: number, deltaSeconds: number): number;
  step(dt: number, count?: number): void;
  renderOfflineFrame(deltaSeconds: number, present?: false, signal?: AbortSignal): Promise<void>;
  sync(): void;
  readField<T extends Float32Array | Float64Array | Uint32Array = Float32Array>(name: FireXFieldName, out?: T): T;
  writeField(name: FireXFieldName, data: Float32Array | Float64Array | Uint32Array): void;
  reduce(name: FireXReductionName): number;
  /** packFields: fill the render volumes read back as packedFlow / packedChemistry / packedMedia. */
  packFields(): void;
  /** Volume ray marcher into the rgba16f scene image (readField('scene')). */
  renderVolume(width: number, height: number, raySteps?: number): void;
  /** A presented frame: width x height rgba8 pixels. */
  renderFrame(width: number, height: number, raySteps?: number): Uint8Array;
  /** Stored format of a field on the device. */
  fieldLayout(name: FireXFieldName): FireXFieldLayout;
  /** The field's stored bytes, verbatim. */
  readFieldRaw(name: FireXFieldName, out?: Uint8Array): Uint8Array;
  /** 64-bit checksum of the field's stored bytes (computed on the device), 16 hex digits. */
  fieldChecksum(name: FireXFieldName): string;
  getGridInfo(): FireXHipGridInfo;
  dispose(): void;
}

export declare function estimateFireXGridMemory(
  dimensions: readonly [number, number, number], particleCount?: number,
): {
  cellCount: number; gridBytes: number; particleBytes: number; fixedBytes: number; totalBytes: number;
  largestStorageBufferBytes: number;
};

export declare function splitOfflineSimulationDelta(deltaSeconds: number, maximumStepSeconds: number): number[];
