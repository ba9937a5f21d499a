'use strict';
/**
 * record_fire_hip — offline capture driver for the MI355X Fire-X solver.
 *
 * The solver half of the reference's recorder (scripts/record-fire.mjs:341-446
 * and the capture page it drives, src/cli/captureApp.ts:124-193): the same job
 * flow — resolve the exact grid, apply a scene preset and field view, warm
 * up, advance an exact 60 fps offline timeline with renderOfflineFrame(1/60),
 * verify from disk, then install the result and its `.json` manifest
 * together from `.partial` files — with an auditable field record in place
 * of the encoded video:
 *
 *   * every frame: a device checksum of each audited field (firex_field_checksum)
 *     and the FP64 residual, divergence and conservation sums (firex_reduce);
 *   * every `dumpEvery` frames (and the last): the stored bytes of the dumped
 *     fields, one `.fxf` file each (format below);
 *   * with `--frame-every N`, every N frames (and the last): the presented
 *     frame at the output resolution (firex_render_frame: ray march, droplet
 *     streaks, bloom and tone-map composite), as an `.fxf` dump of field
 *     "frame" (kind u8, 4 components, dims width x height x 1) plus a `.ppm`
 *     copy of its rgb for viewing;
 *   * verification: every dump is read back from disk and its checksum
 *     recomputed on the host (firex_checksum_bytes) and compared with its
 *     header and with the device checksum of the same frame.
 *
 * Field dump format (`.fxf`, little-endian): a 128-byte header, then the
 * field's stored bytes verbatim.
 *    0  char[8]  magic "FXFIELD1"
 *    8  u32      header bytes (128)
 *   12  u32      field id (FIREX_FIELD_*)
 *   16  u32      storage kind (FIREX_KIND_*: 0 f32, 1 f16, 2 bf16, 3 u32, 4 u8)
 *   20  u32      components per record
 *   24  i32[3]   grid nx, ny, nz (a frame: width, height, 1)
 *   36  u32      0
 *   40  u64      records (cells, or particles)
 *   48  u64      stored bytes
 *   56  u64      frame (1-based, after warm-up)
 *   64  f64      simulation time (s)
 *   72  u64      checksum of the stored bytes
 *   80  char[32] field name (NUL-padded)
 *  112  16 bytes 0
 *
 * CLI: node fire_amd/js/record_fire_hip.js --scene inferno --solver hd --duration 2 --dump-every 30
 * CommonJS, Node >= 12.
 */
const fs = require('fs');
const path = require('path');
const os = require('os');
const { execFileSync } = require('child_process');
const { FireXHipSolver, native } = require('./firex_hip_solver.js');

const FRAME_RATE = 60;
const SCHEMA_VERSION = 1;
const DUMP_MAGIC = 'FXFIELD1';
const DUMP_HEADER_BYTES = 128;
const FIELD_IDS = {
  flow: 0, species: 1, pressure: 2, correction: 3, divergence: 4, reaction: 5, liquid: 6, flowPredicted: 7,
  packedFlow: 8, packedChemistry: 9, packedMedia: 10, particles: 11, particleDensity: 12, frame: 14,
};
// Output presets and their aliases, scripts/record-fire-lib.mjs:32-34, 70-80.
const OUTPUT_PRESETS = { hd: [1920, 1080], qhd: [2560, 1440], uhd: [3840, 2160] };
const OUTPUT_ALIASES = {
  hd: 'hd', 1080: 'hd', '1080p': 'hd', qhd: 'qhd', 1440: 'qhd', '1440p': 'qhd', uhd: 'uhd', 2160: 'uhd', '2160p': 'uhd',
  '4k': 'uhd',
};
const AUDIT_FIELDS = ['flow', 'species', 'pressure', 'correction', 'divergence', 'reaction', 'liquid', 'particles'];
const DUMP_FIELDS = ['flow', 'species', 'pressure', 'liquid'];
const AUDIT_REDUCTIONS = [
  'sumFuel', 'sumOxygen', 'sumProducts', 'sumVapor', 'sumTemperature', 'sumReaction', 'residualL2', 'flowDivergenceL2',
];
const SOLVER_TIERS = ['hd', 'qhd', 'uhd'];

// Field views (value of the viewMode parameter), src/sim/firex/scenes.ts:8-17.
const FIELD_VIEWS = {
  beauty: 0, temperature: 1, reaction: 2, 'fuel-oxygen': 3, products: 4, 'vapor-soot': 5, 'liquid-velocity': 6,
  divergence: 7,
};

// Scene presets, src/sim/firex/scenes.ts:32-69 (parameter values only).
const SCENES = {
  methane: { burnerSize: 1, burnerDepth: 1, sourceThickness: 1, firePower: 1, sourceLift: 1, fuelRate: 0.48, oxygenRate: 0.9, heatEfficiency: 0.78, vorticity: 0.9, buoyancyScale: 1, flamePersistence: 1, airEntrainment: 0, waterFlow: 0, sprayAngle: 10, aimHeight: 0.42, domainWidth: 1, domainHeight: 1, domainDepth: 1, viewZoom: 1, exposure: 1, frontDefinition: 1.15, nozzleType: 0, waterEnabled: false },
  'large-fire': { burnerSize: 2.25, burnerDepth: 2, sourceThickness: 1.25, firePower: 2.1, sourceLift: 1.8, fuelRate: 0.42, oxygenRate: 1, heatEfficiency: 0.92, vorticity: 1.4, buoyancyScale: 1.4, flamePersistence: 1.45, airEntrainment: 1.6, waterFlow: 0, sprayAngle: 10, aimHeight: 0.42, domainWidth: 1, domainHeight: 1, domainDepth: 1, viewZoom: 1.55, exposure: 0.85, frontDefinition: 1.45, nozzleType: 0, waterEnabled: false },
  inferno: { burnerSize: 3.1, burnerDepth: 2.7, sourceThickness: 1.35, firePower: 2.5, sourceLift: 1.6, fuelRate: 0.3, oxygenRate: 1, heatEfficiency: 0.86, vorticity: 3.6, buoyancyScale: 2.15, flamePersistence: 1.45, airEntrainment: 1.65, waterFlow: 0, sprayAngle: 12, aimHeight: 0.42, domainWidth: 1.15, domainHeight: 1.35, domainDepth: 1.15, viewZoom: 2.1, exposure: 1.25, frontDefinition: 1.65, nozzleType: 0, waterEnabled: false },
  rich: { burnerSize: 1, burnerDepth: 1, sourceThickness: 1, firePower: 1, sourceLift: 1, fuelRate: 0.92, oxygenRate: 0.42, heatEfficiency: 0.62, vorticity: 1.6, buoyancyScale: 1, flamePersistence: 1, airEntrainment: 0, waterFlow: 0, sprayAngle: 12, aimHeight: 0.42, domainWidth: 1, domainHeight: 1, domainDepth: 1, viewZoom: 1, exposure: 1, frontDefinition: 1.35, nozzleType: 0, waterEnabled: false },
  'top-jet': { burnerSize: 1, burnerDepth: 1, sourceThickness: 1, firePower: 1, sourceLift: 1, fuelRate: 0.65, oxygenRate: 0.82, heatEfficiency: 0.75, vorticity: 1.1, buoyancyScale: 1, flamePersistence: 1, airEntrainment: 0, waterFlow: 0.7, sprayAngle: 6, aimHeight: 0.64, domainWidth: 1, domainHeight: 1, domainDepth: 1, viewZoom: 1, exposure: 1, frontDefinition: 1.25, nozzleType: 0, waterEnabled: true },
  'base-spray': { burnerSize: 1, burnerDepth: 1, sourceThickness: 1, firePower: 1, sourceLift: 1, fuelRate: 0.65, oxygenRate: 0.82, heatEfficiency: 0.75, vorticity: 1.2, buoyancyScale: 1, flamePersistence: 1, airEntrainment: 0, waterFlow: 0.74, sprayAngle: 30, aimHeight: 0.14, domainWidth: 1, domainHeight: 1, domainDepth: 1, viewZoom: 1, exposure: 1, frontDefinition: 1.25, nozzleType: 1, waterEnabled: true },
};

function usage() {
  return [
    'Usage: node fire_amd/js/record_fire_hip.js [options]',
    '  --scene NAME            methane | large-fire | inferno | rich | top-jet | base-spray (inferno)',
    '  --field VIEW            field view (viewMode): beauty, temperature, ..., divergence (beauty)',
    '  --solver TIER           hd | qhd | uhd (hd)',
    '  --grid XxYxZ            per-axis minimum grid over the tier preset',
    '  --dims XxYxZ            exact grid, below the tier preset too (small previews and checks)',
    '  --duration SECONDS      captured time, a whole number of 60 fps frames (1)',
    '  --warmup SECONDS        warm-up before the capture (0)',
    '  --dump-every N          dump fields every N frames and at the last frame (0: last frame only)',
    '  --dump-fields LIST      comma list of fields to dump (flow,species,pressure,liquid)',
    '  --frame-every N         write the presented frame every N frames and at the last (0: last only; default none)',
    '  --resolution PRESET     hd | qhd | uhd (1080p, 1440p, 2160p, 4k): size of the written frames (hd)',
    '  --camera-yaw R, --camera-pitch R, --camera-distance D   camera (the engine\'s defaults)',
    '  --optical-detail N      0 | 256 | 512 | 1024: optical detail target of the beauty view (1024)',
    '  --audit-fields LIST     comma list of fields checksummed every frame',
    '  --plan reference|fused  kernel plan (fused)',
    '  --species-storage S     f32 | f16 | bf16 (f32)',
    '  --particles N           SPH droplets (the tier\'s count)',
    '  --memory-budget BYTES   guard on the device bytes of the plume',
    '  --output PATH           capture directory (manifest at PATH.json)',
    '  --output-dir DIR        directory of the default output name (captures)',
    '  --overwrite             replace an existing capture',
  ].join('\n');
}

function frameCount(seconds, flag) {
  if (!Number.isFinite(seconds) || seconds < 0) throw new RangeError(`${flag} must be a non-negative number of seconds.`);
  const frames = Math.round(seconds * FRAME_RATE);
  if (Math.abs(seconds * FRAME_RATE - frames) > 1e-7) throw new RangeError(`${flag} must be a whole number of 60 fps frames.`);
  return frames;
}

function parseList(value, allowed, flag) {
  const items = value.split(',').map((item) => item.trim()).filter(Boolean);
  for (const item of items) if (!(item in allowed)) throw new RangeError(`Unknown field "${item}" for ${flag}.`);
  return [...new Set(items)];
}

function defaultOptions() {
  return {
    scene: 'inferno', field: 'beauty', solverTier: 'hd', grid: null, exactDims: null, durationSeconds: 1, warmupSeconds: 0,
    dumpEvery: 0, dumpFields: [...DUMP_FIELDS], auditFields: [...AUDIT_FIELDS], plan: 'fused',
    speciesStorage: 'f32', particleCount: null, gridMemoryBudgetBytes: 0, output: null, outputDir: 'captures',
    overwrite: false, help: false, frameEvery: null, outputPreset: 'hd', camera: {}, opticalDetailTarget: 1024,
  };
}

function nonNegativeInteger(text, name) {
  const n = Number(text);
  if (!Number.isInteger(n) || n < 0) throw new RangeError(`${name} must be a non-negative integer.`);
  return n;
}

function oneOf(text, allowed, message) {
  if (!allowed.includes(text)) throw new RangeError(message);
  return text;
}

// The options: each entry names its spellings, whether it reads a value, and
// how that value updates the options.
const OPTION_SPECS = [
  [['--help', '-h'], false, (o) => { o.help = true; }],
  [['--scene'], true, (o, text) => {
    o.scene = text.toLowerCase();
    if (!(o.scene in SCENES)) throw new RangeError(`Unknown scene "${o.scene}".`);
  }],
  [['--field'], true, (o, text) => {
    o.field = text.toLowerCase();
    if (!(o.field in FIELD_VIEWS)) throw new RangeError(`Unknown field view "${o.field}".`);
  }],
  [['--solver', '--solver-tier'], true, (o, text) => {
    o.solverTier = oneOf(text.toLowerCase(), SOLVER_TIERS, `Unknown solver tier "${text.toLowerCase()}".`);
  }],
  [['--grid', '--dims'], true, (o, text, name) => {
    const dims = text.toLowerCase().split('x').map(Number);
    if (dims.length !== 3 || dims.some((d) => !Number.isInteger(d) || d < 2)) {
      throw new RangeError(`${name} must be XxYxZ with integers of at least 2.`);
    }
    o[name === '--grid' ? 'grid' : 'exactDims'] = dims;
  }],
  [['--duration'], true, (o, text, name) => { o.durationSeconds = Number(text); frameCount(o.durationSeconds, name); }],
  [['--warmup'], true, (o, text, name) => { o.warmupSeconds = Number(text); frameCount(o.warmupSeconds, name); }],
  [['--dump-every'], true, (o, text, name) => { o.dumpEvery = nonNegativeInteger(text, name); }],
  [['--dump-fields'], true, (o, text, name) => { o.dumpFields = parseList(text, FIELD_IDS, name); }],
  [['--audit-fields'], true, (o, text, name) => { o.auditFields = parseList(text, FIELD_IDS, name); }],
  [['--plan'], true, (o, text) => { o.plan = oneOf(text, ['fused', 'reference'], '--plan must be fused or reference.'); }],
  [['--species-storage'], true, (o, text) => {
    o.speciesStorage = oneOf(text, ['f32', 'f16', 'bf16'], '--species-storage must be f32, f16 or bf16.');
  }],
  [['--particles'], true, (o, text, name) => { o.particleCount = nonNegativeInteger(text, name); }],
  [['--memory-budget'], true, (o, text) => {
    const n = Number(text);
    if (!(n >= 0) || !Number.isFinite(n)) throw new RangeError('--memory-budget must be a byte count.');
    o.gridMemoryBudgetBytes = n;
  }],
  [['--frame-every'], true, (o, text, name) => { o.frameEvery = nonNegativeInteger(text, name); }],
  [['--resolution', '--preset'], true, (o, text, name) => {
    const preset = OUTPUT_ALIASES[text.toLowerCase()];
    if (!preset) throw new Error(`Unknown output resolution for ${name}.`);
    o.outputPreset = preset;
  }],
  [['--camera-yaw', '--camera-pitch', '--camera-distance'], true, (o, text, name) => {
    const v = Number(text);
    if (!Number.isFinite(v)) throw new RangeError(`${name} must be a finite number.`);
    const key = `camera${name.slice('--camera-'.length)[0].toUpperCase()}${name.slice('--camera-'.length + 1)}`;
    o.camera[key] = v;
  }],
  [['--optical-detail'], true, (o, text) => {
    const n = Number(text);
    if (![0, 256, 512, 1024].includes(n)) throw new Error('--optical-detail must be 0, 256, 512, or 1024.');
    o.opticalDetailTarget = n;
  }],
  [['--output'], true, (o, text) => { o.output = text; }],
  [['--output-dir'], true, (o, text) => { o.outputDir = text; }],
  [['--overwrite'], false, (o) => { o.overwrite = true; }],
];
const OPTIONS = new Map();
for (const [names, takesValue, apply] of OPTION_SPECS) for (const name of names) OPTIONS.set(name, { takesValue, apply });

// Tokens are "--name", "--name=value" or a value following a name that takes
// one; a value may not itself look like an option.
function parseArguments(argv) {
  const options = defaultOptions();
  const tokens = [...argv];
  while (tokens.length > 0) {
    const token = tokens.shift();
    const match = /^(--[^=]+)=([^]*)$/.exec(token);
    const name = match ? match[1] : token;
    const spec = OPTIONS.get(name);
    if (!spec) throw new Error(`Unknown option ${name}.\n${usage()}`);
    let text = null;
    if (spec.takesValue) {
      if (match) text = match[2];
      else if (tokens.length > 0 && !tokens[0].startsWith('--')) text = tokens.shift();
      else throw new Error(`Option ${name} needs a value.`);
    } else if (match) {
      throw new Error(`Option ${name} takes no value.`);
    }
    spec.apply(options, text, name);
  }
  return options;
}

function resolveOutputPath(options, cwd = process.cwd()) {
  if (options.output) return path.resolve(cwd, options.output);
  const frames = frameCount(options.durationSeconds, '--duration');
  return path.resolve(cwd, options.outputDir,
    `fire-replica-firex-${options.scene}-${options.field}-${options.solverTier}-${frames}f`);
}

function buildCaptureConfig(options, jobId, dimensions) {
  return {
    schemaVersion: SCHEMA_VERSION,
    jobId,
    method: 'firex-hip',
    scene: options.scene,
    sceneValues: SCENES[options.scene],
    fieldView: options.field,
    solverTier: options.solverTier,
    gridDimensions: dimensions,
    requestedGrid: options.grid,
    exactDims: options.exactDims,
    durationSeconds: options.durationSeconds,
    warmupSeconds: options.warmupSeconds,
    frameRate: FRAME_RATE,
    gridMemoryBudgetBytes: options.gridMemoryBudgetBytes,
    plan: options.plan,
    speciesStorage: options.speciesStorage,
    particleCount: options.particleCount,
    dumpEvery: options.dumpEvery,
    dumpFields: options.dumpFields,
    auditFields: options.auditFields,
    frameEvery: options.frameEvery,
    outputPreset: options.outputPreset,
    resolution: OUTPUT_PRESETS[options.outputPreset].join('x'),
    camera: options.camera,
    opticalDetailTarget: options.opticalDetailTarget,
  };
}

function encodeDumpHeader(h) {
  const b = Buffer.alloc(DUMP_HEADER_BYTES);
  b.write(DUMP_MAGIC, 0, 'ascii');
  b.writeUInt32LE(DUMP_HEADER_BYTES, 8);
  b.writeUInt32LE(h.fieldId, 12);
  b.writeUInt32LE(h.kindCode, 16);
  b.writeUInt32LE(h.components, 20);
  for (let i = 0; i < 3; i += 1) b.writeInt32LE(h.dimensions[i], 24 + 4 * i);
  b.writeBigUInt64LE(BigInt(h.records), 40);
  b.writeBigUInt64LE(BigInt(h.storedBytes), 48);
  b.writeBigUInt64LE(BigInt(h.frame), 56);
  b.writeDoubleLE(h.simulationTime, 64);
  b.writeBigUInt64LE(BigInt(`0x${h.checksum}`), 72);
  b.write(h.field.slice(0, 31), 80, 'ascii');
  return b;
}

function decodeDumpHeader(b) {
  if (b.length < DUMP_HEADER_BYTES || b.toString('ascii', 0, 8) !== DUMP_MAGIC) throw new Error('not a Fire-X field dump');
  const headerBytes = b.readUInt32LE(8);
  const name = b.toString('ascii', 80, 112);
  return {
    headerBytes,
    fieldId: b.readUInt32LE(12),
    kindCode: b.readUInt32LE(16),
    components: b.readUInt32LE(20),
    dimensions: [b.readInt32LE(24), b.readInt32LE(28), b.readInt32LE(32)],
    records: Number(b.readBigUInt64LE(40)),
    storedBytes: Number(b.readBigUInt64LE(48)),
    frame: Number(b.readBigUInt64LE(56)),
    simulationTime: b.readDoubleLE(64),
    checksum: b.readBigUInt64LE(72).toString(16).padStart(16, '0'),
    field: name.slice(0, name.indexOf('\0') < 0 ? name.length : name.indexOf('\0')),
  };
}

// Binary PPM (P6) of an rgba8 frame's rgb, row 0 at the top.
function encodePpm(rgba, width, height) {
  const head = Buffer.from(`P6\n${width} ${height}\n255\n`, 'ascii');
  const body = Buffer.alloc(width * height * 3);
  for (let i = 0, j = 0; i < width * height; i += 1, j += 3) {
    body[j] = rgba[4 * i];
    body[j + 1] = rgba[4 * i + 1];
    body[j + 2] = rgba[4 * i + 2];
  }
  return Buffer.concat([head, body]);
}

function readDump(filePath) {
  const bytes = fs.readFileSync(filePath);
  const header = decodeDumpHeader(bytes);
  const data = new Uint8Array(bytes.buffer, bytes.byteOffset + header.headerBytes, bytes.length - header.headerBytes);
  if (data.length !== header.storedBytes) throw new Error(`${filePath}: ${data.length} data bytes, header says ${header.storedBytes}`);
  return { header, data };
}

function optionalCommand(command, args, cwd) {
  try {
    return execFileSync(command, args, { cwd, encoding: 'utf8', stdio: ['ignore', 'pipe', 'ignore'] }).trim();
  } catch (error) {
    return '';
  }
}

function removePath(target) {
  if (fs.existsSync(target)) fs.rmdirSync(target, { recursive: true });
}

/** Runs one capture job; resolves to {outputPath, manifestPath, manifest}. */
async function runCapture(options, log = (line) => process.stdout.write(`${line}\n`)) {
  const outputPath = resolveOutputPath(options);
  const manifestPath = `${outputPath}.json`;
  const partialPath = `${outputPath}.partial`;
  const partialManifestPath = `${manifestPath}.partial`;
  for (const candidate of [outputPath, manifestPath]) {
    if (fs.existsSync(candidate) && !options.overwrite) {
      throw new Error(`${candidate} already exists. Use --overwrite to replace it.`);
    }
  }
  for (const candidate of [partialPath, partialManifestPath]) {
    if (!fs.existsSync(candidate)) continue;
    if (!options.overwrite) throw new Error(`${candidate} already exists. Use --overwrite to replace it.`);
    if (candidate === partialPath) removePath(candidate); else fs.unlinkSync(candidate);
  }

  const dimensions = options.exactDims || native.resolveOfflineDims(options.solverTier, options.grid);
  const expectedFrames = frameCount(options.durationSeconds, '--duration');
  const warmupFrames = frameCount(options.warmupSeconds, '--warmup');
  const jobId = `${options.scene}-${options.field}-${Date.now().toString(36)}`;
  const config = buildCaptureConfig(options, jobId, dimensions);
  const startedAt = new Date();
  const solverOptions = {
    tier: options.solverTier, dims: dimensions, plan: options.plan, speciesStorage: options.speciesStorage,
    memoryBudgetBytes: options.gridMemoryBudgetBytes, jacobiAutotune: options.plan === 'fused' ? 1 : 0,
  };
  if (options.particleCount !== null) solverOptions.particleCount = options.particleCount;
  const solver = new FireXHipSolver(solverOptions);
  for (const [name, value] of Object.entries(SCENES[options.scene])) solver.setParameter(name, value);
  solver.setParameter('viewMode', FIELD_VIEWS[options.field]);
  solver.setParameter('opticalDetailTarget', options.opticalDetailTarget);
  for (const [name, value] of Object.entries(options.camera)) solver.setParameter(name, value);
  const [frameWidth, frameHeight] = OUTPUT_PRESETS[options.outputPreset];

  log(`Capturing ${options.scene} / ${options.field} -> ${outputPath}`);
  log(`  Solver ${options.solverTier}, exact ${dimensions.join('x')} grid, ${expectedFrames} frames at ${FRAME_RATE} fps`);
  fs.mkdirSync(partialPath, { recursive: true });
  const dumps = [];
  const frames = [];
  try {
    await solver.initialize();
    const info = solver.getGridInfo();
    if (info.dimensions.join('x') !== dimensions.join('x')) {
      throw new Error(`Fire-X allocated ${info.dimensions.join('x')}, expected exact ${dimensions.join('x')}.`);
    }
    for (let index = 0; index < warmupFrames; index += 1) await solver.renderOfflineFrame(1 / FRAME_RATE, false);
    const captureStart = process.hrtime.bigint();
    for (let frame = 1; frame <= expectedFrames; frame += 1) {
      await solver.renderOfflineFrame(1 / FRAME_RATE, false);
      const grid = solver.getGridInfo();
      const record = { frame, simulationTime: grid.simulationTime, steps: grid.frameIndex, checksums: {}, reductions: {} };
      for (const name of options.auditFields) record.checksums[name] = solver.fieldChecksum(name);
      for (const name of AUDIT_REDUCTIONS) record.reductions[name] = solver.reduce(name);
      frames.push(record);
      const frameNow = options.frameEvery !== null &&
        (frame === expectedFrames || (options.frameEvery > 0 && frame % options.frameEvery === 0));
      if (frameNow) {
        const pixels = solver.renderFrame(frameWidth, frameHeight);
        const layout = solver.fieldLayout('frame');
        const checksum = solver.fieldChecksum('frame');
        record.checksums.frame = checksum;
        const stem = `frame-${String(frame).padStart(6, '0')}`;
        const header = encodeDumpHeader({
          fieldId: FIELD_IDS.frame, kindCode: layout.kindCode, components: layout.components,
          dimensions: [frameWidth, frameHeight, 1], records: layout.records, storedBytes: layout.storedBytes, frame,
          simulationTime: grid.simulationTime, checksum, field: 'frame',
        });
        fs.writeFileSync(path.join(partialPath, `${stem}-frame.fxf`),
          Buffer.concat([header, Buffer.from(pixels.buffer, pixels.byteOffset, pixels.byteLength)]));
        fs.writeFileSync(path.join(partialPath, `${stem}.ppm`), encodePpm(pixels, frameWidth, frameHeight));
        dumps.push({ frame, field: 'frame', file: `${stem}-frame.fxf`, kind: layout.kind, components: layout.components,
          records: layout.records, storedBytes: layout.storedBytes, checksum, image: `${stem}.ppm` });
      }
      const dumpNow = frame === expectedFrames || (options.dumpEvery > 0 && frame % options.dumpEvery === 0);
      if (!dumpNow) continue;
      for (const name of options.dumpFields) {
        const layout = solver.fieldLayout(name);
        const data = solver.readFieldRaw(name);
        const checksum = record.checksums[name] || solver.fieldChecksum(name);
        const file = `frame-${String(frame).padStart(6, '0')}-${name}.fxf`;
        const header = encodeDumpHeader({
          fieldId: FIELD_IDS[name], kindCode: layout.kindCode, components: layout.components, dimensions,
          records: layout.records, storedBytes: layout.storedBytes, frame, simulationTime: grid.simulationTime,
          checksum, field: name,
        });
        fs.writeFileSync(path.join(partialPath, file), Buffer.concat([header, Buffer.from(data.buffer, data.byteOffset, data.byteLength)]));
        dumps.push({ frame, field: name, file, kind: layout.kind, components: layout.components,
          records: layout.records, storedBytes: layout.storedBytes, checksum });
      }
    }
    const captureSeconds = Number(process.hrtime.bigint() - captureStart) / 1e9;
    const finalInfo = solver.getGridInfo();
    solver.dispose();

    log('  Capture finished. Verifying dumps from disk...');
    let verifiedBytes = 0;
    const mismatches = [];
    for (const d of dumps) {
      const { header, data } = readDump(path.join(partialPath, d.file));
      const onDisk = native.checksumBytes(data);
      const frameRecord = frames[d.frame - 1];
      const device = frameRecord.checksums[d.field] || d.checksum;
      if (onDisk !== header.checksum || onDisk !== device || header.frame !== d.frame || header.field !== d.field) {
        mismatches.push({ file: d.file, onDisk, header: header.checksum, device });
      }
      verifiedBytes += data.length;
    }
    if (mismatches.length) throw new Error(`dump verification failed: ${JSON.stringify(mismatches.slice(0, 4))}`);
    const finishedAt = new Date();
    const manifest = {
      schemaVersion: SCHEMA_VERSION,
      status: 'verified',
      outputPath,
      startedAt: startedAt.toISOString(),
      finishedAt: finishedAt.toISOString(),
      wallSeconds: (finishedAt.getTime() - startedAt.getTime()) / 1000,
      repositoryCommit: optionalCommand('git', ['rev-parse', 'HEAD'], path.join(__dirname, '..', '..')) || null,
      device: { library: 'libfirex_hip.so', abiVersion: native.abiVersion(), host: os.hostname(), deviceBytes: finalInfo.deviceBytes },
      config,
      expectedFrames,
      captureSeconds,
      checksum: 'sum over little-endian u32 words w_i of splitmix64((i << 32) | w_i) mod 2^64 (firex_field_checksum)',
      frames,
      dumps,
      verification: {
        frames: frames.length, dumps: dumps.length, verifiedDumps: dumps.length - mismatches.length,
        verifiedBytes, checksumsMatch: mismatches.length === 0,
      },
    };
    fs.writeFileSync(partialManifestPath, `${JSON.stringify(manifest, null, 2)}\n`, 'utf8');
    // install the verified pair: the capture directory, then its manifest
    if (fs.existsSync(outputPath)) removePath(outputPath);
    fs.renameSync(partialPath, outputPath);
    fs.renameSync(partialManifestPath, manifestPath);
    log(`  Verified ${dumps.length} dumps (${verifiedBytes} bytes), ${frames.length} audited frames.`);
    log(`  Saved ${outputPath}\n  Manifest ${manifestPath}`);
    return { outputPath, manifestPath, manifest };
  } catch (error) {
    solver.dispose();
    removePath(partialPath);
    if (fs.existsSync(partialManifestPath)) fs.unlinkSync(partialManifestPath);
    throw error;
  }
}

async function main(argv) {
  const options = parseArguments(argv);
  if (options.help) {
    process.stdout.write(`${usage()}\n`);
    return 0;
  }
  await runCapture(options);
  return 0;
}

if (require.main === module) {
  main(process.argv.slice(2)).then((code) => { process.exitCode = code; }).catch((error) => {
    process.stderr.write(`${error && error.stack ? error.stack : error}\n`);
    process.exitCode = 1;
  });
}

module.exports = {
  parseArguments, resolveOutputPath, buildCaptureConfig, encodeDumpHeader, decodeDumpHeader, readDump, runCapture,
  encodePpm, SCENES, FIELD_VIEWS, FIELD_IDS, DUMP_HEADER_BYTES, OUTPUT_PRESETS,
};
