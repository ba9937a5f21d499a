'use strict';
// Drives the JS drop-in (fire_amd/js/firex_hip_solver.js) through the
// configuration calls of the reference's capture page, unchanged
// (src/cli/captureApp.ts:124-139), and through transactional re-grids
// (FireXEngine.ts:895-927, 1891-1915), then dumps fields and a JSON record
// for the Python side to compare with the oracle.
//
//   node capture_sequence.js OUT_DIR CONFIG_JSON
//
// CONFIG_JSON.mode:
//   "capture" : the captureApp sequence, then config.frames offline frames
//   "regrid"  : a live plume, a re-grid refused by the memory guard (the plume
//               must stay live and bit-identical), a re-grid that succeeds
//   "offline" : setter rules before initialize (no device work)
const fs = require('fs');
const path = require('path');
const fx = require(path.join(__dirname, '..', '..', 'fire_amd', 'js', 'firex_hip_solver.js'));

const { FireXHipSolver, resolveFireXOfflineDimensions } = fx;
const AUDITED = ['flow', 'species', 'pressure', 'correction', 'divergence', 'reaction'];

function sameDimensions(a, b) {
  return a.length === b.length && a.every((value, index) => value === b[index]);
}

function dump(outDir, engine, prefix) {
  for (const name of ['flow', 'species', 'reaction', 'divergence', 'pressure']) {
    const data = engine.readField(name);
    fs.writeFileSync(path.join(outDir, `${prefix}${name}.f32`), Buffer.from(data.buffer, data.byteOffset, data.byteLength));
  }
}

function checksums(engine) {
  const out = {};
  for (const name of AUDITED) out[name] = engine.fieldChecksum(name);
  return out;
}

async function captureMode(outDir, config, record) {
  const scene = { values: config.sceneValues };
  const field = { value: config.fieldView };
  const engine = new FireXHipSolver();
  // src/cli/captureApp.ts:124-139, the same calls in the same order
  engine.setGridMemoryBudget(config.gridMemoryBudgetBytes);
  await engine.setQuality('cinematic');
  await engine.setGridDimensions(config.gridDimensions);
  await engine.setOfflineRenderTier(config.solverTier);
  engine.setOfflineOpticalDetailTarget(config.opticalDetailTarget);
  for (const [name, value] of Object.entries(scene.values)) engine.setParameter(name, value);
  engine.setParameter('viewMode', field.value);
  engine.setParameter('cameraYaw', config.camera.yaw);
  engine.setParameter('cameraPitch', config.camera.pitch);
  engine.setParameter('cameraDistance', config.camera.distance);
  await engine.initialize();
  const expectedDimensions = resolveFireXOfflineDimensions(config.solverTier, config.gridDimensions);
  const gridInfo = engine.getGridInfo();
  if (!sameDimensions(gridInfo.dimensions, expectedDimensions)) {
    throw new Error(`Fire-X allocated ${gridInfo.dimensions.join('×')}, expected exact ${expectedDimensions.join('×')}.`);
  }
  for (let index = 0; index < config.frames; index += 1) await engine.renderOfflineFrame(1 / 60, false);
  dump(outDir, engine, '');
  record.dimensions = gridInfo.dimensions;
  record.expectedDimensions = expectedDimensions;
  record.tier = fx.native.configFromTier(config.solverTier);
  record.opticalDetailTarget = engine.getParameter('opticalDetailTarget');
  record.frameIndex = engine.getGridInfo().frameIndex;
  engine.dispose();
}

async function regridMode(outDir, config, record) {
  const engine = new FireXHipSolver();
  engine.setGridMemoryBudget(config.initialBudgetBytes);
  await engine.setQuality(config.quality);
  for (const [name, value] of Object.entries(config.sceneValues)) engine.setParameter(name, value);
  await engine.initialize();
  record.tier = fx.native.configFromTier(config.quality);
  record.initialDimensions = engine.getGridInfo().dimensions;
  for (let index = 0; index < config.framesBefore; index += 1) await engine.renderOfflineFrame(1 / 60, false);
  const before = { checksums: checksums(engine), info: engine.getGridInfo() };
  // a re-grid above the guard: refused, nothing changes
  try {
    await engine.setGridDimensions(config.refusedDimensions);
    record.refusedError = null;
  } catch (error) {
    record.refusedError = error.message;
  }
  const after = { checksums: checksums(engine), info: engine.getGridInfo() };
  record.unchangedChecksums = JSON.stringify(before.checksums) === JSON.stringify(after.checksums);
  record.unchangedInfo = sameDimensions(before.info.dimensions, after.info.dimensions)
    && before.info.frameIndex === after.info.frameIndex && before.info.simulationTime === after.info.simulationTime;
  // the plume stays live: more frames continue its trajectory
  for (let index = 0; index < config.framesAfterRefusal; index += 1) await engine.renderOfflineFrame(1 / 60, false);
  dump(outDir, engine, 'kept_');
  record.keptFrameIndex = engine.getGridInfo().frameIndex;
  // a re-grid inside a raised guard: a fresh plume on the new grid
  engine.setGridMemoryBudget(config.raisedBudgetBytes);
  await engine.setGridDimensions(config.acceptedDimensions);
  const rebuilt = engine.getGridInfo();
  record.rebuiltDimensions = rebuilt.dimensions;
  record.rebuiltFrameIndex = rebuilt.frameIndex;
  for (let index = 0; index < config.framesAfterRegrid; index += 1) await engine.renderOfflineFrame(1 / 60, false);
  dump(outDir, engine, 'regrid_');
  // the same grid again is a no-op (FireXEngine.ts:904-908)
  const frameBefore = engine.getGridInfo().frameIndex;
  await engine.setGridDimensions(config.acceptedDimensions);
  record.sameGridKeepsState = engine.getGridInfo().frameIndex === frameBefore;
  engine.dispose();
}

async function offlineMode(outDir, config, record) {
  const engine = new FireXHipSolver();
  engine.setGridMemoryBudget(1);
  record.clampedLow = engine.gridMemoryBudgetBytes;
  engine.setGridMemoryBudget(1e15);
  record.clampedHigh = engine.gridMemoryBudgetBytes;
  const unchanged = engine.gridMemoryBudgetBytes;
  engine.setGridMemoryBudget(Number.NaN);
  engine.setGridMemoryBudget(-5);
  record.ignoredInvalid = engine.gridMemoryBudgetBytes === unchanged;
  await engine.setQuality('high');
  record.highDimensions = engine.getGridInfo().dimensions;
  await engine.setGridDimensions([48, 64, 40]);
  record.overrideDimensions = engine.getGridInfo().dimensions;
  await engine.setOfflineRenderTier('qhd');
  record.offlineDimensions = engine.getGridInfo().dimensions;
  try {
    await engine.setQuality('low');
    record.qualityError = null;
  } catch (error) {
    record.qualityError = error.message;
  }
  try {
    await engine.setGridDimensions([64, 64, 64]);
    record.gridError = null;
  } catch (error) {
    record.gridError = error.message;
  }
  engine.setOfflineOpticalDetailTarget(700);
  record.snappedTarget = engine.offlineOpticalDetailTarget;
  await engine.setOfflineRenderTier(null);
  record.liveAgainDimensions = engine.getGridInfo().dimensions;
  engine.dispose();
}

async function main() {
  const [outDir, configJson] = process.argv.slice(2);
  const config = JSON.parse(configJson);
  const record = { mode: config.mode };
  if (config.mode === 'capture') await captureMode(outDir, config, record);
  else if (config.mode === 'regrid') await regridMode(outDir, config, record);
  else await offlineMode(outDir, config, record);
  fs.writeFileSync(path.join(outDir, 'record.json'), JSON.stringify(record));
}

main().catch((e) => { console.error(e && e.stack ? e.stack : e); process.exit(1); });
