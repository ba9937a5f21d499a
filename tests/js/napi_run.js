'use strict';
// Drives the JS drop-in (fire_amd/js/firex_hip_solver.js) the way the
// reference's offline capture drives FireXEngine (src/cli/captureApp.ts:124-155):
// construct, apply a scene, initialize, renderOfflineFrame(dt) per frame, then
// dumps fields for the Python side to compare with the oracle.
const fs = require('fs');
const path = require('path');
const { FireXHipSolver } = require(path.join(__dirname, '..', '..', 'fire_amd', 'js', 'firex_hip_solver.js'));

async function main() {
  const [outDir, configJson] = process.argv.slice(2);
  const cfg = JSON.parse(configJson);
  const solver = new FireXHipSolver(cfg.options);
  for (const [k, v] of Object.entries(cfg.scene || {})) solver.setParameter(k, v);
  await solver.initialize();
  for (let i = 0; i < cfg.frames; i += 1) await solver.renderOfflineFrame(cfg.delta, false);
  for (const name of cfg.fields) {
    const data = solver.readField(name);
    fs.writeFileSync(path.join(outDir, `${name}.f32`), Buffer.from(data.buffer, data.byteOffset, data.byteLength));
  }
  const controller = { aborted: true };
  let abortName = null;
  try { await solver.renderOfflineFrame(cfg.delta, false, controller); } catch (e) { abortName = e.name; }
  fs.writeFileSync(path.join(outDir, 'info.json'), JSON.stringify({ info: solver.getGridInfo(), abortName }));
  solver.dispose();
}

main().catch((e) => { console.error(e && e.stack ? e.stack : e); process.exit(1); });
