"""The Node-API drop-in: the addon loads in Node, its host-side functions give
the reference's known answers, and (on a GPU) the JS FireXHipSolver class
driven like the reference's offline capture produces fields bit-identical to
the FP32 oracle."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from fire_amd import device_count
from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ADDON = os.path.join(ROOT, "fire_amd", "firex_napi.node")
NODE = shutil.which("node")

needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the N-API addon is absent")


def node_eval(script: str) -> str:
    res = subprocess.run([NODE, "-e", script], capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert res.returncode == 0, res.stderr
    return res.stdout


@needs_node
def test_addon_host_functions_match_reference_answers():
    out = json.loads(node_eval(
        "const fx = require('./fire_amd/js/firex_hip_solver.js');"
        "console.log(JSON.stringify({m: fx.estimateFireXGridMemory([80,120,80],1536),"
        "s: fx.splitOfflineSimulationDelta(1/15, 1/30), z: fx.splitOfflineSimulationDelta(0, 1/30),"
        "v: fx.native.abiVersion()}))"))
    assert out["m"]["totalBytes"] == 768_000 * 116 + 1536 * 80 + 64 ** 3 * 4  # gridConfiguration.test.ts:129-139
    assert len(out["s"]) == 2 and max(out["s"]) <= 1 / 30  # offlineTimeline.test.ts:17-22
    assert out["z"] == []
    assert out["v"] == 1


@needs_node
def test_addon_reports_errors_as_js_exceptions():
    script = ("const fx = require('./fire_amd/js/firex_hip_solver.js');"
              "try { fx.native.create({tier:'hd', dims:[16,24,16], memoryBudgetBytes: 1024}); }"
              "catch (e) { console.log(e.code + '|' + e.message); }"
              "try { fx.splitOfflineSimulationDelta(-1, 1/30); } catch (e) { console.log(e.constructor.name); }")
    lines = node_eval(script).strip().splitlines()
    assert lines[0].startswith("FIREX_ERR_3|") and "guard" in lines[0]
    assert lines[1] == "RangeError"


@pytest.mark.gpu
@needs_node
def test_js_drop_in_matches_oracle_bitwise(tmp_path):
    if device_count() < 1:
        pytest.fail("no HIP device")
    dims = (24, 32, 24)
    scene = {"fuelRate": 0.48, "oxygenRate": 0.9, "vorticity": 0.9, "airEntrainment": 0.0, "firePower": 1.0}
    cfg = {"options": {"tier": "hd", "dims": list(dims), "pressureIterations": 6, "correctionIterations": 6,
                       "plan": "fused", "heavyMode": "off"},
           "scene": scene, "frames": 3, "delta": 1 / 60, "fields": ["flow", "species", "reaction", "divergence"]}
    res = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "napi_run.js"), str(tmp_path), json.dumps(cfg)],
                         capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert res.returncode == 0, res.stderr
    o = orc.Oracle(dims, precision=32, pressure_iterations=6, correction_iterations=6)
    o.set_bool("cinematic", True)
    o.set_heavy_mode(orc.HEAVY_OFF)
    for k, v in scene.items():
        o.set_parameter(k, v)
    o.reset()
    o.step(1 / 60, 3)
    for name, fid in (("flow", orc.FIELD_FLOW), ("species", orc.FIELD_SPECIES), ("reaction", orc.FIELD_REACTION),
                      ("divergence", orc.FIELD_DIVERGENCE)):
        got = np.fromfile(tmp_path / f"{name}.f32", dtype=np.float32)
        want = o.read(fid).astype(np.float32).ravel()
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32), err_msg=name)
    info = json.loads((tmp_path / "info.json").read_text())
    assert info["abortName"] == "AbortError"
    assert info["info"]["dimensions"] == list(dims)
    assert info["info"]["frameIndex"] == 3
