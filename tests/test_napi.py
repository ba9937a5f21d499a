"""The Node-API drop-in: the addon loads in Node, its host-side functions give
the reference's known answers, and (on a GPU) the JS FireXHipSolver class
driven like the reference's offline capture produces fields bit-identical to
the FP32 oracle."""
import json
import re
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


def _run_sequence(tmp_path, config, timeout=600):
    res = subprocess.run([NODE, os.path.join(ROOT, "tests", "js", "capture_sequence.js"), str(tmp_path),
                          json.dumps(config)], capture_output=True, text=True, timeout=timeout, cwd=ROOT)
    assert res.returncode == 0, res.stderr
    return json.loads((tmp_path / "record.json").read_text())


def _oracle_for(dims, tier, scene, steps):
    o = orc.Oracle(dims, precision=32, pressure_iterations=int(tier["pressureIterations"]),
                   correction_iterations=int(tier["correctionIterations"]), time_step=tier["timeStep"])
    o.set_bool("cinematic", bool(tier["cinematic"]))
    for k, v in scene.items():
        o.set_parameter(k, v)
    o.reset()
    o.step(1 / 60, steps)
    return o


def _assert_dump_equals(tmp_path, prefix, o):
    for name, fid in (("flow", orc.FIELD_FLOW), ("species", orc.FIELD_SPECIES), ("reaction", orc.FIELD_REACTION),
                      ("divergence", orc.FIELD_DIVERGENCE), ("pressure", orc.FIELD_PRESSURE)):
        got = np.fromfile(tmp_path / f"{prefix}{name}.f32", dtype=np.float32)
        want = o.read(fid).astype(np.float32).ravel()
        np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32), err_msg=prefix + name)


@needs_node
def test_setters_follow_the_reference_before_initialize(tmp_path):
    """setGridMemoryBudget clamps to [256 MiB, 32 GiB] and ignores invalid
    values (FireXEngine.ts:929-935); a live quality's grid, an override, and
    an offline tier's resolved grid (per-axis maximum of preset and
    override, :279-290); setQuality and setGridDimensions refuse while an
    offline tier is active (:865-867, :900-902); the optical-detail target
    snaps to 0 / 256 / 512 / 1024 (:834-841)."""
    r = _run_sequence(tmp_path, {"mode": "offline"})
    assert r["clampedLow"] == 256 * 2 ** 20 and r["clampedHigh"] == 32 * 2 ** 30 and r["ignoredInvalid"]
    assert r["highDimensions"] == [64, 96, 64] and r["overrideDimensions"] == [48, 64, 40]
    assert r["offlineDimensions"] == [160, 240, 160]
    assert r["qualityError"] == "Compute quality cannot change while an exact offline solver bundle is active."
    assert r["gridError"] == "Grid dimensions cannot change while an exact offline solver bundle is active."
    assert r["snappedTarget"] == 512 and r["liveAgainDimensions"] == [48, 64, 40]


SCENE = {"fuelRate": 0.42, "oxygenRate": 0.95, "vorticity": 2.4, "firePower": 2.0, "buoyancyScale": 1.8}


@pytest.mark.gpu
@needs_node
def test_capture_call_sequence_gets_the_offline_grid(tmp_path):
    """The capture page's calls, unchanged (captureApp.ts:124-139): budget,
    setQuality('cinematic'), setGridDimensions, setOfflineRenderTier('hd'),
    the optical-detail target, the scene, then initialize.  The solver
    allocates resolveFireXOfflineDimensions('hd', [136, 160, 96]) =
    136 x 192 x 128 and two offline frames equal the FP32 oracle with the hd
    tier's settings bit for bit."""
    if device_count() < 1:
        pytest.fail("no HIP device")
    cfg = {"mode": "capture", "gridMemoryBudgetBytes": 4 * 2 ** 30, "gridDimensions": [136, 160, 96],
           "solverTier": "hd", "opticalDetailTarget": 600, "sceneValues": SCENE, "fieldView": 0,
           "camera": {"yaw": -0.3, "pitch": 0.05, "distance": 4.6}, "frames": 2}
    print("capture sequence: hd tier, grid override [136, 160, 96]", flush=True)
    r = _run_sequence(tmp_path, cfg)
    assert r["dimensions"] == r["expectedDimensions"] == [136, 192, 128]
    assert r["opticalDetailTarget"] == 512 and r["frameIndex"] == 2
    _assert_dump_equals(tmp_path, "", _oracle_for((136, 192, 128), r["tier"], SCENE, 2))


@pytest.mark.gpu
@needs_node
def test_regrid_refused_by_the_guard_keeps_the_plume(tmp_path):
    """A re-grid whose current + replacement estimate exceeds the guard is
    refused with the reference's message and leaves the plume live and
    bit-identical (FireXEngine.ts:1891-1915, 2823-2853): its fields keep their
    checksums and two more frames continue the trajectory the FP32 oracle
    gives for five.  A re-grid inside a raised guard installs a fresh plume on
    the new grid (time and fields reset, parameters kept) that again equals
    the oracle; asking for the same grid again changes nothing."""
    if device_count() < 1:
        pytest.fail("no HIP device")
    cfg = {"mode": "regrid", "quality": "high", "initialBudgetBytes": 256 * 2 ** 20, "sceneValues": SCENE,
           "framesBefore": 3, "refusedDimensions": [160, 240, 160], "framesAfterRefusal": 2,
           "raisedBudgetBytes": 2 ** 30, "acceptedDimensions": [48, 72, 48], "framesAfterRegrid": 2}
    print("regrid: high quality 64x96x64, refused 160x240x160, accepted 48x72x48", flush=True)
    r = _run_sequence(tmp_path, cfg)
    assert r["initialDimensions"] == [64, 96, 64]
    assert r["refusedError"] is not None
    assert re.fullmatch(r"Transactional rebuild peak requires about [0-9.]+ MiB \([0-9.]+ MiB current \+ "
                        r"[0-9.]+ MiB replacement\), above the selected 256 MiB guard\.", r["refusedError"])
    assert r["unchangedChecksums"] and r["unchangedInfo"] and r["keptFrameIndex"] == 5
    _assert_dump_equals(tmp_path, "kept_", _oracle_for((64, 96, 64), r["tier"], SCENE, 5))
    assert r["rebuiltDimensions"] == [48, 72, 48] and r["rebuiltFrameIndex"] == 0
    _assert_dump_equals(tmp_path, "regrid_", _oracle_for((48, 72, 48), r["tier"], SCENE, 2))
    assert r["sameGridKeepsState"]
