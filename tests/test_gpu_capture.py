"""GPU: the device field checksum against the host checksum of the raw
stored bytes for every storage format, and the Node capture driver
(fire_amd/js/record_fire_hip.js) end to end: its per-frame device checksums
equal the checksums of the FP32 oracle's fields at the same frames, its dumps
hold the oracle's bits, and the manifest re-verifies from disk."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from fire_amd import FireXHipSolver, checksum_bytes, device_count
from fire_amd.capture import read_field_dump, verify_capture
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
DT = 1.0 / 60.0

# base-spray preset, src/sim/firex/scenes.ts:63-67: water on, spray nozzle
BASE_SPRAY = {"burnerSize": 1, "burnerDepth": 1, "sourceThickness": 1, "firePower": 1, "sourceLift": 1,
              "fuelRate": 0.65, "oxygenRate": 0.82, "heatEfficiency": 0.75, "vorticity": 1.2, "buoyancyScale": 1,
              "flamePersistence": 1, "airEntrainment": 0, "waterFlow": 0.74, "sprayAngle": 30, "aimHeight": 0.14,
              "domainWidth": 1, "domainHeight": 1, "domainDepth": 1, "nozzleType": 1}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if device_count() < 1:
        pytest.fail("no HIP device visible to the capture tests")


def _oracle_bytes(o, fid, kind="f32"):
    v = o.read(fid)
    return (v.astype(np.uint32) if kind == "u32" else v.astype(np.float32)).tobytes()


@pytest.mark.parametrize("storage,heavy", [("f32", "off"), ("f16", "off"), ("bf16", "off"), ("f32", "f16")])
def test_device_checksum_equals_host_checksum_of_stored_bytes(storage, heavy):
    print(f"checksum storage {storage} heavy {heavy}", flush=True)
    g = FireXHipSolver((33, 40, 21), tier="hd", pressure_iterations=6, correction_iterations=6, plan="fused",
                       species_storage=storage, heavy_mode=heavy, particle_count=300)
    for k, v in BASE_SPRAY.items():
        g.set_parameter(k, v)
    g.set_parameter("waterEnabled", True)
    g.step(DT, 4)
    g.pack_fields()
    names = ["flow", "species", "pressure", "correction", "divergence", "reaction", "liquid", "particles",
             "particle_density", "packed_flow", "packed_chemistry", "packed_media"]
    for name in names:
        raw = g.read_field_raw(name)
        layout = g.field_layout(name)
        assert len(raw) == layout["stored_bytes"]
        assert g.field_checksum(name) == checksum_bytes(raw), name
        # the raw bytes are the stored values the converting read returns
        as_read = g.read_field(name, dtype=np.uint32 if layout["kind"] == "u32" else np.float32).ravel()
        if layout["kind"] == "f16":
            stored = np.frombuffer(raw, dtype=np.float16).astype(np.float32)
        elif layout["kind"] == "bf16":
            stored = (np.frombuffer(raw, dtype=np.uint16).astype(np.uint32) << 16).view(np.float32)
        else:
            stored = np.frombuffer(raw, dtype=np.uint32 if layout["kind"] == "u32" else np.float32)
        np.testing.assert_array_equal(stored.view(np.uint32) if stored.dtype == np.float32 else stored,
                                      as_read.view(np.uint32) if as_read.dtype == np.float32 else as_read)
    g.close()


@pytest.mark.skipif(NODE is None, reason="node is absent")
def test_capture_driver_matches_oracle(tmp_path):
    dims = (32, 48, 32)
    out = tmp_path / "cap"
    cmd = [NODE, os.path.join(ROOT, "fire_amd", "js", "record_fire_hip.js"), "--scene", "base-spray", "--solver", "hd",
           "--dims", "x".join(map(str, dims)), "--duration", "0.2", "--dump-every", "6", "--particles", "640",
           "--dump-fields", "flow,species,liquid,particles", "--output", str(out)]
    print("capture:", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=ROOT)
    assert res.returncode == 0, res.stdout + res.stderr
    manifest = json.loads((tmp_path / "cap.json").read_text())
    assert manifest["status"] == "verified" and manifest["expectedFrames"] == 12
    assert manifest["verification"]["checksumsMatch"] and manifest["verification"]["dumps"] == 8
    assert manifest["config"]["gridDimensions"] == list(dims)
    assert verify_capture(str(tmp_path / "cap.json"))["dumps"] == 8
    assert not os.path.exists(str(out) + ".partial")

    # the same run on the FP32 oracle: hd tier (64 + 64 sweeps, cinematic), the
    # particle state written at creation and again at reset with the scene
    o = orc.Oracle(dims, precision=32, pressure_iterations=64, correction_iterations=64)
    o.set_bool("cinematic", True)
    o.set_heavy_mode(orc.HEAVY_OFF)
    o.set_particle_count(640)
    for k, v in BASE_SPRAY.items():
        o.set_parameter(k, v)
    o.set_bool("waterEnabled", True)
    o.reset()
    fields = [("flow", orc.FIELD_FLOW, "f32"), ("species", orc.FIELD_SPECIES, "f32"),
              ("pressure", orc.FIELD_PRESSURE, "f32"), ("correction", orc.FIELD_CORRECTION, "f32"),
              ("divergence", orc.FIELD_DIVERGENCE, "f32"), ("reaction", orc.FIELD_REACTION, "f32"),
              ("liquid", orc.FIELD_LIQUID, "u32"), ("particles", orc.FIELD_PARTICLES, "f32")]
    dumps = {(d["frame"], d["field"]): d for d in manifest["dumps"]}
    for rec in manifest["frames"]:
        o.step(DT, 1)
        for name, fid, kind in fields:
            want = checksum_bytes(_oracle_bytes(o, fid, kind))
            assert int(rec["checksums"][name], 16) == want, f"frame {rec['frame']}: {name}"
        for name, fid, kind in fields:
            d = dumps.get((rec["frame"], name))
            if d is None:
                continue
            header, arr = read_field_dump(str(out / d["file"]))
            assert header["data"] == _oracle_bytes(o, fid, kind), f"dump {d['file']}"
    assert o.read(orc.FIELD_LIQUID).sum() > 0
    o.close()


@pytest.mark.skipif(NODE is None, reason="node is absent")
def test_capture_frames_match_oracle(tmp_path):
    """--frame-every: the presented hd frames (ray march, droplet streaks of
    the spray, bloom and composite) written by the driver hold the FP32
    oracle's frame bytes at the same frames; the .ppm holds their rgb."""
    dims = (32, 48, 32)
    out = tmp_path / "cap"
    cmd = [NODE, os.path.join(ROOT, "fire_amd", "js", "record_fire_hip.js"), "--scene", "base-spray", "--solver", "hd",
           "--dims", "x".join(map(str, dims)), "--duration", "0.1", "--frame-every", "3", "--particles", "640",
           "--dump-fields", "liquid", "--camera-yaw", "0.4", "--output", str(out)]
    print("capture:", " ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True, timeout=240, cwd=ROOT)
    assert res.returncode == 0, res.stdout + res.stderr
    manifest = json.loads((tmp_path / "cap.json").read_text())
    assert manifest["config"]["resolution"] == "1920x1080" and manifest["config"]["frameEvery"] == 3
    frame_dumps = {d["frame"]: d for d in manifest["dumps"] if d["field"] == "frame"}
    assert sorted(frame_dumps) == [3, 6]
    assert verify_capture(str(tmp_path / "cap.json"))["dumps"] == len(manifest["dumps"])

    o = orc.Oracle(dims, precision=32, pressure_iterations=64, correction_iterations=64)
    o.set_bool("cinematic", True)
    o.set_heavy_mode(orc.HEAVY_OFF)
    o.set_particle_count(640)
    for k, v in BASE_SPRAY.items():
        o.set_parameter(k, v)
    # the preset's view values (scenes.ts:63-67), which the step above does not read
    for k, v in (("viewZoom", 1.0), ("exposure", 1.0), ("frontDefinition", 1.25), ("cameraYaw", 0.4)):
        o.set_parameter(k, v)
    o.set_bool("waterEnabled", True)
    o.reset()
    for frame in range(1, 7):
        o.step(DT, 1)
        if frame not in frame_dumps:
            continue
        ref = o.render_frame(1920, 1080, 256)
        header, arr = read_field_dump(str(out / frame_dumps[frame]["file"]))
        assert header["dims"] == (1920, 1080, 1) and header["kind"] == 4
        got = arr.reshape(1080, 1920, 4)
        if not np.array_equal(got, ref):
            bad = np.argwhere(got != ref)
            raise AssertionError(f"frame {frame}: {len(bad)} bytes differ, first {bad[0].tolist()}")
        ppm = (out / frame_dumps[frame]["image"]).read_bytes()
        assert ppm[:15] == b"P6\n1920 1080\n25" and ppm[-1920 * 1080 * 3:] == ref[..., :3].tobytes()
    o.close()
