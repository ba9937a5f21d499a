"""The auditable capture format on the CPU: the field checksum (the C
library's host implementation against an independent numpy restatement of
its definition, and its sensitivity to any change), the `.fxf` dump header
written by the Node driver and read by fire_amd.capture, and the driver's
argument checks."""
import json
import os
import shutil
import subprocess

import numpy as np
import pytest

from fire_amd import checksum_bytes
from fire_amd.capture import read_field_dump

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
ADDON = os.path.join(ROOT, "fire_amd", "firex_napi.node")
needs_node = pytest.mark.skipif(NODE is None or not os.path.exists(ADDON), reason="node or the N-API addon is absent")


def splitmix64(z):
    z = z + np.uint64(0x9E3779B97F4A7C15)
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def checksum_numpy(data: bytes) -> int:
    """sum_i splitmix64((i << 32) | w_i) mod 2^64 over little-endian u32 words
    (a final partial word zero-padded)."""
    words = np.frombuffer(data + b"\0" * (-len(data) % 4), dtype="<u4").astype(np.uint64)
    i = np.arange(words.size, dtype=np.uint64)
    with np.errstate(over="ignore"):
        return int(splitmix64((i << np.uint64(32)) | words).sum(dtype=np.uint64))


@pytest.mark.parametrize("n", [0, 1, 2, 3, 4, 5, 7, 8, 1000, 4099])
def test_host_checksum_matches_its_definition(n):
    data = np.random.default_rng(n).integers(0, 256, n, dtype=np.uint8).tobytes()
    assert checksum_bytes(data) == checksum_numpy(data)


def test_checksum_sees_every_change():
    rng = np.random.default_rng(3)
    data = bytearray(rng.integers(0, 256, 4096, dtype=np.uint8).tobytes())
    base = checksum_bytes(bytes(data))
    for pos in rng.integers(0, len(data), 64):
        flipped = bytearray(data)
        flipped[pos] ^= 1 << int(rng.integers(0, 8))
        assert checksum_bytes(bytes(flipped)) != base
    swapped = bytearray(data)
    swapped[0:4], swapped[8:12] = data[8:12], data[0:4]
    assert data[0:4] != data[8:12] and checksum_bytes(bytes(swapped)) != base


def node(script, *args):
    res = subprocess.run([NODE, "-e", script, *args], capture_output=True, text=True, timeout=60, cwd=ROOT)
    return res


@needs_node
def test_dump_header_written_by_the_driver_reads_back(tmp_path):
    data = np.arange(24, dtype=np.float16).tobytes()
    src = tmp_path / "payload.bin"
    src.write_bytes(data)
    out = tmp_path / "frame-000007-species.fxf"
    script = ("const r = require('./fire_amd/js/record_fire_hip.js'); const fs = require('fs');"
              "const fx = require('./fire_amd/js/firex_hip_solver.js');"
              "const data = fs.readFileSync(process.argv[1]);"
              "const h = r.encodeDumpHeader({fieldId: 1, kindCode: 1, components: 4, dimensions: [3, 2, 1], records: 6,"
              " storedBytes: data.length, frame: 7, simulationTime: 0.125, checksum: fx.native.checksumBytes(data),"
              " field: 'species'});"
              "fs.writeFileSync(process.argv[2], Buffer.concat([h, data]));")
    res = node(script, str(src), str(out))
    assert res.returncode == 0, res.stderr
    header, arr = read_field_dump(str(out))
    assert (header["field"], header["field_id"], header["kind"], header["components"]) == ("species", 1, 1, 4)
    assert header["dims"] == (3, 2, 1) and header["records"] == 6 and header["frame"] == 7
    assert header["simulation_time"] == 0.125 and header["stored_bytes"] == len(data)
    assert header["checksum"] == checksum_numpy(data)
    np.testing.assert_array_equal(arr.ravel(), np.arange(24, dtype=np.float16))


@needs_node
def test_driver_arguments():
    ok = node("const r = require('./fire_amd/js/record_fire_hip.js');"
              "console.log(JSON.stringify(r.parseArguments(['--scene', 'base-spray', '--duration', '0.25', "
              "'--dims', '32x48x32', '--dump-every', '5', '--dump-fields', 'flow,particles'])));")
    assert ok.returncode == 0, ok.stderr
    opts = json.loads(ok.stdout)
    assert opts["scene"] == "base-spray" and opts["exactDims"] == [32, 48, 32] and opts["dumpEvery"] == 5
    assert opts["dumpFields"] == ["flow", "particles"]
    for bad, msg in ((["--scene", "bonfire"], "Unknown scene"), (["--duration", "0.01"], "whole number of 60 fps"),
                     (["--dump-fields", "velocity"], "Unknown field"), (["--solver", "4k"], "Unknown solver tier")):
        res = node("const r = require('./fire_amd/js/record_fire_hip.js');"
                   f"try {{ r.parseArguments({json.dumps(bad)}); }} catch (e) {{ console.log(e.message); }}")
        assert msg in res.stdout, (bad, res.stdout, res.stderr)


@needs_node
def test_driver_frame_arguments():
    ok = node("const r = require('./fire_amd/js/record_fire_hip.js');"
              "console.log(JSON.stringify(r.parseArguments(['--frame-every', '3', '--resolution', '4K', "
              "'--camera-yaw', '0.5', '--camera-distance', '3.5', '--optical-detail', '256'])));")
    assert ok.returncode == 0, ok.stderr
    opts = json.loads(ok.stdout)
    assert opts["frameEvery"] == 3 and opts["outputPreset"] == "uhd" and opts["opticalDetailTarget"] == 256
    assert opts["camera"] == {"cameraYaw": 0.5, "cameraDistance": 3.5}
    # frames are off unless asked for; the resolution aliases follow the reference recorder's
    res = node("const r = require('./fire_amd/js/record_fire_hip.js');"
               "console.log(JSON.stringify([r.parseArguments([]).frameEvery, "
               "['1080p', '1440', 'qhd', '2160p'].map((v) => r.parseArguments(['--resolution', v]).outputPreset)]));")
    assert json.loads(res.stdout) == [None, ["hd", "qhd", "qhd", "uhd"]]
    for bad, msg in ((["--resolution", "8k"], "Unknown output resolution"), (["--frame-every", "-1"], "non-negative"),
                     (["--optical-detail", "300"], "optical-detail must be"), (["--camera-yaw", "x"], "finite")):
        res = node("const r = require('./fire_amd/js/record_fire_hip.js');"
                   f"try {{ r.parseArguments({json.dumps(bad)}); }} catch (e) {{ console.log(e.message); }}")
        assert msg in res.stdout, (bad, res.stdout, res.stderr)


@needs_node
def test_ppm_of_a_frame(tmp_path):
    out = tmp_path / "f.ppm"
    res = node("const r = require('./fire_amd/js/record_fire_hip.js'); const fs = require('fs');"
               "const px = new Uint8Array([1, 2, 3, 255, 4, 5, 6, 255, 7, 8, 9, 255, 10, 11, 12, 255, 13, 14, 15, 255, "
               "16, 17, 18, 255]);"
               "fs.writeFileSync(process.argv[1], r.encodePpm(px, 3, 2));", str(out))
    assert res.returncode == 0, res.stderr
    assert out.read_bytes() == b"P6\n3 2\n255\n" + bytes(range(1, 19))
