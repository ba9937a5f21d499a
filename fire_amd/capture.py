"""Readers for the auditable offline capture written by the Node driver
fire_amd/js/record_fire_hip.js: `.fxf` field dumps and the `.json`
manifest.  verify_capture() re-checks a capture from disk with the library's
host checksum (firex_checksum_bytes).

Dump layout (little-endian): a 128-byte header — magic "FXFIELD1", header
bytes, field id, storage kind (0 f32, 1 f16, 2 bf16, 3 u32, 4 u8), components,
grid nx/ny/nz (a presented frame: width, height, 1), records, stored bytes, frame, simulation time, checksum, field
name — then the field's stored bytes verbatim.
"""
from __future__ import annotations

import json
import os
import struct

import numpy as np

from .solver import checksum_bytes

DUMP_MAGIC = b"FXFIELD1"
_HEADER = struct.Struct("<8sIIII3iIQQQdQ32s16x")
_KIND_DTYPES = {0: np.float32, 1: np.float16, 2: np.uint16, 3: np.uint32, 4: np.uint8}  # bf16 as its bit pattern


def read_field_dump(path: str):
    """(header dict, array of shape (records, components)) of one `.fxf`
    dump; bf16 data come back as uint16 bit patterns."""
    with open(path, "rb") as f:
        raw = f.read()
    if raw[:8] != DUMP_MAGIC:
        raise ValueError(f"{path}: not a Fire-X field dump")
    (magic, header_bytes, field_id, kind, comps, nx, ny, nz, _pad, records, stored, frame, sim_time, checksum,
     name) = _HEADER.unpack_from(raw, 0)
    data = raw[header_bytes:]
    if len(data) != stored:
        raise ValueError(f"{path}: {len(data)} data bytes, header says {stored}")
    header = {"field_id": field_id, "kind": kind, "components": comps, "dims": (nx, ny, nz), "records": records,
              "stored_bytes": stored, "frame": frame, "simulation_time": sim_time, "checksum": checksum,
              "field": name.rstrip(b"\0").decode("ascii"), "data": data}
    arr = np.frombuffer(data, dtype=_KIND_DTYPES[kind]).reshape(records, comps)
    return header, arr


def verify_capture(manifest_path: str) -> dict:
    """Re-verify a capture: every dump's checksum, recomputed from disk,
    equals its header and the device checksum the manifest recorded for that
    frame.  Returns counts; raises ValueError on a mismatch."""
    with open(manifest_path) as f:
        manifest = json.load(f)
    if manifest.get("status") != "verified":
        raise ValueError(f"{manifest_path}: status {manifest.get('status')!r}")
    root = manifest["outputPath"]
    if not os.path.isdir(root):
        root = manifest_path[:-len(".json")]
    frames = {rec["frame"]: rec for rec in manifest["frames"]}
    total = 0
    for d in manifest["dumps"]:
        header, _ = read_field_dump(os.path.join(root, d["file"]))
        on_disk = checksum_bytes(header["data"])
        device = int(frames[d["frame"]]["checksums"].get(d["field"], d["checksum"]), 16)
        if not (on_disk == header["checksum"] == device == int(d["checksum"], 16)):
            raise ValueError(f"{d['file']}: checksum {on_disk:016x} on disk, {header['checksum']:016x} in the "
                             f"header, {device:016x} from the device")
        total += header["stored_bytes"]
    return {"dumps": len(manifest["dumps"]), "bytes": total, "frames": len(manifest["frames"])}
