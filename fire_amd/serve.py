"""Python mirror of the serving C ABI (include/firex_serve.h).

`FireXServer` keeps many plumes resident on one MI355X, admits queued
requests against an HBM byte budget and a slot limit, and advances every
admitted plume concurrently (a pool of HIP streams, hipGraph-captured steps,
at most a few steps in flight per plume; with batch_plumes, small plumes share
streams and step together, each Jacobi launch covering all of them).  Requests submitted while others
run join at the next pump.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np

from .solver import (DTYPE_F32, DTYPE_F64, DTYPE_U32, FIELDS, PLANS, STORAGE, VECTOR_FIELDS, Emitter, FireXError,
                     _check, load_library)

QUEUED, ACTIVE, DONE, FAILED, RELEASED = range(5)
STATE_NAMES = {QUEUED: "queued", ACTIVE: "active", DONE: "done", FAILED: "failed", RELEASED: "released"}


class ServerConfig(ctypes.Structure):
    _fields_ = [
        ("device", ctypes.c_int32),
        ("memory_budget_bytes", ctypes.c_uint64),
        ("max_active", ctypes.c_int32),
        ("streams", ctypes.c_int32),
        ("plan", ctypes.c_int32),
        ("capture_graphs", ctypes.c_int32),
        ("max_inflight_steps", ctypes.c_int32),
        ("jacobi_autotune", ctypes.c_int32),
        ("batch_plumes", ctypes.c_int32),
        ("batch_max_cells", ctypes.c_int32),
        ("batch_streams", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 1),
    ]


class Request(ctypes.Structure):
    _fields_ = [
        ("dims", ctypes.c_int32 * 3),
        ("pressure_iterations", ctypes.c_int32),
        ("correction_iterations", ctypes.c_int32),
        ("steps", ctypes.c_int32),
        ("dt", ctypes.c_double),
        ("species_storage", ctypes.c_int32),
        ("emitter_count", ctypes.c_int32),
        ("emitters", Emitter * 8),
        ("jitter", ctypes.c_float),
        ("velocity_noise", ctypes.c_float),
        ("seed", ctypes.c_uint64),
        ("keep_result", ctypes.c_int32),
        ("reserved", ctypes.c_int32 * 5),
    ]


class RequestStatus(ctypes.Structure):
    _fields_ = [
        ("state", ctypes.c_int32),
        ("steps_done", ctypes.c_int32),
        ("submit_time", ctypes.c_double),
        ("admit_time", ctypes.c_double),
        ("finish_time", ctypes.c_double),
        ("latency_ms_p50", ctypes.c_double),
        ("latency_ms_p99", ctypes.c_double),
        ("latency_ms_max", ctypes.c_double),
        ("device_bytes", ctypes.c_uint64),
    ]


class ServerStats(ctypes.Structure):
    _fields_ = [
        ("queued", ctypes.c_int32), ("active", ctypes.c_int32), ("done", ctypes.c_int32), ("failed", ctypes.c_int32),
        ("resident_bytes", ctypes.c_uint64),
        ("memory_budget_bytes", ctypes.c_uint64),
        ("steps_completed", ctypes.c_int64),
        ("cell_updates", ctypes.c_double),
        ("elapsed_seconds", ctypes.c_double),
        ("latency_ms_p50", ctypes.c_double),
        ("latency_ms_p99", ctypes.c_double),
        ("latency_ms_max", ctypes.c_double),
        ("grouped_steps", ctypes.c_int64),
        ("group_steps", ctypes.c_int64),
    ]


_P = ctypes.c_void_p
SERVE_SIGNATURES = {
    "firex_server_create": (ctypes.c_int, [ctypes.POINTER(ServerConfig), ctypes.POINTER(_P)]),
    "firex_server_destroy": (None, [_P]),
    "firex_server_submit": (ctypes.c_int, [_P, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_int64)]),
    "firex_server_pump": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_int32)]),
    "firex_server_drain": (ctypes.c_int, [_P]),
    "firex_server_request_status": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.POINTER(RequestStatus)]),
    "firex_server_get_stats": (ctypes.c_int, [_P, ctypes.POINTER(ServerStats)]),
    "firex_server_read_field": (ctypes.c_int, [_P, ctypes.c_int64, ctypes.c_int32, ctypes.c_void_p, ctypes.c_size_t,
                                               ctypes.c_int32]),
    "firex_server_release": (ctypes.c_int, [_P, ctypes.c_int64]),
    "firex_server_clock": (ctypes.c_double, [_P]),
}

_bound = False


def _lib():
    global _bound
    lib = load_library()
    if not _bound:
        for name, (res, args) in SERVE_SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _bound = True
    return lib


def make_request(dims: Sequence[int], steps: int, *, pressure_iterations: int = 96, correction_iterations: int = 96,
                 dt: float = 1.0 / 60.0, species_storage: str = "f16", emitters=(), jitter: float = 0.0,
                 velocity_noise: float = 0.0, seed: int = 0, keep_result: bool = False) -> Request:
    r = Request()
    for i in range(3):
        r.dims[i] = int(dims[i])
    r.pressure_iterations, r.correction_iterations = int(pressure_iterations), int(correction_iterations)
    r.steps, r.dt = int(steps), float(dt)
    r.species_storage = STORAGE[species_storage]
    em = list(emitters)
    if len(em) > 8:
        raise ValueError("at most 8 emitters")
    r.emitter_count = len(em)
    for i, e in enumerate(em):
        r.emitters[i].center[0], r.emitters[i].center[1], r.emitters[i].center[2] = float(e[0]), float(e[1]), float(e[2])
        r.emitters[i].radius, r.emitters[i].amplitude = float(e[3]), float(e[4])
    r.jitter, r.velocity_noise = float(jitter), float(velocity_noise)
    r.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    r.keep_result = int(bool(keep_result))
    return r


class FireXServer:
    def __init__(self, memory_budget_bytes: int = 0, max_active: int = 16, streams: int = 4, plan: str = "fused",
                 capture_graphs: bool = True, max_inflight_steps: int = 1, device: int = 0,
                 jacobi_autotune: bool = False, batch_plumes: bool = False, batch_max_cells: int = 0,
                 batch_streams: int = 0):
        self._lib = _lib()
        cfg = ServerConfig()
        cfg.device = int(device)
        cfg.memory_budget_bytes = int(memory_budget_bytes)
        cfg.max_active = int(max_active)
        cfg.streams = int(streams)
        cfg.plan = PLANS[plan]
        cfg.capture_graphs = int(bool(capture_graphs))
        cfg.max_inflight_steps = int(max_inflight_steps)
        cfg.jacobi_autotune = int(bool(jacobi_autotune))
        cfg.batch_plumes = int(bool(batch_plumes))
        cfg.batch_max_cells = int(batch_max_cells)
        cfg.batch_streams = int(batch_streams)
        h = ctypes.c_void_p()
        _check(self._lib.firex_server_create(ctypes.byref(cfg), ctypes.byref(h)))
        self._h = h
        self.dims = {}

    def close(self) -> None:
        if getattr(self, "_h", None):
            self._lib.firex_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, request: Request) -> int:
        rid = ctypes.c_int64()
        _check(self._lib.firex_server_submit(self._h, ctypes.byref(request), ctypes.byref(rid)))
        self.dims[rid.value] = tuple(request.dims)
        return rid.value

    def pump(self) -> bool:
        busy = ctypes.c_int32()
        _check(self._lib.firex_server_pump(self._h, ctypes.byref(busy)))
        return bool(busy.value)

    def drain(self) -> None:
        _check(self._lib.firex_server_drain(self._h))

    def clock(self) -> float:
        return self._lib.firex_server_clock(self._h)

    def status(self, rid: int) -> dict:
        st = RequestStatus()
        _check(self._lib.firex_server_request_status(self._h, int(rid), ctypes.byref(st)))
        d = {n: getattr(st, n) for n, _ in RequestStatus._fields_}
        d["state_name"] = STATE_NAMES.get(st.state, "?")
        return d

    def stats(self) -> dict:
        st = ServerStats()
        _check(self._lib.firex_server_get_stats(self._h, ctypes.byref(st)))
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}

    def read_field(self, rid: int, name: str, dtype=np.float32) -> np.ndarray:
        nx, ny, nz = self.dims[rid]
        shape = (nz, ny, nx, 4) if name in VECTOR_FIELDS else (nz, ny, nx)
        dt = np.dtype(dtype)
        code = {np.dtype(np.float32): DTYPE_F32, np.dtype(np.float64): DTYPE_F64, np.dtype(np.uint32): DTYPE_U32}[dt]
        out = np.empty(shape, dtype=dt)
        _check(self._lib.firex_server_read_field(self._h, int(rid), FIELDS[name], out.ctypes.data, out.nbytes, code))
        return out

    def release(self, rid: int) -> None:
        _check(self._lib.firex_server_release(self._h, int(rid)))


def poisson_stream(count: int, rate_per_s: float, seed: int, grids: Optional[Sequence] = None,
                   steps_range=(60, 600)) -> list:
    """Seeded Poisson arrivals: (arrival_seconds, dims, steps, emitter seed)."""
    rng = np.random.default_rng(seed)
    grids = grids or [(64, 96, 64), (128, 192, 128), (160, 240, 160), (192, 288, 192)]
    t = 0.0
    out = []
    for i in range(count):
        t += float(rng.exponential(1.0 / rate_per_s))
        g = grids[int(rng.integers(0, len(grids)))]
        steps = int(rng.integers(steps_range[0], steps_range[1] + 1))
        out.append((t, tuple(g), steps, int(rng.integers(0, 2 ** 31))))
    return out


__all__ = ["FireXServer", "FireXError", "make_request", "poisson_stream", "QUEUED", "ACTIVE", "DONE", "FAILED",
           "RELEASED"]
