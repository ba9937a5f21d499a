"""ctypes front for the CPU oracle (test infrastructure only).

The oracle restates the reference Fire-X step (src/sim/firex/shaders.ts,
src/sim/firex/FireXEngine.ts) on the CPU in FP32 (WGSL operation order) and
FP64.  Only the tests, the smoke check and the CPU-baseline leg of bench.py
use it; the product never imports it.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# FIREX_ORACLE_LIB points the tests at another build of the oracle (used to
# check that planted mistakes in the restatement fail the pinning tests).
LIB_PATH = os.environ.get("FIREX_ORACLE_LIB") or os.path.join(_HERE, "liboracle_firex.so")

FIELD_FLOW, FIELD_SPECIES, FIELD_PRESSURE, FIELD_CORRECTION = 0, 1, 2, 3
FIELD_DIVERGENCE, FIELD_REACTION, FIELD_LIQUID, FIELD_FLOW_PREDICTED = 4, 5, 6, 7
VECTOR_FIELDS = (FIELD_FLOW, FIELD_SPECIES, FIELD_FLOW_PREDICTED)
# per-particle fields: (count, 8) positionAge + velocityMass; (count, 4) density
FIELD_PARTICLES, FIELD_PARTICLE_DENSITY = 11, 12
PARTICLE_FIELDS = {FIELD_PARTICLES: 8, FIELD_PARTICLE_DENSITY: 4}
DROPLET_DENSITY, DROPLET_INTEGRATE, DROPLET_DEPOSIT = 0, 1, 2

SPECIES_F32, SPECIES_F16, SPECIES_BF16 = 0, 1, 2
HEAVY_AUTO, HEAVY_OFF, HEAVY_F16, HEAVY_F32 = 0, 1, 2, 3

_lib = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"oracle library not built: {LIB_PATH} (run make -C oracle)")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i, d, f = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float
    u64, u32 = ctypes.c_uint64, ctypes.c_uint32
    dp = ctypes.POINTER(ctypes.c_double)
    fp = ctypes.POINTER(ctypes.c_float)
    sig = {
        "orc_create": (vp, [i, i, i, i, i, i]),
        "orc_destroy": (None, [vp]),
        "orc_set_parameter": (i, [vp, ctypes.c_char_p, d]),
        "orc_set_bool": (i, [vp, ctypes.c_char_p, i]),
        "orc_get_parameter": (d, [vp, ctypes.c_char_p]),
        "orc_set_time_step": (None, [vp, d]),
        "orc_set_time": (None, [vp, d, d]),
        "orc_set_species_storage": (None, [vp, i]),
        "orc_set_heavy_mode": (None, [vp, i]),
        "orc_set_emitters": (i, [vp, i, fp, f, f, u64]),
        "orc_reset": (None, [vp]),
        "orc_step": (None, [vp, d, i]),
        "orc_read_field": (i, [vp, i, dp]),
        "orc_write_field": (i, [vp, i, dp]),
        "orc_simulation_time": (d, [vp]),
        "orc_species_index": (i, [vp]),
        "orc_pressure_index": (i, [vp]),
        "orc_pass_log": (i, [vp, ctypes.c_char_p, i]),
        "orc_uniforms": (None, [vp, d, fp]),
        "orc_kernel_divergence": (None, [vp, d, i]),
        "orc_kernel_jacobi": (None, [vp, d, i]),
        "orc_kernel_project": (None, [vp, d, i]),
        "orc_kernel_simulate": (None, [vp, d]),
        "orc_trilinear_flow": (None, [vp, d, d, d, d, dp]),
        "orc_emitter": (d, [vp, d, d, d, d]),
        "orc_round_f16": (f, [f]),
        "orc_round_bf16": (f, [f]),
        "orc_seeded_uniform": (f, [u64, u64, u64, u32]),
        "orc_f32_exp": (f, [f]),
        "orc_f32_sin": (f, [f]),
        "orc_f32_cos": (f, [f]),
        "orc_f32_pow": (f, [f, f]),
        "orc_f32_log2": (f, [f]),
        "orc_f32_exp2": (f, [f]),
        "orc_f32_eval": (None, [i, fp, fp, fp, ctypes.c_int64]),
        "orc_stencil_weights": (None, [vp, d, dp]),
        "orc_pack_fields": (i, [vp, dp, dp, dp]),
        "orc_set_particle_count": (i, [vp, i]),
        "orc_particle_count": (i, [vp]),
        "orc_particle_index": (i, [vp]),
        "orc_kernel_droplets": (i, [vp, d, i]),
        "orc_write_particle_initial_state": (None, [vp, i]),
        "orc_write_particle_buffers": (None, [vp, dp, dp]),
        "orc_render_volume": (i, [vp, i, i, i, dp]),
        "orc_detail_volume": (None, [ctypes.POINTER(ctypes.c_uint8)]),
        "orc_render_frame": (i, [vp, i, i, i, ctypes.POINTER(ctypes.c_uint8)]),
        "orc_present": (i, [ctypes.POINTER(ctypes.c_float), i, i, ctypes.c_float, i, ctypes.POINTER(ctypes.c_uint8),
                            ctypes.POINTER(ctypes.c_float)]),
        "orc_set_threads": (None, [i]),
        "orc_max_threads": (i, []),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def detail_volume() -> np.ndarray:
    """The 64^3 rgba8 periodic detail volume (createPeriodicDetailVolume),
    shape (64, 64, 64, 4) indexed [z, y, x, channel]."""
    out = np.empty((64, 64, 64, 4), dtype=np.uint8)
    load().orc_detail_volume(out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return out


def present(scene, exposure: float, cinematic: bool):
    """The presentation passes over a (height, width, 4) float scene image:
    returns (frame as (height, width, 4) uint8, bloom as (ceil(height/4),
    ceil(width/4), 4) float32, or None for the standard presentation)."""
    scene = np.ascontiguousarray(scene, dtype=np.float32)
    h, w = scene.shape[:2]
    frame = np.empty((h, w, 4), dtype=np.uint8)
    bloom = np.zeros((max((h + 3) // 4, 1), max((w + 3) // 4, 1), 4), dtype=np.float32) if cinematic else None
    fp = ctypes.POINTER(ctypes.c_float)
    if load().orc_present(scene.ctypes.data_as(fp), w, h, float(exposure), int(bool(cinematic)),
                          frame.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                          bloom.ctypes.data_as(fp) if bloom is not None else None):
        raise ValueError("present needs a non-empty scene image")
    return frame, bloom


def max_threads() -> int:
    """OpenMP threads the oracle's loops use (OMP_NUM_THREADS or the host's cores)."""
    return load().orc_max_threads()


def set_threads(n: int) -> int:
    """Set the OpenMP thread count of the oracle's loops; returns the previous one."""
    lib = load()
    prev = lib.orc_max_threads()
    lib.orc_set_threads(int(n))
    return prev


F32_FUNCS = {"exp": 0, "exp2": 1, "log2": 2, "sin": 3, "cos": 4, "pow": 5, "smoothstep": 6}


def f32_eval(name: str, x, y=None) -> np.ndarray:
    """Evaluate an f32 WGSL built-in of the oracle's FP32 layer over an array."""
    lib = load()
    xa = np.ascontiguousarray(x, dtype=np.float32)
    ya = np.ascontiguousarray(y if y is not None else np.zeros(1), dtype=np.float32)
    out = np.empty_like(xa)
    fp = ctypes.POINTER(ctypes.c_float)
    lib.orc_f32_eval(F32_FUNCS[name], xa.ctypes.data_as(fp), ya.ctypes.data_as(fp), out.ctypes.data_as(fp), xa.size)
    return out


class Oracle:
    """One oracle solver instance (FP32 or FP64)."""

    def __init__(self, dims, precision=64, pressure_iterations=24, correction_iterations=24,
                 time_step=1.0 / 60.0):
        self.lib = load()
        self.dims = tuple(int(v) for v in dims)
        self.cells = self.dims[0] * self.dims[1] * self.dims[2]
        self.precision = precision
        self.h = self.lib.orc_create(*self.dims, precision, pressure_iterations, correction_iterations)
        if not self.h:
            raise ValueError("invalid oracle configuration")
        self.lib.orc_set_time_step(self.h, time_step)

    def close(self):
        if self.h:
            self.lib.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_parameter(self, name: str, value: float) -> None:
        self.lib.orc_set_parameter(self.h, name.encode(), float(value))

    def set_bool(self, name: str, value: bool) -> None:
        self.lib.orc_set_bool(self.h, name.encode(), int(bool(value)))

    def get_parameter(self, name: str) -> float:
        return self.lib.orc_get_parameter(self.h, name.encode())

    def set_time(self, simulation_time: float, frame_index: int = 0) -> None:
        """Set the clock the next step's uniforms carry (time.y, time.z)."""
        self.lib.orc_set_time(self.h, float(simulation_time), float(frame_index))

    def set_species_storage(self, mode: int) -> None:
        self.lib.orc_set_species_storage(self.h, mode)

    def set_heavy_mode(self, mode: int) -> None:
        self.lib.orc_set_heavy_mode(self.h, mode)

    def set_emitters(self, emitters, jitter=0.0, velocity_noise=0.0, seed=0) -> None:
        arr = np.ascontiguousarray(np.asarray(emitters, dtype=np.float32).reshape(-1, 5))
        rc = self.lib.orc_set_emitters(self.h, arr.shape[0], arr.ctypes.data_as(ctypes.POINTER(ctypes.c_float)),
                                       float(jitter), float(velocity_noise), int(seed))
        if rc:
            raise ValueError("at most 8 emitters")

    def reset(self) -> None:
        self.lib.orc_reset(self.h)

    def step(self, dt: float, count: int = 1) -> None:
        self.lib.orc_step(self.h, float(dt), int(count))

    def set_particle_count(self, count: int) -> None:
        """Particle count of the resource set (384-4096 by tier); rebuilds
        the particle state (FireXEngine.ts:1917-1931, 1976-2034)."""
        if self.lib.orc_set_particle_count(self.h, int(count)):
            raise ValueError("particle count must be non-negative")

    @property
    def particle_count(self) -> int:
        return self.lib.orc_particle_count(self.h)

    @property
    def particle_index(self) -> int:
        return self.lib.orc_particle_index(self.h)

    def kernel_droplets(self, dt: float, which: int) -> None:
        """One droplet pass (DROPLET_DENSITY / _INTEGRATE / _DEPOSIT) on the
        current particle buffers with the uniforms of a dt step."""
        if self.lib.orc_kernel_droplets(self.h, float(dt), int(which)):
            raise ValueError(f"unknown droplet pass {which}")

    def render_volume(self, width: int, height: int, ray_steps: int) -> np.ndarray:
        """The volume ray marcher (qualityShaders.ts:374-575) over freshly
        packed volumes: (height, width, 4) rgba, f16-rounded, row 0 at the
        top.  FP32 oracle only."""
        out = np.empty((height, width, 4), dtype=np.float64)
        if self.lib.orc_render_volume(self.h, int(width), int(height), int(ray_steps),
                                      out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))):
            raise ValueError("render_volume needs the FP32 oracle and a positive size and step count")
        return out

    def render_frame(self, width: int, height: int, ray_steps: int) -> np.ndarray:
        """A presented frame (the ray march, then bloom + filmic composite +
        dither with the cinematic presentation, else the scene image) as
        (height, width, 4) uint8.  FP32 oracle only."""
        out = np.empty((height, width, 4), dtype=np.uint8)
        if self.lib.orc_render_frame(self.h, int(width), int(height), int(ray_steps),
                                     out.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))):
            raise ValueError("render_frame needs the FP32 oracle and a positive size and step count")
        return out

    def write_particle_buffers(self, current, previous) -> None:
        """The live particle buffer and the other one (the previous state
        the deposition sweeps from), each (count, 8)."""
        shape = (self.particle_count, 8)
        c = np.ascontiguousarray(np.asarray(current, dtype=np.float64).reshape(shape))
        q = np.ascontiguousarray(np.asarray(previous, dtype=np.float64).reshape(shape))
        dp = ctypes.POINTER(ctypes.c_double)
        self.lib.orc_write_particle_buffers(self.h, c.ctypes.data_as(dp), q.ctypes.data_as(dp))

    def write_particle_initial_state(self, active: bool) -> None:
        self.lib.orc_write_particle_initial_state(self.h, int(bool(active)))

    def _shape(self, field):
        if field in PARTICLE_FIELDS:
            return (self.particle_count, PARTICLE_FIELDS[field])
        nx, ny, nz = self.dims
        return (nz, ny, nx, 4) if field in VECTOR_FIELDS else (nz, ny, nx)

    def read(self, field: int) -> np.ndarray:
        shape = self._shape(field)
        out = np.empty(shape, dtype=np.float64)
        rc = self.lib.orc_read_field(self.h, field, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc:
            raise ValueError(f"unknown field {field}")
        return out

    def pack_fields(self):
        """packFields (qualityShaders.ts:37-73): (flow, chemistry, media)
        volumes of shape (nz, ny, nx, 4); the FP32 oracle rounds each
        component to f16 as the rgba16float texture store does."""
        nx, ny, nz = self.dims
        outs = [np.empty((nz, ny, nx, 4), dtype=np.float64) for _ in range(3)]
        dp = ctypes.POINTER(ctypes.c_double)
        self.lib.orc_pack_fields(self.h, *(o.ctypes.data_as(dp) for o in outs))
        return tuple(outs)

    def write(self, field: int, values) -> None:
        arr = np.ascontiguousarray(np.asarray(values, dtype=np.float64).reshape(self._shape(field)))
        rc = self.lib.orc_write_field(self.h, field, arr.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        if rc:
            raise ValueError(f"field {field} is not writable")

    @property
    def simulation_time(self) -> float:
        return self.lib.orc_simulation_time(self.h)

    @property
    def species_index(self) -> int:
        return self.lib.orc_species_index(self.h)

    @property
    def pressure_index(self) -> int:
        return self.lib.orc_pressure_index(self.h)

    def pass_log(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        self.lib.orc_pass_log(self.h, buf, 4096)
        return buf.value.decode()

    def uniforms(self, dt: float) -> np.ndarray:
        out = np.zeros(44, dtype=np.float32)
        self.lib.orc_uniforms(self.h, float(dt), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        return out

    # single kernels -------------------------------------------------------
    def kernel_divergence(self, dt, from_predicted=False):
        self.lib.orc_kernel_divergence(self.h, float(dt), int(from_predicted))

    def kernel_jacobi(self, dt, f16=False):
        self.lib.orc_kernel_jacobi(self.h, float(dt), int(f16))

    def kernel_project(self, dt, src=0):
        self.lib.orc_kernel_project(self.h, float(dt), int(src))

    def kernel_simulate(self, dt):
        self.lib.orc_kernel_simulate(self.h, float(dt))

    def trilinear_flow(self, dt, position):
        out = (ctypes.c_double * 4)()
        self.lib.orc_trilinear_flow(self.h, float(dt), *[float(v) for v in position], out)
        return np.array(out[:])

    def stencil_weights(self, dt=0.0):
        out = (ctypes.c_double * 3)()
        self.lib.orc_stencil_weights(self.h, float(dt), out)
        return tuple(out[:])

    def emitter(self, dt, normalized_position):
        return self.lib.orc_emitter(self.h, float(dt), *[float(v) for v in normalized_position])
