"""Per-field error of the f32 step with reduced-precision species storage
against the FP64 oracle, over a long run (test infrastructure: a
measurement, not part of the product).

The HIP kernels equal the FP32 oracle bit for bit at every storage, so the
FP32 oracle with the same species storage stands for the GPU path.  Both
oracles run the same workload side by side; at each checkpoint N the script
reports, per field and channel, the relative L2 error ||a - b|| / ||b|| and
the max-abs error of the FP32 state against the FP64 state (accumulated
error), and at the shadow checkpoints also the error of ONE step taken from
the FP64 state by both (the FP64 fields written into a fresh FP32 oracle,
species rounded to the storage as a store does): the per-step error.

    python oracle/precision_study.py --dims 192 288 192 --steps 600 --storage f16

Default workload: the benchmark's UHD mixed configuration (three seeded
emitters of radius 6-14 cells, seed 2026, 10 % jitter, 96 + 96 Jacobi
sweeps, dt 1/60, Inferno parameters, cinematic uniforms, warm-started f32
pressure).  Prints one JSON line per checkpoint.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from oracle import oracle as orc  # noqa: E402

DT = 1.0 / 60.0
FIELDS = (("flow", orc.FIELD_FLOW), ("species", orc.FIELD_SPECIES), ("pressure", orc.FIELD_PRESSURE),
          ("reaction", orc.FIELD_REACTION))
STORAGE = {"f32": None, "f16": orc.SPECIES_F16, "bf16": orc.SPECIES_BF16}


def seeded(dims, count, radius_range, seed):
    """The benchmark's emitter set (splitmix64 stream; the same sequence as
    fire_amd.solver.seeded_emitters, restated so that this script needs only
    the oracle)."""
    state = [seed & 0xFFFFFFFFFFFFFFFF]

    def nxt():
        state[0] = (state[0] + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
        z = state[0]
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
        z ^= z >> 31
        return (z >> 11) / float(1 << 53)

    nx, ny, nz = dims
    out = []
    for _ in range(count):
        r = radius_range[0] + (radius_range[1] - radius_range[0]) * nxt()
        cx = nx * (0.25 + 0.5 * nxt())
        cz = nz * (0.25 + 0.5 * nxt())
        out.append((cx, 1.5 * r / np.sqrt(2.0), cz, r, 1.0))
    return out


def make(dims, precision, storage, iters, emitters, seed):
    o = orc.Oracle(dims, precision=precision, pressure_iterations=iters, correction_iterations=iters)
    o.set_bool("cinematic", True)
    o.set_heavy_mode(orc.HEAVY_OFF)
    if precision == 32 and STORAGE[storage] is not None:
        o.set_species_storage(STORAGE[storage])
    o.reset()
    o.set_emitters(emitters, 0.1, 0.0, seed)
    return o


def read(o):
    return {name: o.read(fid) for name, fid in FIELDS}


def errors(a, b):
    out = {}
    for name, _ in FIELDS:
        x, y = a[name], b[name]
        chans = [(f"{name}.{'xyzw'[c]}", x[..., c], y[..., c]) for c in range(4)] if x.ndim == 4 else [(name, x, y)]
        for key, u, v in chans:
            d = u - v
            out[key] = [float(np.sqrt(np.sum(d * d)) / max(np.sqrt(np.sum(v * v)), 1e-300)),
                        float(np.max(np.abs(d)))]
    return out


def shadow(dims, storage, iters, emitters, seed, state64, sim_time, frame):
    """One step from the FP64 state, taken by a fresh FP32 and a fresh FP64
    oracle; returns the FP32 error of that step."""
    outs = []
    for precision in (32, 64):
        o = make(dims, precision, storage, iters, emitters, seed)
        o.write(orc.FIELD_FLOW, state64["flow"])
        o.write(orc.FIELD_SPECIES, state64["species"])
        o.write(orc.FIELD_PRESSURE, state64["pressure"])
        o.set_time(sim_time, frame)
        o.step(DT, 1)
        outs.append(read(o))
        o.close()
    return errors(outs[0], outs[1])


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--dims", type=int, nargs=3, default=(192, 288, 192))
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--iters", type=int, default=96)
    ap.add_argument("--storage", default="f16", choices=sorted(STORAGE))
    ap.add_argument("--checkpoints", type=int, nargs="*", default=[1, 10, 30, 60, 150, 300, 450, 600])
    ap.add_argument("--shadows", type=int, nargs="*", default=[60, 300, 600])
    ap.add_argument("--seed", type=int, default=2026)
    args = ap.parse_args()
    dims = tuple(args.dims)
    em = seeded(dims, 3, (6.0, 14.0), args.seed)
    a = make(dims, 32, args.storage, args.iters, em, args.seed)
    b = make(dims, 64, args.storage, args.iters, em, args.seed)
    print(json.dumps({"dims": dims, "steps": args.steps, "iters": [args.iters, args.iters], "storage": args.storage,
                      "emitters": em, "threads": orc.max_threads()}), flush=True)
    t0 = time.time()
    done = 0
    for n in sorted(set(args.checkpoints) | set(args.shadows)):
        if n > args.steps:
            break
        a.step(DT, n - done)
        b.step(DT, n - done)
        done = n
        s64 = read(b)
        rec = {"step": n, "elapsed_s": round(time.time() - t0, 1),
               "fp64_max": {"T": float(s64["flow"][..., 3].max()), "reaction": float(s64["reaction"].max()),
                            "fuel": float(s64["species"][..., 0].max())}}
        if n in args.checkpoints:
            rec["accumulated"] = errors(read(a), s64)
        if n in args.shadows:
            rec["one_step"] = shadow(dims, args.storage, args.iters, em, args.seed, s64, b.simulation_time, n)
        print(json.dumps(rec), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
