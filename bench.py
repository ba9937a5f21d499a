"""Benchmark of the Fire-X fixed step on one MI355X.

Workload (default "uhd-mixed"): 192 x 288 x 192 grid (10,616,832 cells), 96
primary + 96 residual Jacobi sweeps per step, dt = 1/60, species stored f16
(velocity, temperature, pressure and all arithmetic f32), three seeded
Gaussian emitters of radius 6-14 cells near the grid base, Inferno scene
parameters, zero initial velocity.  Inputs are synthetic and resident on the
device before the timed region.

Output: one JSON line with cell-updates/s (cells x steps / wall time of the
timed steps), the roofline of the dominant kernel (HIP-event timing of its
launches on the solver's stream), per-pass device times, the plain f32
reference-structured GPU path for comparison, and the FP32 CPU oracle
timed on the host's cores over a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E, published
DT = 1.0 / 60.0

# Algorithmic (compulsory) bytes per cell of each pass, f32 reference
# structure: simulate reads flow 16 + species 16 and writes flow 16 +
# species 16 + reaction 4 (no liquid input); divergence 16 + 4; one Jacobi
# sweep reads p and div and writes p; projection reads flow 16 + p 4 and
# writes flow 16; the fused projection also writes div 4; copyFlow 32.
# Algorithmic HBM bytes per cell and launch (each input read once, each
# output written once).  A temporally blocked Jacobi launch reads p and div
# and writes p once per cell whatever its number of sweeps: 12 B (f32 p).
BYTES_PER_CELL = {
    "simulate_f32": 68, "simulate_f16": 52, "divergence": 20, "jacobi_launch": 12,
    "project": 36, "fill_copy": 32,
}

WORKLOADS = {
    "uhd-mixed": dict(dims=(192, 288, 192), iters=(96, 96), storage="f16", emitters=3, radius=(6.0, 14.0),
                      seed=2026, jitter=0.1, velocity_noise=0.0),
    "hd-plain": dict(dims=(128, 192, 128), iters=(96, 96), storage="f32", emitters=1, radius=(8 * 2 ** 0.5,) * 2,
                     seed=2026, jitter=0.1, velocity_noise=0.0),
    "cpu-plain": dict(dims=(64, 96, 64), iters=(24, 24), storage="f32", emitters=1, radius=(4 * 2 ** 0.5,) * 2,
                      seed=2026, jitter=0.1, velocity_noise=0.0),
    # 131.1 M cells: the >= 64 Mi cell path (f16 pressure), bf16 scalars,
    # sigma 32 cells, +-0.5 seeded noise on the injected velocity
    "plume-508": dict(dims=(508, 508, 508), iters=(96, 96), storage="bf16", emitters=1,
                      radius=(32 * 2 ** 0.5,) * 2, seed=2026, jitter=0.1, velocity_noise=0.5),
}


def make_solver(wl: dict, plan: str, sweeps: int = 0, autotune: bool = False, chunk: int = 0, heavy: str = "off",
                waves: int = 0):
    from fire_amd import FireXHipSolver, seeded_emitters

    s = FireXHipSolver(wl["dims"], tier="uhd", pressure_iterations=wl["iters"][0],
                       correction_iterations=wl["iters"][1], species_storage=wl["storage"], plan=plan,
                       heavy_mode=heavy, jacobi_sweeps_per_launch=sweeps, jacobi_chunk_z=chunk,
                       jacobi_autotune=autotune, jacobi_ring_waves=waves)
    em = seeded_emitters(wl["dims"], wl["emitters"], radius_range=wl["radius"], seed=wl["seed"])
    s.set_emitters(em, jitter=wl["jitter"], velocity_noise=wl["velocity_noise"], seed=wl["seed"])
    return s


def time_steps(solver, steps: int, warmup: int) -> float:
    solver.step(DT, warmup)
    solver.sync()
    t0 = time.perf_counter()
    solver.step(DT, steps)
    solver.sync()
    return time.perf_counter() - t0


def pass_profile(solver, steps: int, cells: int, storage: str, f16_pressure: bool = False) -> dict:
    # f16 pressure (the >= 64 Mi cell path) moves 2 B instead of 4 per p access
    pressure_bytes = {"jacobi_launch": 8, "project": 34} if f16_pressure else {}
    solver.set_pass_timing(True)
    solver.step(DT, steps)
    t = solver.pass_timing()
    solver.set_pass_timing(False)
    out = {}
    for name in ("simulate", "divergence", "jacobi", "project", "fill_copy"):
        ms, n = t[name]["ms"], t[name]["launches"]
        if n == 0:
            continue
        if name == "simulate":
            bpc = BYTES_PER_CELL["simulate_f16" if storage != "f32" else "simulate_f32"]
            units = n
        else:
            key = "jacobi_launch" if name == "jacobi" else name
            bpc = pressure_bytes.get(key, BYTES_PER_CELL[key])
            units = n
        gbs = bpc * cells * units / (ms * 1e-3) / 1e9 if ms > 0 else 0.0
        out[name] = {"ms_per_step": ms / t["steps"], "launches_per_step": n / t["steps"],
                     "avg_launch_us": ms / n * 1e3, "algorithmic_GBps": round(gbs, 1)}
        if name == "jacobi" and ms > 0:
            # what one-sweep-per-launch kernels would have to stream to match
            out[name]["sweep_equivalent_GBps"] = round(bpc * cells * t["jacobi_sweeps"] / (ms * 1e-3) / 1e9, 1)
    out["jacobi_sweeps_per_step"] = t["jacobi_sweeps"] / max(1, t["steps"])
    out["_raw"] = t
    return out


def pack_profile(solver, cells: int, storage: str, reps: int = 10) -> dict:
    """packFields (the render volumes) timed alone: flow 16 B + species
    (8 or 16 B) + reaction 4 B read, three f16 texels (24 B) written."""
    bpc = 16 + (16 if storage == "f32" else 8) + 4 + 24
    solver.set_pass_timing(True)
    for _ in range(reps):
        solver.pack_fields()
    t = solver.pass_timing()
    solver.set_pass_timing(False)
    ms, n = t["pack"]["ms"], t["pack"]["launches"]
    return {"avg_launch_us": ms / n * 1e3, "algorithmic_bytes_per_cell": bpc,
            "algorithmic_GBps": round(bpc * cells * n / (ms * 1e-3) / 1e9, 1), "launches": n}


def droplet_profile(wl: dict, steps: int, warmup: int, plan: dict, render: bool = True) -> dict:
    """The same workload with the hose on (waterFlow 0.8, spray nozzle,
    waterEnabled): the uhd tier's 4096 SPH droplets advance every step
    (density and integration each test all N^2 ordered pairs; deposition
    scatters into the liquid field that simulate reads).  The Jacobi plan is
    the one the dry run used."""
    s = make_solver(wl, "fused", plan["sweeps_per_launch"], chunk=plan["chunk_z"], waves=plan.get("waves", 0))
    for k, v in (("waterFlow", 0.8), ("nozzleType", 1), ("sprayAngle", 24.0)):
        s.set_parameter(k, v)
    s.set_parameter("waterEnabled", True)
    n = s.info["particle_count"]
    el = time_steps(s, steps, warmup)
    s.set_pass_timing(True)
    s.step(DT, 10)
    t = s.pass_timing()
    s.set_pass_timing(False)
    liquid = float(s.read_field("liquid", dtype="uint32").sum()) / 4096.0
    frame = None
    if render:
        # a 4K presented frame with the droplet streaks drawn over the scene
        w, h, rsteps = RENDER_4K
        s.render_frame(w, h, rsteps, read=False)  # warm: buffers
        s.set_pass_timing(True)
        for _ in range(3):
            s.render_frame(w, h, rsteps, read=False)
        ft = s.pass_timing()
        s.set_pass_timing(False)
        frame = {k: round(ft[k]["ms"] / 3, 3) for k in ("pack", "render", "streaks", "present")}
        frame["streak_launches"] = ft["streaks"]["launches"] // 3
        frame["output"] = f"{w}x{h}"
    s.close()
    d = t["droplets"]
    us = d["ms"] / t["steps"] * 1e3
    return {"particles": n, "nozzle": "spray", "steps": steps, "ms_per_step_with_water": el / steps * 1e3,
            "droplet_passes_us_per_step": round(us, 2), "launches_per_step": d["launches"] / t["steps"],
            "pair_tests_per_s": 2.0 * n * n / (us * 1e-6) if us > 0 else None,
            "deposited_liquid_mass": round(liquid, 3), "presented_4k_frame_ms": frame}


RENDER_4K = (3840, 2160, 384)  # the reference's 4K offline output and the uhd tier's ray budget
REFERENCE_508_FRAMES_PER_S = 0.21  # the reference's README.md:64-67: one 508^3 step + 4K presentation per frame


def render_profile(solver, frames: int = 3) -> dict:
    """The volume ray marcher at the 4K offline size (pack + march), in the
    cinematic beauty view (empty-space skipping on) and the temperature view
    (every sample shaded), then the whole presented beauty frame (pack, march,
    bloom and composite): device time per frame from HIP events."""
    w, h, steps = RENDER_4K
    out = {"width": w, "height": h, "ray_steps": steps}
    for view, name in ((0, "beauty"), (1, "temperature")):
        solver.set_parameter("viewMode", view)
        solver.render_volume(w, h, steps)  # warm: allocations, detail volume
        solver.set_pass_timing(True)
        for _ in range(frames):
            solver.render_volume(w, h, steps)
        t = solver.pass_timing()
        solver.set_pass_timing(False)
        ms = t["render"]["ms"] / frames
        out[name] = {"render_ms": round(ms, 3), "pack_ms": round(t["pack"]["ms"] / frames, 3),
                     "mpixel_per_s": round(w * h / (ms * 1e-3) / 1e6, 1),
                     "ray_budget_samples_per_s": w * h * steps / (ms * 1e-3)}
    solver.set_parameter("viewMode", 0)
    # the presented frame: ray march, bloom and composite (cinematic)
    solver.render_frame(w, h, steps)  # warm: presentation buffers
    solver.set_pass_timing(True)
    for _ in range(frames):
        solver.render_frame(w, h, steps, read=False)
    t = solver.pass_timing()
    solver.set_pass_timing(False)
    ms = {k: t[k]["ms"] / frames for k in ("pack", "render", "present")}
    out["presented_frame"] = {"pack_ms": round(ms["pack"], 3), "render_ms": round(ms["render"], 3),
                              "present_ms": round(ms["present"], 3), "device_ms": round(sum(ms.values()), 3),
                              "frames_per_s_device": round(1e3 / sum(ms.values()), 1)}
    return out


def audit_profile(solver, reps: int = 20) -> dict:
    """The capture audit: device checksums of every solver field (wall time
    per call, launch + 8-byte readback + sync included) and of the flow field
    alone as an HBM stream (16 B/cell read once)."""
    names = ("flow", "species", "pressure", "correction", "divergence", "reaction")
    out = {}
    for name in names:
        solver.field_checksum(name)  # warm
        t0 = time.perf_counter()
        for _ in range(reps):
            solver.field_checksum(name)
        us = (time.perf_counter() - t0) / reps * 1e6
        nbytes = solver.field_layout(name)["stored_bytes"]
        out[name] = {"us_per_call": round(us, 2), "bytes": nbytes, "GBps_wall": round(nbytes / (us * 1e-6) / 1e9, 1)}
    out["all_fields_us"] = round(sum(v["us_per_call"] for v in out.values()), 2)
    return out


def measured_traffic(workload: str, plan: dict):
    """HBM bytes per Jacobi launch for this plan from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a copy kernel), or
    None when the plan was not profiled."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                        "k_jacobi_ring_uhd_mixed_hbm_traffic.json")
    if workload != "uhd-mixed" or not os.path.exists(path):
        return None
    with open(path) as f:
        table = json.load(f)
    def key(p):
        return (p["kernel"], p["sweeps_per_launch"], p["chunk_z"], p["prefetch"], p.get("tile_height", 32),
                p.get("waves", 8))

    for p in table["plans"]:
        if key(p) == key(plan):
            return p["hbm_bytes_per_launch"]
    return None


def cpu_baseline(wl: dict, sample_steps: int) -> dict:
    """The FP32 oracle (the reference step restated on the CPU, same passes
    and order) on the host: per-step median over sample_steps steps after two
    warm-up steps on all OpenMP threads, then one step on a single thread."""
    import statistics

    from fire_amd import seeded_emitters
    from oracle import oracle as orc

    cores = orc.max_threads()
    o = orc.Oracle(wl["dims"], precision=32, pressure_iterations=wl["iters"][0],
                   correction_iterations=wl["iters"][1])
    o.set_bool("cinematic", True)
    o.set_heavy_mode(orc.HEAVY_OFF)
    if wl["storage"] != "f32":
        o.set_species_storage({"f16": orc.SPECIES_F16, "bf16": orc.SPECIES_BF16}[wl["storage"]])
        o.reset()
    em = seeded_emitters(wl["dims"], wl["emitters"], radius_range=wl["radius"], seed=wl["seed"])
    o.set_emitters(em, wl["jitter"], wl["velocity_noise"], wl["seed"])
    o.step(DT, 2)  # warm-up (page-in)
    times = []
    for _ in range(sample_steps):
        t0 = time.perf_counter()
        o.step(DT, 1)
        times.append(time.perf_counter() - t0)
    orc.set_threads(1)
    t0 = time.perf_counter()
    o.step(DT, 1)
    single = time.perf_counter() - t0
    orc.set_threads(cores)
    cells = wl["dims"][0] * wl["dims"][1] * wl["dims"][2]
    o.close()
    med = statistics.median(times)
    return {"value": cells / med, "unit": "cell-updates/s", "cores": cores, "kind": "port",
            "sample": f"median of {sample_steps} steps of the same workload after 2 warm-up steps, FP32 CPU oracle "
                      f"(OpenMP over z, {cores} threads; {med:.2f} s per step)",
            "single_thread": {"value": cells / single, "unit": "cell-updates/s", "cores": 1,
                              "sample": f"1 step after the sample above, 1 thread, {single:.1f} s"}}


SERVE_GRIDS = [(64, 96, 64), (128, 192, 128), (160, 240, 160), (192, 288, 192)]
SMALL_SERVE_GRIDS = [(64, 96, 64), (96, 64, 80), (64, 128, 64)]


def serving_stream(requests: int, rate: float, seed: int, streams: int = 4, batch: bool = False,
                   batch_streams: int = 0, grids=None, inflight: int = 1, max_active: int = 16,
                   graphs: bool = True) -> dict:
    """Mixed-size plumes arriving as a seeded Poisson stream, served with
    continuous batching (admission by HBM budget and slots, captured steps,
    one step in flight per plume).  Wall-clock arrivals: a request is
    submitted when its arrival time has passed.  batch: plumes of at most
    4 Mi cells go to batch_streams streams and those on one stream step
    together, each Jacobi launch one launch over all of them (16-wave ring
    plan of depth 4 on 8-plane z chunks); larger plumes step alone on the
    other streams."""
    from fire_amd import FireXHipSolver, seeded_emitters
    from fire_amd.serve import FireXServer, make_request, poisson_stream

    grids = grids or SERVE_GRIDS
    arrivals = poisson_stream(requests, rate, seed, grids=grids)
    # Autotune every grid shape once up front (the plan cache is per
    # process), so the timed stream does not include tuning.
    for g in sorted({a[1] for a in arrivals}):
        small = batch and g[0] * g[1] * g[2] <= 4 << 20
        FireXHipSolver(g, tier="uhd", pressure_iterations=96, correction_iterations=96, species_storage="f16",
                       plan="fused", heavy_mode="off", jacobi_autotune=True,
                       **({"jacobi_sweeps_per_launch": 4, "jacobi_ring_waves": 16, "jacobi_chunk_z": 8}
                          if small else {})).close()
    srv = FireXServer(max_active=max_active, streams=streams, plan="fused", capture_graphs=graphs,
                      max_inflight_steps=inflight,
                      jacobi_autotune=True, batch_plumes=batch, batch_streams=batch_streams)
    ids = []
    t0 = time.perf_counter()
    pending = list(arrivals)
    busy = True
    while pending or busy:
        now = time.perf_counter() - t0
        while pending and pending[0][0] <= now:
            _, dims, steps, rseed = pending.pop(0)
            em = seeded_emitters(dims, 3, radius_range=(6.0, 14.0), seed=rseed)
            ids.append((srv.submit(make_request(dims, steps, emitters=em, jitter=0.1, seed=rseed)), dims, steps))
        busy = srv.pump()
        if not busy and pending:
            time.sleep(max(0.0, min(0.002, pending[0][0] - (time.perf_counter() - t0))))
    srv.drain()
    makespan = time.perf_counter() - t0
    stats = srv.stats()
    done = [srv.status(i) for i, _, _ in ids]
    srv.close()
    turnaround = sorted(d["finish_time"] - d["submit_time"] for d in done)
    waits = sorted(d["admit_time"] - d["submit_time"] for d in done)

    def pct(v, q):
        return float(v[min(len(v) - 1, int(q * len(v)))]) if v else 0.0

    return {
        "requests": len(ids), "arrival_rate_per_s": rate, "seed": seed, "grids": [list(g) for g in grids],
        "steps_range": [60, 600], "jacobi": [96, 96], "species_storage": "f16",
        "completed": int(stats["done"]), "failed": int(stats["failed"]),
        "makespan_s": round(makespan, 3), "steps_completed": int(stats["steps_completed"]),
        "cell_updates_per_s": stats["cell_updates"] / makespan,
        "step_latency_ms": {"p50": round(stats["latency_ms_p50"], 3), "p99": round(stats["latency_ms_p99"], 3),
                            "max": round(stats["latency_ms_max"], 3)},
        "request_turnaround_s": {"p50": round(pct(turnaround, 0.5), 3), "p99": round(pct(turnaround, 0.99), 3)},
        "admission_wait_s": {"p50": round(pct(waits, 0.5), 3), "p99": round(pct(waits, 0.99), 3)},
        "server": {"max_active": max_active, "streams": streams, "capture_graphs": graphs, "max_inflight_steps": inflight,
                   "batch_plumes": batch, "batch_streams": batch_streams if batch else 0,
                   "memory_budget_bytes": int(stats["memory_budget_bytes"])},
        "group_steps": int(stats["group_steps"]), "grouped_plume_steps": int(stats["grouped_steps"]),
    }


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=600)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--workload", default="uhd-mixed", choices=sorted(WORKLOADS))
    ap.add_argument("--plan", default="fused", choices=["fused", "reference"])
    ap.add_argument("--heavy", default=None, choices=["auto", "off", "f16", "f32"],
                    help="heavy path (>= 64 Mi cells): default auto for plume-508, off otherwise")
    ap.add_argument("--sweeps", type=int, default=0, help="Jacobi sweeps per launch (0 = autotuned)")
    ap.add_argument("--jacobi-chunk", type=int, default=0, help="Jacobi z planes per block (0 = autotuned)")
    ap.add_argument("--jacobi-waves", type=int, default=0,
                    help="waves per ring-kernel block: 8 (4 rows each) or 16 (2 rows each); 0 = autotuned or default")
    ap.add_argument("--no-autotune", action="store_true",
                    help="no Jacobi plan search (fixed --sweeps / --jacobi-chunk or the defaults): for profiles")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="only run a plain device-to-device copy of BYTES three times (counter calibration)")
    ap.add_argument("--bandwidth-probe", action="store_true",
                    help="only time plain device reads (sum) and copies from 64 MiB to 4 GiB: the achievable rates "
                         "with and without the 256 MB last-level cache")
    ap.add_argument("--serve-only", choices=["mixed", "small"], default=None,
                    help="only run one serving stream (the mixed grids or the small grids), per plume and batched")
    ap.add_argument("--no-serve-graphs", dest="serve_graphs", action="store_false",
                    help="serve without captured step graphs (every launch eager)")
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--plain-steps", type=int, default=60, help="steps timed on the plain f32 path (0 = skip)")
    ap.add_argument("--cpu-sample-steps", type=int, default=5)
    ap.add_argument("--large-steps", type=int, default=10,
                    help="steps of the 508^3 plume for its per-pass HBM GB/s (0 = skip)")
    ap.add_argument("--droplet-steps", type=int, default=60,
                    help="steps of the same workload with the hose on (SPH droplets; 0 = skip)")
    ap.add_argument("--no-render", dest="render", action="store_false",
                    help="skip the 4K volume-render measurements")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=16, help="serving-stream sample size (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=20.0, help="Poisson arrivals per second")
    args = ap.parse_args()

    if args.copy_calibration > 0:
        # A copy kernel of known size: FETCH_SIZE / WRITE_SIZE of this run
        # scale the same counters of the solver's kernels to bytes.
        import torch

        n = args.copy_calibration // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        dst.copy_(src)
        a.record()
        for _ in range(3):
            dst.copy_(src)
        b.record()
        torch.cuda.synchronize()
        ms = a.elapsed_time(b) / 3
        print(json.dumps({"copy_calibration_bytes": n * 4, "copies": 4, "copy_ms": ms,
                          "copy_GBps_read_plus_write": 2 * n * 4 / (ms * 1e-3) / 1e9}), flush=True)
        return 0

    if args.bandwidth_probe:
        # What a plain stream reaches on this device: a sum reads the buffer
        # once, a copy reads and writes it; 20 back-to-back launches each.
        import torch

        a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

        def launch_us(fn, reps=20):
            fn()
            torch.cuda.synchronize()
            a.record()
            for _ in range(reps):
                fn()
            b.record()
            torch.cuda.synchronize()
            return a.elapsed_time(b) / reps * 1e3

        rows = []
        for mib in (64, 128, 256, 1024, 4096):
            n = mib * 2**20 // 4
            x = torch.rand(n, device="cuda")
            y = torch.empty_like(x)
            read_us, copy_us = launch_us(lambda: x.sum()), launch_us(lambda: y.copy_(x))
            rows.append({"MiB": mib, "read_TBps": round(n * 4 / read_us / 1e6, 2),
                         "copy_TBps_read_plus_write": round(2 * n * 4 / copy_us / 1e6, 2)})
            print(json.dumps(rows[-1]), flush=True)
            del x, y
        print(json.dumps({"bandwidth_probe": rows}), flush=True)
        return 0

    if args.serve_only:
        n, rate, kw = args.serve_requests, args.serve_rate, {}
        if args.serve_only == "small":
            n, rate, kw = 2 * n, 10 * rate, {"grids": SMALL_SERVE_GRIDS}
        out = {"per_plume": serving_stream(n, rate, seed=2026, graphs=args.serve_graphs, **kw),
               "batched": serving_stream(n, rate, seed=2026, streams=3 if kw else 4, batch=True,
                                         batch_streams=3 if kw else 2, graphs=args.serve_graphs, **kw)}
        print(json.dumps({"serving_stream": args.serve_only, **out}), flush=True)
        return 0

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: the Fire-X solver here is a single-GPU build "
              "(a plume would shard by z slabs across GPUs; that path is not built).", file=sys.stderr)
        return 2

    from fire_amd import device_count

    if device_count() < 1:
        print("bench.py needs a HIP device", file=sys.stderr)
        return 2
    wl = WORKLOADS[args.workload]
    cells = wl["dims"][0] * wl["dims"][1] * wl["dims"][2]
    print(f"bench: workload {args.workload} {wl['dims']} plan {args.plan} steps {args.steps} warmup {args.warmup}",
          file=sys.stderr, flush=True)

    t_setup = time.perf_counter()
    heavy = args.heavy or ("auto" if args.workload == "plume-508" else "off")
    solver = make_solver(wl, args.plan, args.sweeps, autotune=not args.no_autotune and args.plan == "fused",
                         chunk=args.jacobi_chunk, waves=args.jacobi_waves, heavy=heavy)
    jacobi_plan = solver.info["jacobi"]
    f16_pressure = bool(solver.info["f16_pressure"])
    print(f"bench: solver ready in {time.perf_counter() - t_setup:.2f} s, Jacobi plan {jacobi_plan}",
          file=sys.stderr, flush=True)
    elapsed = time_steps(solver, args.steps, args.warmup)
    ms_per_step = elapsed / args.steps * 1e3
    value = cells * args.steps / elapsed
    print(f"bench: timed {args.steps} steps in {elapsed:.3f} s ({ms_per_step:.3f} ms/step)", file=sys.stderr, flush=True)
    passes = pass_profile(solver, args.profile_steps, cells, wl["storage"], f16_pressure=f16_pressure)
    passes["pack_fields"] = pack_profile(solver, cells, wl["storage"])
    audit = audit_profile(solver)
    render = render_profile(solver) if args.render else None
    t_fuel = solver.reduce("sum_fuel")
    solver.close()

    jac = passes.get("jacobi", {})
    raw = passes.pop("_raw")
    sweeps_per_launch = raw["jacobi_sweeps"] / max(1, raw["jacobi"]["launches"])
    avg_launch_s = jac.get("avg_launch_us", 0.0) * 1e-6
    bytes_per_launch = (8 if f16_pressure else BYTES_PER_CELL["jacobi_launch"]) * cells  # f16 p: 2 + 4 + 2 B
    achieved = bytes_per_launch / avg_launch_s / 1e9 if avg_launch_s > 0 else 0.0
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": measured_traffic(args.workload, jacobi_plan),
                "kernel": f"k_jacobi_{jacobi_plan['kernel']} (Jacobi, {sweeps_per_launch:g} sweeps per launch, "
                          f"z chunk {jacobi_plan['chunk_z']})",
                "algorithmic_bytes_per_launch": int(bytes_per_launch),
                "avg_launch_us": round(jac.get("avg_launch_us", 0.0), 2)}

    plain = None
    if args.plain_steps > 0:
        print("bench: plain f32 reference-structured path", file=sys.stderr, flush=True)
        pw = dict(wl, storage="f32")
        ps = make_solver(pw, "reference")
        pel = time_steps(ps, args.plain_steps, 5)
        ps.close()
        plain = {"cell_updates_per_s": cells * args.plain_steps / pel, "ms_per_step": pel / args.plain_steps * 1e3,
                 "steps": args.plain_steps}

    large = None
    if args.large_steps > 0 and args.workload == "uhd-mixed":
        lw = WORKLOADS["plume-508"]
        lcells = lw["dims"][0] * lw["dims"][1] * lw["dims"][2]
        print(f"bench: 508^3 plume ({lcells} cells), {args.large_steps} steps", file=sys.stderr, flush=True)
        ls = make_solver(lw, "fused", autotune=True, heavy="auto")
        linfo = ls.info
        lel = time_steps(ls, args.large_steps, 2)
        lpasses = pass_profile(ls, 2, lcells, lw["storage"], f16_pressure=linfo["f16_pressure"])
        lpasses.pop("_raw")
        frame = None
        if args.render:
            # the reference's published 508^3 figure: one step + a 4K presentation per output frame
            w, h, rsteps = RENDER_4K
            host_frame = ls.render_frame(w, h, rsteps)
            ls.sync()
            t0 = time.perf_counter()
            for _ in range(3):
                ls.step(DT, 1)
                ls.render_frame(w, h, rsteps, read=False)
            ls.sync()
            fs = 3 / (time.perf_counter() - t0)
            # the same with each frame copied back to host memory (a recorder)
            t0 = time.perf_counter()
            for _ in range(3):
                ls.step(DT, 1)
                ls.render_frame(w, h, rsteps, read=False)
                ls.read_frame(host_frame)
            fr = 3 / (time.perf_counter() - t0)
            frame = {"frames_per_s": fs, "ms_per_frame": 1e3 / fs, "output": f"{w}x{h}", "ray_steps": rsteps,
                     "frames_per_s_with_readback": fr, "ms_per_frame_with_readback": 1e3 / fr,
                     "reference_published_frames_per_s": REFERENCE_508_FRAMES_PER_S,
                     "over_reference_published": fs / REFERENCE_508_FRAMES_PER_S,
                     "note": "presentation here = packFields, the volume ray march, the quarter-size bloom and the "
                             "filmic composite into an rgba8 frame; the reference's also draws droplet sprites "
                             "(none in this dry plume); its hardware is unspecified"}
        ls.close()
        # the configuration as stated: FP32 pressure (the compact heavy path,
        # FireXEngine.ts:2214-2246: shared pressure fields, cold start)
        print("bench: 508^3 plume with FP32 pressure", file=sys.stderr, flush=True)
        fs32 = make_solver(lw, "fused", autotune=True, heavy="f32")
        finfo = fs32.info
        fel = time_steps(fs32, args.large_steps, 2)
        fpasses = pass_profile(fs32, 2, lcells, lw["storage"], f16_pressure=False)
        fpasses.pop("_raw")
        fs32.close()
        fp32_pressure = {"pressure": "f32 (compact heavy path)", "ms_per_step": fel / args.large_steps * 1e3,
                         "cell_updates_per_s": lcells * args.large_steps / fel, "jacobi_plan": finfo["jacobi"],
                         "passes": {k: dict(v, frac_of_peak=round(v["algorithmic_GBps"] / PEAK_HBM_GBS, 4))
                                    if isinstance(v, dict) else v for k, v in fpasses.items()}}
        large = {"grid": list(lw["dims"]), "cells": lcells, "steps": args.large_steps, "jacobi": list(lw["iters"]),
                 "species_storage": lw["storage"], "pressure": "f16" if linfo["f16_pressure"] else "f32",
                 "velocity_noise": lw["velocity_noise"], "ms_per_step": lel / args.large_steps * 1e3,
                 "cell_updates_per_s": lcells * args.large_steps / lel, "jacobi_plan": linfo["jacobi"],
                 "passes": {k: dict(v, frac_of_peak=round(v["algorithmic_GBps"] / PEAK_HBM_GBS, 4))
                            if isinstance(v, dict) else v for k, v in lpasses.items()},
                 "step_plus_4k_frame": frame, "fp32_pressure": fp32_pressure}

    droplets = None
    if args.droplet_steps > 0:
        print(f"bench: hose on, {args.droplet_steps} steps", file=sys.stderr, flush=True)
        droplets = droplet_profile(wl, args.droplet_steps, 10, jacobi_plan, render=args.render)
        droplets["step_time_over_dry"] = round(droplets["ms_per_step_with_water"] / ms_per_step, 4)

    serving = None
    if args.serve_requests > 0:
        print(f"bench: serving stream ({args.serve_requests} requests at {args.serve_rate}/s)", file=sys.stderr,
              flush=True)
        serving = serving_stream(args.serve_requests, args.serve_rate, seed=2026)
        # Small plumes batched: grids up to 4 Mi cells share two of the four
        # streams and step in groups (one Jacobi launch per group), larger
        # ones step alone on the other two.
        print("bench: serving stream, small plumes batched", file=sys.stderr, flush=True)
        serving["batched"] = serving_stream(args.serve_requests, args.serve_rate, seed=2026, streams=4, batch=True,
                                            batch_streams=2)
        # A stream of small grids only (where one plume's launches fill a
        # fraction of the device): per-plume streams against batched groups.
        small = {"grids": SMALL_SERVE_GRIDS, "requests": 2 * args.serve_requests, "rate": 10 * args.serve_rate}
        print("bench: small-grid serving stream, per plume and batched", file=sys.stderr, flush=True)
        serving["small_grids"] = {
            "per_plume": serving_stream(small["requests"], small["rate"], seed=2026, grids=SMALL_SERVE_GRIDS),
            "batched": serving_stream(small["requests"], small["rate"], seed=2026, streams=3, batch=True,
                                      batch_streams=3, grids=SMALL_SERVE_GRIDS),
        }

    cpu = None
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (FP32 oracle)", file=sys.stderr, flush=True)
        cpu = cpu_baseline(wl, args.cpu_sample_steps)

    line = {
        "metric": "cell-updates/s at 192x288x192 (96+96 Jacobi, f16 species)",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": args.workload, "grid": list(wl["dims"]), "cells": cells,
                   "jacobi": list(wl["iters"]), "species_storage": wl["storage"], "dt": DT,
                   "emitters": wl["emitters"], "seed": wl["seed"], "plan": args.plan, "jacobi_plan": jacobi_plan,
                   "pressure_storage": "f16" if f16_pressure else "f32", "parallelism": "single GPU"},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "passes": passes,
        "plain_f32_gpu": plain,
        "multiple_over_plain_f32": (value / plain["cell_updates_per_s"]) if plain else None,
        "serving": serving,
        "plume_508": large,
        "droplets": droplets,
        "capture_audit": audit,
        "render_4k": render,
        "check": {"sum_fuel_after_run": t_fuel},
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
