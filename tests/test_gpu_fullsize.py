"""Full-size parity: the CPU-runnable 64x96x64 configuration run in full, the
plain FP32 128x192x128 configuration, the UHD mixed-precision workload
(192x288x192) and the 508^3 plume, stepped on the GPU and by the FP32 oracle
on the host, compared bit for bit (+0 and -0 differ).

The FP32 oracle restates the reference step (src/sim/firex/shaders.ts,
FireXEngine.ts:2115-2276); at these sizes the ring Jacobi kernel runs with
many tiles per row and column, interior and wall tiles, both z stream
directions, and (508^3) the >= 64 Mi cell path with f16 pressure.  One step
of the oracle at 508^3 takes seconds on the host's cores.
"""
import numpy as np
import pytest

from fire_amd import FireXHipSolver, device_count, seeded_emitters
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
FIELDS = [("flow", orc.FIELD_FLOW), ("species", orc.FIELD_SPECIES), ("pressure", orc.FIELD_PRESSURE),
          ("correction", orc.FIELD_CORRECTION), ("divergence", orc.FIELD_DIVERGENCE),
          ("reaction", orc.FIELD_REACTION)]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if device_count() < 1:
        pytest.fail("no HIP device visible to the full-size parity tests")


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def _compare(g, o, label, names):
    for name, fid in FIELDS:
        if name not in names:
            continue
        a = _bits(g.read_field(name))
        b = _bits(o.read(fid).astype(np.float32))
        bad = int(np.count_nonzero(a != b))
        assert bad == 0, f"{label}: field {name} differs in {bad} of {a.size} values"


def _run_pair(dims, storage, heavy, emitters, noise, steps, iters=96, plan="fused"):
    em = emitters
    o = orc.Oracle(dims, precision=32, pressure_iterations=iters, correction_iterations=iters)
    o.set_bool("cinematic", True)
    o.set_heavy_mode({"off": orc.HEAVY_OFF, "auto": orc.HEAVY_AUTO}[heavy])
    if storage != "f32":
        o.set_species_storage({"f16": orc.SPECIES_F16, "bf16": orc.SPECIES_BF16}[storage])
        o.reset()
    o.set_emitters(em, 0.1, noise, 2026)
    g = FireXHipSolver(dims, tier="uhd", pressure_iterations=iters, correction_iterations=iters, plan=plan,
                       species_storage=storage, heavy_mode=heavy, jacobi_autotune=plan == "fused")
    g.set_emitters(em, 0.1, noise, 2026)
    print(f"{dims}: GPU plan {g.info['jacobi']}, heavy {g.info['heavy']}, f16 pressure {g.info['f16_pressure']}",
          flush=True)
    g.step(DT, steps)
    o.step(DT, steps)
    return g, o


@pytest.mark.parametrize("plan", ["reference", "fused"])
def test_cpu_plain_config_full_run_bitwise(plan):
    """The CPU-runnable configuration in full: 64x96x64, f32 fields, one seeded
    Gaussian emitter (sigma 4 cells) with +-10 % jitter, 24 + 24 sweeps,
    dt 1/60, 60 steps."""
    dims = (64, 96, 64)
    em = seeded_emitters(dims, 1, radius_range=(4 * 2 ** 0.5, 4 * 2 ** 0.5), seed=2026)
    g, o = _run_pair(dims, "f32", "off", em, 0.0, 60, iters=24, plan=plan)
    _compare(g, o, f"64x96x64 {plan} after 60 steps",
             {"flow", "species", "pressure", "correction", "divergence", "reaction"})
    g.close()
    o.close()


@pytest.mark.parametrize("plan", ["reference", "fused"])
def test_hd_plain_config_bitwise(plan):
    """The plain FP32 GPU configuration: 128x192x128, f32 fields, one seeded
    Gaussian emitter (sigma 8 cells), 96 + 96 sweeps, three steps, on the
    reference-structured plan (one launch per reference pass) and the fused plan."""
    dims = (128, 192, 128)
    em = seeded_emitters(dims, 1, radius_range=(8 * 2 ** 0.5, 8 * 2 ** 0.5), seed=2026)
    g, o = _run_pair(dims, "f32", "off", em, 0.0, 3, iters=96, plan=plan)
    _compare(g, o, f"128x192x128 {plan} after 3 steps",
             {"flow", "species", "pressure", "correction", "divergence", "reaction"})
    g.close()
    o.close()


def test_uhd_mixed_workload_bitwise():
    """192x288x192, f16 species, three seeded emitters of radius 6-14, 96 + 96 sweeps, two steps."""
    dims = (192, 288, 192)
    em = seeded_emitters(dims, 3, radius_range=(6.0, 14.0), seed=2026)
    g, o = _run_pair(dims, "f16", "off", em, 0.0, 2)
    _compare(g, o, "UHD mixed", {"flow", "species", "pressure", "correction", "divergence", "reaction"})
    g.pack_fields()
    for name, ref in zip(("packed_flow", "packed_chemistry", "packed_media"), o.pack_fields()):
        bad = int(np.count_nonzero(_bits(g.read_field(name)) != _bits(ref.astype(np.float32))))
        assert bad == 0, f"UHD {name} differs in {bad} values"
    g.close()
    o.close()


def test_508_cubed_turbulent_plume_bitwise():
    """508^3 (131.1 M cells), bf16 species, a seeded emitter with +-0.5 velocity
    noise, 96 + 96 sweeps on the >= 64 Mi cell path (f16 pressure), one step."""
    dims = (508, 508, 508)
    em = seeded_emitters(dims, 1, radius_range=(24.0, 24.0), seed=2026)
    g, o = _run_pair(dims, "bf16", "auto", em, 0.5, 1)
    assert g.info["heavy"] and g.info["f16_pressure"]
    _compare(g, o, "508^3", {"flow", "species", "pressure", "correction", "divergence", "reaction"})
    g.close()
    o.close()
