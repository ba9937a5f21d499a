"""The benchmark configurations at their full sizes and late in their runs.

Late-state shadows: the GPU runs a configuration for its full step count,
its state (flow, species in its storage, warm-start pressure, clock and frame
index) is written into the FP32 oracle, and both take one more step, which
must agree bit for bit.  This checks the kernels where the trajectory is
strongly burning and turbulent (the ignition and normalize thresholds are
crossed everywhere in the plume), which a short run from rest does not reach:
  * 128 x 192 x 128, f32, 96 + 96 sweeps, after 300 steps (plain FP32);
  * 192 x 288 x 192, f16 species, 96 + 96, three emitters, after 600 steps;
  * 508^3, bf16 species, FP32 pressure on the compact heavy path
    (FireXEngine.ts:2214-2246: shared pressure fields, cold start), turbulent
    emitter, 96 + 96, after 120 steps.
Serving at its workload: plumes on the serving stream's four grids at 96 + 96
sweeps give the same fields in the mix as alone, and a served plume equals
the FP32 oracle.
"""
import numpy as np
import pytest

from fire_amd import FireXHipSolver, device_count, seeded_emitters
from fire_amd import serve as sv
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
STATE = (("flow", orc.FIELD_FLOW), ("species", orc.FIELD_SPECIES), ("pressure", orc.FIELD_PRESSURE))
COMPARED = STATE + (("correction", orc.FIELD_CORRECTION), ("divergence", orc.FIELD_DIVERGENCE),
                    ("reaction", orc.FIELD_REACTION))
STORAGE = {"f32": None, "f16": orc.SPECIES_F16, "bf16": orc.SPECIES_BF16}
HEAVY = {"off": orc.HEAVY_OFF, "f32": orc.HEAVY_F32, "f16": orc.HEAVY_F16}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if device_count() < 1:
        pytest.fail("no HIP device visible to the configuration tests")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def late_state_shadow(dims, steps, storage, heavy, emitters, jitter, noise, seed, shadow_steps=1):
    print(f"late state: {dims} {steps} steps, species {storage}, heavy {heavy}", flush=True)
    g = FireXHipSolver(dims, tier="uhd", pressure_iterations=96, correction_iterations=96, species_storage=storage,
                       heavy_mode=heavy, plan="fused")
    g.set_emitters(emitters, jitter=jitter, velocity_noise=noise, seed=seed)
    g.step(DT, steps)
    info = g.info
    assert info["frame_index"] == steps
    o = orc.Oracle(dims, precision=32, pressure_iterations=96, correction_iterations=96)
    o.set_bool("cinematic", True)
    o.set_heavy_mode(HEAVY[heavy])
    if STORAGE[storage] is not None:
        o.set_species_storage(STORAGE[storage])
    o.reset()
    o.set_emitters(emitters, jitter, noise, seed)
    for name, fid in STATE:
        o.write(fid, g.read_field(name))
    o.set_time(info["simulation_time"], steps)
    t = g.read_field("flow")[..., 3]
    hot = float((t > 930.0).mean())
    print(f"  state copied: T max {t.max():.0f} K, {hot:.4f} of cells above 930 K; stepping the oracle", flush=True)
    assert hot > 0  # burning
    del t
    g.step(DT, shadow_steps)
    o.step(DT, shadow_steps)
    for name, fid in COMPARED:
        a = bits(g.read_field(name))
        b = bits(o.read(fid).astype(np.float32))
        assert np.array_equal(a, b), f"{dims} step {steps + shadow_steps}: {name} differs in {(a != b).sum()} values"
        del a, b
    print("  bit-exact", flush=True)
    o.close()
    g.close()
    return info


def test_late_state_hd_plain_after_300_steps():
    dims = (128, 192, 128)
    em = seeded_emitters(dims, 1, radius_range=(8 * 2 ** 0.5,) * 2, seed=2026)
    late_state_shadow(dims, 300, "f32", "off", em, 0.1, 0.0, 2026, shadow_steps=2)


def test_late_state_uhd_mixed_after_600_steps():
    dims = (192, 288, 192)
    em = seeded_emitters(dims, 3, radius_range=(6.0, 14.0), seed=2026)
    late_state_shadow(dims, 600, "f16", "off", em, 0.1, 0.0, 2026)


@pytest.mark.timeout(1200)
def test_late_state_508_fp32_pressure_after_120_steps():
    """508^3 as the configuration states it: FP32 pressure (the compact heavy
    path, no f16 storage), bf16 scalars, seeded turbulent emitter."""
    dims = (508, 508, 508)
    em = seeded_emitters(dims, 1, radius_range=(32 * 2 ** 0.5,) * 2, seed=2026)
    info = late_state_shadow(dims, 120, "bf16", "f32", em, 0.1, 0.5, 2026)
    assert info["heavy"] and not info["f16_pressure"]


SERVE_MIX = [((64, 96, 64), 8), ((128, 192, 128), 5), ((160, 240, 160), 4), ((192, 288, 192), 3),
             ((64, 96, 64), 6)]


def test_serving_at_workload_grids_alone_equals_mix_and_oracle():
    """Five requests on the serving stream's grids (64 x 96 x 64 up to
    192 x 288 x 192) at 96 + 96 sweeps, f16 species, seeded emitters: every
    served plume equals the same plume alone, and the first equals the FP32
    oracle."""
    with sv.FireXServer(max_active=4, streams=4, plan="fused", capture_graphs=True) as srv:
        ids = []
        for k, (dims, steps) in enumerate(SERVE_MIX):
            em = seeded_emitters(dims, 3, radius_range=(6.0, 14.0), seed=300 + k)
            ids.append((srv.submit(sv.make_request(dims, steps, emitters=em, jitter=0.1, seed=300 + k,
                                                   keep_result=True)), dims, steps, em, 300 + k))
            srv.pump()
        srv.drain()
        assert srv.stats()["done"] == len(SERVE_MIX)
        for i, (rid, dims, steps, em, seed) in enumerate(ids):
            print(f"served plume {dims} {steps} steps", flush=True)
            with FireXHipSolver(dims, tier="uhd", pressure_iterations=96, correction_iterations=96,
                                species_storage="f16", plan="fused", heavy_mode="off") as g:
                g.set_emitters(em, jitter=0.1, seed=seed)
                g.step(DT, steps)
                for name in ("flow", "species", "pressure", "reaction"):
                    assert np.array_equal(bits(srv.read_field(rid, name)), bits(g.read_field(name))), (dims, name)
            if i == 0:
                o = orc.Oracle(dims, precision=32, pressure_iterations=96, correction_iterations=96)
                o.set_bool("cinematic", True)
                o.set_heavy_mode(orc.HEAVY_OFF)
                o.set_species_storage(orc.SPECIES_F16)
                o.reset()
                o.set_emitters(em, 0.1, 0.0, seed)
                o.step(DT, steps)
                for name, fid in (("flow", orc.FIELD_FLOW), ("species", orc.FIELD_SPECIES),
                                  ("pressure", orc.FIELD_PRESSURE), ("reaction", orc.FIELD_REACTION)):
                    assert np.array_equal(bits(srv.read_field(rid, name)), bits(o.read(fid).astype(np.float32))), name
                o.close()
            srv.release(rid)
