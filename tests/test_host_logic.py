"""Known-answer tests of the host-side logic, from the reference's own tests.

Values come from src/sim/firex/gridConfiguration.test.ts,
src/sim/qualityProfiles.test.ts, src/sim/offlineTimeline.test.ts,
src/sim/FixedStepClock.test.ts and scripts/record-fire-lib.test.mjs; every
call goes to libfirex_hip.so (no GPU needed).
"""
import math

import pytest

from fire_amd import solver as fs

GENEROUS = {"maxTextureDimension3D": 2048, "maxStorageBufferBindingSize": 128 * 1024 * 1024,
            "maxBufferSize": 256 * 1024 * 1024}  # gridConfiguration.test.ts:22-26
DETAIL = 64 * 64 * 64 * 4
GIB = 1024 ** 3


# ---------------------------------------------------------------- memory model
def test_heavy_threshold_switches_bytes_per_cell():
    # gridConfiguration.test.ts:119-127, record-fire-lib.test.mjs:62-68
    below = fs.estimate_memory((1020, 256, 256))
    at = fs.estimate_memory((1024, 256, 256))
    assert below.cell_count < 64 * 1024 * 1024
    assert below.grid_bytes == below.cell_count * 116
    assert at.cell_count == 64 * 1024 * 1024
    assert at.grid_bytes == at.cell_count * 108


def test_estimate_accounts_for_every_allocation():
    # gridConfiguration.test.ts:129-139
    e = fs.estimate_memory((80, 120, 80), 1536)
    assert e.cell_count == 768_000
    assert e.grid_bytes == 768_000 * 116
    assert e.particle_bytes == 1536 * 80
    assert e.fixed_bytes == DETAIL
    assert e.total_bytes == 768_000 * 116 + 1536 * 80 + DETAIL
    assert e.largest_storage_buffer_bytes == 768_000 * 16


def test_reference_508_cubed():
    # record-fire-lib.test.mjs:56-60
    e = fs.estimate_memory((508, 508, 508), 4096)
    assert e.cell_count == 131_096_512
    assert round(e.total_bytes / GIB, 2) == pytest.approx(13.19, abs=0.006)
    assert e.total_bytes / GIB == pytest.approx(13.187, abs=0.005)


@pytest.mark.parametrize("axis,cells,total,largest,budget", [
    (256, 16_777_216, 1_947_328_512, 268_435_456, 2 * GIB),
    (320, 32_768_000, 3_802_259_456, 524_288_000, 4 * GIB),
    (384, 56_623_104, 6_569_451_520, 905_969_664, 8 * GIB),
    (512, 134_217_728, 14_496_686_080, 2_147_483_648, 16 * GIB),
])
def test_dense_cubes_pass_with_exact_limits(axis, cells, total, largest, budget):
    # gridConfiguration.test.ts:162-188
    e = fs.estimate_memory((axis,) * 3, 1536)
    assert (e.cell_count, e.total_bytes, e.largest_storage_buffer_bytes) == (cells, total, largest)
    limits = {"maxTextureDimension3D": 2048, "maxStorageBufferBindingSize": largest, "maxBufferSize": largest}
    assert fs.preflight((axis,) * 3, limits, 1536, budget) == []


def test_preflight_accepts_hd_experimental_preset():
    # gridConfiguration.test.ts:141-150
    assert fs.preflight((128, 192, 128), GENEROUS, 3072) == []
    assert fs.estimate_memory((128, 192, 128), 3072).total_bytes < 512 * 1024 * 1024


def test_preflight_192_cubed_needs_expert_budget():
    # gridConfiguration.test.ts:152-160
    assert fs.preflight((192, 192, 192), GENEROUS, 1536, 1024 * 1024 * 1024) == []
    assert fs.estimate_memory((192, 192, 192), 1536).total_bytes > 512 * 1024 * 1024


def test_preflight_reports_every_failure_kind():
    # gridConfiguration.test.ts:190-208
    issues = fs.preflight((162, 240, 160), {"maxTextureDimension3D": 192, "maxStorageBufferBindingSize": 32 * 1024 * 1024,
                                            "maxBufferSize": 48 * 1024 * 1024}, 4096, 256 * 1024 * 1024)
    for code in ("unaligned-dimension", "texture-dimension-limit", "storage-binding-limit", "buffer-size-limit",
                 "memory-budget"):
        assert code in issues


def test_preflight_budget_guard_can_be_disabled():
    # gridConfiguration.test.ts:210-221
    limits = {"maxTextureDimension3D": 2048, "maxStorageBufferBindingSize": 128 * 1024 * 1024,
              "maxBufferSize": 256 * 1024 * 1024}
    assert fs.preflight((160, 240, 160), limits, 0, None) == []


def test_preflight_1024_cubed_is_not_normalised():
    # gridConfiguration.test.ts:223-252
    e = fs.estimate_memory((1024, 1024, 1024), 1536)
    assert e.cell_count == 1_073_741_824
    assert e.grid_bytes == 1_073_741_824 * 108 == 108 * GIB
    assert e.largest_storage_buffer_bytes == 16 * GIB
    issues = fs.preflight((1024, 1024, 1024), GENEROUS, 1536)
    assert {"storage-binding-limit", "buffer-size-limit", "memory-budget"} <= set(issues)
    assert "texture-dimension-limit" not in issues


def test_preflight_unaligned_axis():
    # gridConfiguration.test.ts:254-259
    assert "unaligned-dimension" in fs.preflight((1023, 1024, 1024), GENEROUS)


# ----------------------------------------------------------------------- tiers
@pytest.mark.parametrize("tier,dims,iters,dt", [
    ("maximum", (80, 120, 80), 48, 1 / 30),   # qualityProfiles.test.ts:29-36
    ("cinematic", (80, 120, 80), 48, 1 / 30),  # qualityProfiles.test.ts:49-56
    ("hd", (128, 192, 128), 64, 1 / 60),       # qualityProfiles.test.ts:105-113
    ("qhd", (160, 240, 160), 80, 1 / 60),      # qualityProfiles.test.ts:114-122
    ("uhd", (192, 288, 192), 96, 1 / 60),      # qualityProfiles.test.ts:123-131
])
def test_tier_tables(tier, dims, iters, dt):
    cfg = fs.config_from_tier(tier)
    assert tuple(cfg.dims) == dims
    assert cfg.pressure_iterations == iters and cfg.correction_iterations == iters
    assert cfg.time_step == dt
    assert cfg.maximum_substeps == 1


def test_tier_ordering_and_cinematic_flag():
    high, maximum = fs.config_from_tier("high"), fs.config_from_tier("maximum")
    assert math.prod(maximum.dims) > math.prod(high.dims)
    assert maximum.pressure_iterations > high.pressure_iterations
    assert maximum.correction_iterations > high.correction_iterations
    # usesCinematicPresentation: cinematic quality and every offline tier
    assert fs.config_from_tier("cinematic").cinematic == 1
    assert fs.config_from_tier("uhd").cinematic == 1
    assert fs.config_from_tier("balanced").cinematic == 0
    live = math.prod(fs.config_from_tier("cinematic").dims)
    for t in ("hd", "qhd", "uhd"):
        assert math.prod(fs.config_from_tier(t).dims) > live
    with pytest.raises(fs.FireXError):
        fs.config_from_tier("ultra-mega")


def test_offline_dimension_override():
    # qualityProfiles.test.ts:140-141
    assert fs.resolve_offline_dims("hd", (160, 128, 192)) == (160, 192, 192)
    assert fs.resolve_offline_dims("uhd") == (192, 288, 192)


# -------------------------------------------------------------------- timeline
def test_offline_split_one_step_per_60fps_frame():
    # offlineTimeline.test.ts:5-15
    samples = [fs.split_offline_delta(1 / 60, 1 / 30) for _ in range(60)]
    assert all(len(s) == 1 and s[0] > 0 for s in samples)
    assert sum(sum(s) for s in samples) == pytest.approx(1, abs=1e-12)


def test_offline_split_slow_intervals():
    # offlineTimeline.test.ts:17-22
    steps = fs.split_offline_delta(1 / 15, 1 / 30)
    assert len(steps) == 2
    assert max(steps) <= 1 / 30
    assert sum(steps) == pytest.approx(1 / 15, abs=1e-12)


def test_offline_split_zero_and_invalid():
    # offlineTimeline.test.ts:24-28
    assert fs.split_offline_delta(0, 1 / 30) == []
    with pytest.raises(ValueError):
        fs.split_offline_delta(-1, 1 / 30)
    with pytest.raises(ValueError):
        fs.split_offline_delta(1 / 60, 0)


def test_clock_deterministic_steps():
    # FixedStepClock.test.ts:5-14
    clock = fs.FixedStepClock(1 / 120, 8)
    elapsed = [0.0]

    def step(dt):
        elapsed[0] += dt

    clock.advance(1 / 60, step)
    clock.advance(1 / 60, step)
    assert elapsed[0] == pytest.approx(1 / 30, abs=1e-8)


def test_clock_bounds_catch_up():
    # FixedStepClock.test.ts:16-21
    clock = fs.FixedStepClock(1 / 120, 4)
    count = [0]
    assert clock.advance(10, lambda dt: count.__setitem__(0, count[0] + 1)) == 4
    assert count[0] == 4


def test_clock_rejects_invalid_and_ignores_nonfinite():
    # FixedStepClock.test.ts:23-28
    with pytest.raises(ValueError):
        fs.FixedStepClock(float("nan"), 4)
    with pytest.raises(ValueError):
        fs.FixedStepClock(1 / 120, 1.5)
    clock = fs.FixedStepClock(1 / 120, 4)
    assert clock.advance(float("inf")) == 0


# ------------------------------------------------------------- parameter clamps
@pytest.mark.parametrize("name,value,previous,expected", [
    ("burnerSize", 10, 3.1, 3.2), ("burnerSize", 0.1, 3.1, 0.35), ("burnerDepth", -5, 2.7, 0.35),
    ("sourceThickness", 9, 1.35, 2.5), ("firePower", 0, 2.5, 0.1), ("firePower", 5, 2.5, 4),
    ("sourceLift", -1, 1.6, 0), ("buoyancyScale", 7, 2.15, 4), ("flamePersistence", 0.1, 1.45, 0.4),
    ("airEntrainment", 3.5, 1.65, 3), ("aimHeight", 1, 0.42, 0.82), ("domainWidth", 0.1, 1.15, 0.5),
    ("domainHeight", 7, 1.35, 6), ("domainDepth", 5, 1.15, 4), ("viewMode", 3.6, 0, 4),
    ("nozzleType", 0.5, 0, 1), ("nozzleType", 0.49, 1, 0), ("opticalDetailTarget", 500, 1024, 512),
    ("opticalDetailTarget", 100, 1024, 0), ("fuelRate", 0.92, 0.3, 0.92), ("oxygenRate", -2, 1, -2),
    ("vorticity", float("nan"), 3.6, 3.6), ("firePower", float("inf"), 2.5, 2.5),
])
def test_set_parameter_clamps(name, value, previous, expected):
    # FireXEngine.ts:937-1032 (finiteNumber keeps the previous value)
    assert fs.clamp_parameter(name, value, previous) == pytest.approx(expected)


def test_unknown_parameter_key():
    with pytest.raises(fs.FireXError):
        fs.clamp_parameter("notAKey", 1, 1)
