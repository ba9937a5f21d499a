"""GPU parity: the HIP solver against the FP32 oracle, bit for bit.

The FP32 oracle restates the reference step (src/sim/firex/shaders.ts,
FireXEngine.ts:2115-2276) in WGSL operation order; the HIP kernels keep the
same f32 operations, so every field must match exactly (bitwise on the f32
pattern; +0 and -0 count as different).  Grids are small enough for the
oracle to finish in seconds.
"""
import numpy as np
import pytest

from fire_amd import FireXHipSolver, device_count, seeded_emitters
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
FIELD_PAIRS = [
    ("flow", orc.FIELD_FLOW),
    ("species", orc.FIELD_SPECIES),
    ("pressure", orc.FIELD_PRESSURE),
    ("correction", orc.FIELD_CORRECTION),
    ("divergence", orc.FIELD_DIVERGENCE),
    ("reaction", orc.FIELD_REACTION),
]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if device_count() < 1:
        pytest.fail("no HIP device visible to the GPU parity tests")


def _bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_fields_equal(gpu, oracle, label, names=None):
    for name, fid in FIELD_PAIRS:
        if names and name not in names:
            continue
        g = gpu.read_field(name)
        o = oracle.read(fid).astype(np.float32)
        if not np.array_equal(_bits(g), _bits(o)):
            diff = np.abs(g.astype(np.float64) - o.astype(np.float64))
            bad = np.argwhere(_bits(g) != _bits(o))
            raise AssertionError(
                f"{label}: field {name} differs in {len(bad)} of {g.size} values; max |diff| {diff.max():.3e}; "
                f"first at {bad[0].tolist()} gpu={g[tuple(bad[0])]!r} oracle={o[tuple(bad[0])]!r}")


def make_pair(dims, iters=(8, 8), plan="reference", storage="f32", heavy="off", sweeps=0, emitters=None,
              params=None):
    o = orc.Oracle(dims, precision=32, pressure_iterations=iters[0], correction_iterations=iters[1])
    o.set_bool("cinematic", True)
    if storage != "f32":
        o.set_species_storage({"f16": orc.SPECIES_F16, "bf16": orc.SPECIES_BF16}[storage])
        o.reset()
    o.set_heavy_mode({"off": orc.HEAVY_OFF, "f16": orc.HEAVY_F16, "f32": orc.HEAVY_F32, "auto": orc.HEAVY_AUTO}[heavy])
    g = FireXHipSolver(dims, tier="hd", pressure_iterations=iters[0], correction_iterations=iters[1], plan=plan,
                       species_storage=storage, heavy_mode=heavy, jacobi_sweeps_per_launch=sweeps)
    if emitters is not None:
        em, jitter, noise, seed = emitters
        o.set_emitters(em, jitter, noise, seed)
        g.set_emitters(em, jitter, noise, seed)
    for k, v in (params or {}).items():
        o.set_parameter(k, v)
        g.set_parameter(k, v)
    return g, o


def test_initial_state_matches():
    g, o = make_pair((16, 24, 16))
    assert_fields_equal(g, o, "after reset")
    g.close()


@pytest.mark.parametrize("dims", [(16, 24, 16), (20, 28, 12), (17, 29, 13)])
def test_reference_plan_steps_bitwise(dims):
    g, o = make_pair(dims, iters=(8, 8))
    for step in range(3):
        o.step(DT)
        g.step(DT)
        assert_fields_equal(g, o, f"reference plan {dims} step {step + 1}")
    g.close()


@pytest.mark.parametrize("sweeps", [1, 2, 3, 4, 5, 6, 7, 8])
def test_fused_plan_sweep_depths_bitwise(sweeps):
    dims = (40, 72, 36)
    g, o = make_pair(dims, iters=(11, 9), plan="fused", sweeps=sweeps)
    for step in range(2):
        o.step(DT)
        g.step(DT)
    assert_fields_equal(g, o, f"fused plan K={sweeps}", names=["flow", "species", "divergence", "reaction"])
    # pressure/correction live in whichever buffer the last launch wrote
    np.testing.assert_array_equal(_bits(g.read_field("pressure")), _bits(o.read(orc.FIELD_PRESSURE)))
    np.testing.assert_array_equal(_bits(g.read_field("correction")), _bits(o.read(orc.FIELD_CORRECTION)))
    g.close()


@pytest.mark.parametrize("plan", ["reference", "fused"])
def test_hd_shape_many_steps_bitwise(plan):
    dims = (64, 96, 64)
    g, o = make_pair(dims, iters=(24, 24), plan=plan)
    o.step(DT, 10)
    g.step(DT, 10)
    assert_fields_equal(g, o, f"{plan} 64x96x64 after 10 steps")
    g.close()


@pytest.mark.parametrize("storage", ["f16", "bf16"])
@pytest.mark.parametrize("plan", ["reference", "fused"])
def test_reduced_precision_species_bitwise(storage, plan):
    dims = (24, 36, 24)
    g, o = make_pair(dims, iters=(8, 8), plan=plan, storage=storage)
    o.step(DT, 4)
    g.step(DT, 4)
    assert_fields_equal(g, o, f"{storage} species {plan}")
    g.close()


@pytest.mark.parametrize("plan", ["reference", "fused"])
def test_seeded_emitters_bitwise(plan):
    dims = (32, 48, 32)
    em = seeded_emitters(dims, 3, radius_range=(3.0, 7.0), seed=7)
    g, o = make_pair(dims, iters=(8, 8), plan=plan, emitters=(em, 0.1, 0.5, 1234))
    o.step(DT, 4)
    g.step(DT, 4)
    assert_fields_equal(g, o, f"seeded emitters {plan}")
    g.close()


@pytest.mark.parametrize("heavy", ["f16", "f32"])
@pytest.mark.parametrize("plan", ["reference", "fused"])
def test_heavy_paths_forced_on_small_grid_bitwise(heavy, plan):
    dims = (24, 32, 20)
    g, o = make_pair(dims, iters=(8, 8), plan=plan, heavy=heavy)
    o.step(DT, 3)
    g.step(DT, 3)
    assert_fields_equal(g, o, f"heavy {heavy} {plan}", names=["flow", "species", "divergence", "reaction"])
    np.testing.assert_array_equal(_bits(g.read_field("pressure")), _bits(o.read(orc.FIELD_PRESSURE)))
    np.testing.assert_array_equal(_bits(g.read_field("correction")), _bits(o.read(orc.FIELD_CORRECTION)))
    g.close()


def test_scene_parameters_bitwise():
    dims = (24, 32, 24)
    params = {"burnerSize": 1.0, "burnerDepth": 1.0, "sourceThickness": 1.0, "firePower": 1.0, "sourceLift": 1.0,
              "fuelRate": 0.92, "oxygenRate": 0.42, "heatEfficiency": 0.62, "vorticity": 1.6, "buoyancyScale": 1.0,
              "flamePersistence": 1.0, "airEntrainment": 0.0, "domainWidth": 1.0, "domainHeight": 1.0,
              "domainDepth": 1.0}
    g, o = make_pair(dims, iters=(6, 6), params=params)
    # the oxygen level of the reset state follows oxygenRate
    g.reset()
    o.reset()
    o.step(DT, 3)
    g.step(DT, 3)
    assert_fields_equal(g, o, "rich scene")
    g.close()


def test_liquid_input_bitwise():
    dims = (16, 24, 16)
    g, o = make_pair(dims, iters=(6, 6))
    rng = np.random.default_rng(3)
    liquid = rng.integers(0, 4096 * 2, size=(dims[2], dims[1], dims[0])).astype(np.uint32)
    g.write_field("liquid", liquid)
    o.write(orc.FIELD_LIQUID, liquid.astype(np.float64))
    o.step(DT, 2)
    g.step(DT, 2)
    assert_fields_equal(g, o, "liquid cooling")
    g.close()


def test_seeded_state_write_and_step_bitwise():
    dims = (20, 24, 16)
    g, o = make_pair(dims, iters=(5, 7))
    rng = np.random.default_rng(11)
    flow = np.empty((dims[2], dims[1], dims[0], 4))
    flow[..., :3] = rng.uniform(-1.2, 1.2, size=flow[..., :3].shape)
    flow[..., 3] = rng.uniform(300, 2000, size=flow[..., 3].shape)
    flow = flow.astype(np.float32).astype(np.float64)
    species = rng.uniform(0, 1.5, size=(dims[2], dims[1], dims[0], 4)).astype(np.float32).astype(np.float64)
    pressure = rng.normal(0, 1e-3, size=(dims[2], dims[1], dims[0])).astype(np.float32).astype(np.float64)
    for name, fid, arr in (("flow", orc.FIELD_FLOW, flow), ("species", orc.FIELD_SPECIES, species),
                           ("pressure", orc.FIELD_PRESSURE, pressure)):
        o.write(fid, arr)
        g.write_field(name, arr)
    assert_fields_equal(g, o, "after write", names=["flow", "species", "pressure"])
    o.step(DT)
    g.step(DT)
    assert_fields_equal(g, o, "random state, one step")
    g.close()


@pytest.mark.parametrize("plan", ["reference", "fused"])
def test_captured_graph_replay_bitwise(plan):
    """The post-simulate part of the step replayed from captured hipGraphs
    (one graph per buffer-index state) gives the eager results, step after
    step, with eager host work and synchronisation between the replays."""
    dims = (24, 40, 20)
    o = orc.Oracle(dims, precision=32, pressure_iterations=7, correction_iterations=5)
    o.set_bool("cinematic", True)
    o.set_heavy_mode(orc.HEAVY_OFF)
    g = FireXHipSolver(dims, tier="hd", pressure_iterations=7, correction_iterations=5, plan=plan,
                       heavy_mode="off", capture_graphs=True, jacobi_sweeps_per_launch=3)
    for step in range(5):
        o.step(DT)
        g.step(DT)
        g.sync()
        assert_fields_equal(g, o, f"graph replay {plan} step {step + 1}")
    g.close()


def _ring_pair(dims, iters, sweeps, heavy="off", chunk=0, height=0, prefetch=None, waves=0):
    o = orc.Oracle(dims, precision=32, pressure_iterations=iters[0], correction_iterations=iters[1])
    o.set_bool("cinematic", True)
    o.set_heavy_mode({"off": orc.HEAVY_OFF, "f16": orc.HEAVY_F16}[heavy])
    g = FireXHipSolver(dims, tier="hd", pressure_iterations=iters[0], correction_iterations=iters[1], plan="fused",
                       heavy_mode=heavy, jacobi_sweeps_per_launch=sweeps, jacobi_kernel="ring",
                       jacobi_chunk_z=chunk, jacobi_tile_height=height, jacobi_prefetch=prefetch,
                       jacobi_ring_waves=waves)
    return g, o


@pytest.mark.parametrize("dims,sweeps,chunk,heavy", [
    ((80, 72, 30), 1, 7, "off"), ((80, 72, 30), 2, 5, "off"), ((80, 72, 30), 3, 0, "off"), ((80, 72, 30), 4, 7, "off"),
    ((64, 32, 16), 4, 5, "off"), ((128, 96, 40), 4, 9, "off"), ((70, 100, 11), 3, 4, "off"), ((64, 48, 24), 4, 6, "f16"),
    ((192, 72, 20), 4, 6, "off"), ((192, 40, 12), 3, 5, "f16"), ((160, 72, 20), 4, 6, "off"), ((150, 56, 13), 2, 4, "off")])
def test_ring_kernel_sixteen_wave_blocks_bitwise(dims, sweeps, chunk, heavy):
    """The 32-row ring tile as 16 waves of 2 rows (K <= 4), f32 and f16
    pressure, against the oracle bit for bit.  Most of these grids leave at
    most 32 - 2K columns at the right wall, which the narrow 32-column tile
    covers with two y tiles per wave (an odd y-tile count leaves the last
    pair's upper half unused); 160 columns at K = 4 and 150 at K = 2 take
    full tiles only."""
    g, o = _ring_pair(dims, (11, 9), sweeps, heavy=heavy, chunk=chunk, waves=16)
    info = g.info["jacobi"]
    assert info["waves"] == 16 and info["tile_height"] == 32 and info["sweeps_per_launch"] == sweeps
    o.step(DT, 2)
    g.step(DT, 2)
    assert_fields_equal(g, o, f"ring 16x2 {dims} K={sweeps} {heavy}", names=["flow", "species", "divergence", "reaction"])
    np.testing.assert_array_equal(_bits(g.read_field("pressure")), _bits(o.read(orc.FIELD_PRESSURE)))
    g.close()


def _layout_cases():
    """Grids drawn (seeded) across the widths where the 16-wave layout does
    and does not take the 32-column wall tile, with odd and even y-tile
    counts, short z chunks and every depth."""
    rng = np.random.default_rng(20261204)
    cases = []
    for _ in range(12):
        k = int(rng.integers(1, 5))
        nx = int(rng.integers(64, 260))
        ny = int(rng.integers(32, 120))
        nz = int(rng.integers(k + 2, 24))
        chunk = int(rng.integers(2, 9))
        cases.append(((nx, ny, nz), k, chunk))
    return cases


@pytest.mark.parametrize("dims,sweeps,chunk", _layout_cases())
def test_ring_kernel_sixteen_wave_layouts_bitwise(dims, sweeps, chunk):
    g, o = _ring_pair(dims, (7, 5), sweeps, chunk=chunk, waves=16)
    o.step(DT, 1)
    g.step(DT, 1)
    assert_fields_equal(g, o, f"ring 16x2 layout {dims} K={sweeps} chunk {chunk}",
                        names=["flow", "species", "divergence", "reaction"])
    np.testing.assert_array_equal(_bits(g.read_field("pressure")), _bits(o.read(orc.FIELD_PRESSURE)))
    np.testing.assert_array_equal(_bits(g.read_field("correction")), _bits(o.read(orc.FIELD_CORRECTION)))
    g.close()


@pytest.mark.parametrize("dims,sweeps,chunk,waves,heavy", [
    ((64, 96, 64), 4, 1, 16, "off"), ((64, 96, 64), 4, 2, 16, "off"), ((64, 96, 64), 4, 3, 16, "f16"),
    ((96, 64, 80), 5, 2, 8, "off"), ((96, 64, 80), 5, 1, 8, "off"), ((64, 128, 64), 3, 2, 8, "f16"),
    ((70, 40, 23), 2, 1, 16, "off")])
def test_ring_kernel_short_chunks_bitwise(dims, sweeps, chunk, waves, heavy):
    """Chunks of 1-3 planes, which the tuner picks on small grids (each
    block streams its chunk plus K halo planes on either side)."""
    g, o = _ring_pair(dims, (9, 7), sweeps, heavy=heavy, chunk=chunk, waves=waves)
    info = g.info["jacobi"]
    assert info["chunk_z"] == chunk and info["sweeps_per_launch"] == sweeps and info["waves"] == waves
    o.step(DT, 1)
    g.step(DT, 1)
    assert_fields_equal(g, o, f"ring short chunk {dims} K={sweeps} chunk {chunk} waves {waves} {heavy}",
                        names=["flow", "species", "divergence", "reaction"])
    np.testing.assert_array_equal(_bits(g.read_field("pressure")), _bits(o.read(orc.FIELD_PRESSURE)))
    g.close()


@pytest.mark.parametrize("height,dims,sweeps,chunk,prefetch", [
    (48, (80, 100, 20), 3, 7, False), (48, (80, 100, 20), 2, 5, True), (48, (64, 48, 12), 5, 0, False),
    (64, (64, 130, 17), 3, 6, False), (64, (64, 130, 17), 3, 6, True), (64, (128, 96, 40), 4, 9, False),
    (64, (96, 64, 9), 2, 4, True), (32, (72, 40, 14), 3, 5, True), (64, (70, 200, 11), 5, 3, False)])
def test_ring_kernel_tall_tiles_bitwise(height, dims, sweeps, chunk, prefetch):
    """Ring tiles of 48 and 64 rows (8 waves of 6 or 8 rows), with and
    without the one-plane prefetch, against the oracle bit for bit."""
    g, o = _ring_pair(dims, (11, 9), sweeps, chunk=chunk, height=height, prefetch=prefetch)
    info = g.info["jacobi"]
    assert info["tile_height"] == height and info["prefetch"] == int(prefetch) and info["sweeps_per_launch"] == sweeps
    o.step(DT, 2)
    g.step(DT, 2)
    assert_fields_equal(g, o, f"ring H={height} {dims} K={sweeps} pf={prefetch}",
                        names=["flow", "species", "divergence", "reaction"])
    np.testing.assert_array_equal(_bits(g.read_field("pressure")), _bits(o.read(orc.FIELD_PRESSURE)))
    np.testing.assert_array_equal(_bits(g.read_field("correction")), _bits(o.read(orc.FIELD_CORRECTION)))
    g.close()


@pytest.mark.parametrize("sweeps", [1, 2, 3, 4, 5, 6, 7, 8])
def test_ring_kernel_sweep_depths_bitwise(sweeps):
    """Register-ring Jacobi with wall-aligned, overlapping tiles (nx, ny not
    multiples of the tile interior) and chunks that do not divide nz."""
    dims = (80, 72, 30)
    g, o = _ring_pair(dims, (11, 9), sweeps, chunk=7)
    o.step(DT, 2)
    g.step(DT, 2)
    assert_fields_equal(g, o, f"ring K={sweeps}", names=["flow", "species", "divergence", "reaction"])
    np.testing.assert_array_equal(_bits(g.read_field("pressure")), _bits(o.read(orc.FIELD_PRESSURE)))
    np.testing.assert_array_equal(_bits(g.read_field("correction")), _bits(o.read(orc.FIELD_CORRECTION)))
    g.close()


@pytest.mark.parametrize("dims,chunk", [((64, 32, 16), 0), ((64, 32, 16), 5), ((128, 96, 40), 0), ((96, 160, 8), 3)])
def test_ring_kernel_tile_shapes_bitwise(dims, chunk):
    g, o = _ring_pair(dims, (8, 8), 3, chunk=chunk)
    o.step(DT, 2)
    g.step(DT, 2)
    assert_fields_equal(g, o, f"ring {dims} chunk {chunk}")
    g.close()


def test_ring_kernel_f16_pressure_bitwise():
    dims = (64, 48, 24)
    g, o = _ring_pair(dims, (9, 7), 4, heavy="f16", chunk=6)
    o.step(DT, 3)
    g.step(DT, 3)
    assert_fields_equal(g, o, "ring f16 pressure", names=["flow", "species", "divergence", "reaction"])
    np.testing.assert_array_equal(_bits(g.read_field("pressure")), _bits(o.read(orc.FIELD_PRESSURE)))
    g.close()


def test_autotuned_plan_bitwise_and_reported():
    """The autotuner picks some ring plan on a ring-sized grid; whatever it
    picks, the step stays bit-identical to the oracle, and a second solver of
    the same shape reuses the cached choice."""
    dims = (96, 64, 40)
    o = orc.Oracle(dims, precision=32, pressure_iterations=13, correction_iterations=10)
    o.set_bool("cinematic", True)
    g = FireXHipSolver(dims, tier="hd", pressure_iterations=13, correction_iterations=10, plan="fused",
                       jacobi_autotune=True)
    info = g.info["jacobi"]
    assert info["kernel"] == "ring" and 1 <= info["sweeps_per_launch"] <= 5 and info["chunk_z"] >= 1
    assert info["tuned_us_per_sweep"] > 0
    o.step(DT, 2)
    g.step(DT, 2)
    assert_fields_equal(g, o, f"autotuned {info}", names=["flow", "species", "divergence", "reaction"])
    np.testing.assert_array_equal(_bits(g.read_field("pressure")), _bits(o.read(orc.FIELD_PRESSURE)))
    g2 = FireXHipSolver(dims, tier="hd", pressure_iterations=13, correction_iterations=10, plan="fused",
                        jacobi_autotune=True)
    assert g2.info["jacobi"] == info
    g.close()
    g2.close()


def test_more_than_1024_jacobi_launches_bitwise():
    """1100 one-sweep launches in one solve (a schedule longer than any fixed
    launch table) still run every sweep."""
    dims = (16, 24, 16)
    g, o = make_pair(dims, iters=(1100, 3), plan="fused", sweeps=1)
    o.step(DT, 1)
    g.step(DT, 1)
    assert_fields_equal(g, o, "1100 sweeps", names=["flow", "species", "divergence", "reaction"])
    np.testing.assert_array_equal(_bits(g.read_field("pressure")), _bits(o.read(orc.FIELD_PRESSURE)))
    g.close()


@pytest.mark.parametrize("storage", ["f32", "f16", "bf16"])
@pytest.mark.parametrize("view", [0, 7])
def test_pack_fields_bitwise(storage, view):
    """packFields (qualityShaders.ts:37-73) after a few burning steps with a
    liquid field: the three render volumes, f16 texels, bit for bit; view 7
    is the divergence view (refreshed divergence, zero boundary shell)."""
    dims = (20, 28, 16)
    em = seeded_emitters(dims, 2, radius_range=(3.0, 5.0), seed=11)
    g, o = make_pair(dims, iters=(8, 8), plan="fused", storage=storage, emitters=(em, 0.1, 0.0, 11))
    liquid = np.random.default_rng(2).integers(0, 6000, size=(dims[2], dims[1], dims[0])).astype(np.uint32)
    g.write_field("liquid", liquid)
    o.write(orc.FIELD_LIQUID, liquid)
    o.step(DT, 4)
    g.step(DT, 4)
    g.set_parameter("viewMode", view)
    o.set_parameter("viewMode", view)
    g.pack_fields()
    pf, pc, pm = o.pack_fields()
    for name, ref in (("packed_flow", pf), ("packed_chemistry", pc), ("packed_media", pm)):
        got = g.read_field(name)
        assert np.array_equal(_bits(got), _bits(ref.astype(np.float32))), \
            f"{name} ({storage}, view {view}) differs in {int((_bits(got) != _bits(ref.astype(np.float32))).sum())} values"
    assert pm[..., 1].max() > 0 or pm[..., 0].max() > 0  # the plume burns
    g.close()
