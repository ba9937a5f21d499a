"""firex_reconfigure on the GPU: per-step settings changed in place keep the
fields, and a rebuilt buffer set is a fresh plume on the new grid with the
parameters kept (FireXEngine.ts:799-927, 1891-1931); both equal the FP32
oracle bit for bit."""
import numpy as np
import pytest

from fire_amd import FireXError, FireXHipSolver, device_count
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
FIELDS = (("flow", orc.FIELD_FLOW), ("species", orc.FIELD_SPECIES), ("pressure", orc.FIELD_PRESSURE),
          ("reaction", orc.FIELD_REACTION), ("divergence", orc.FIELD_DIVERGENCE))
PARAMS = {"fuelRate": 0.45, "vorticity": 2.8, "firePower": 2.2}


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if device_count() < 1:
        pytest.fail("no HIP device visible to the reconfigure tests")


def bits(a):
    return np.ascontiguousarray(a, dtype=np.float32).view(np.uint32)


def assert_equal(g, o, label):
    for name, fid in FIELDS:
        a, b = bits(g.read_field(name)), bits(o.read(fid).astype(np.float32))
        assert np.array_equal(a, b), f"{label}: {name} differs in {(a != b).sum()} values"


def oracle(dims, iters, cinematic):
    o = orc.Oracle(dims, precision=32, pressure_iterations=iters[0], correction_iterations=iters[1])
    o.set_bool("cinematic", cinematic)
    o.set_heavy_mode(orc.HEAVY_OFF)
    for k, v in PARAMS.items():
        o.set_parameter(k, v)
    o.reset()
    return o


def solver(dims, iters, **kw):
    g = FireXHipSolver(dims, tier="hd", pressure_iterations=iters[0], correction_iterations=iters[1], plan="fused",
                       heavy_mode="off", **kw)
    for k, v in PARAMS.items():
        g.set_parameter(k, v)
    g.reset()
    return g


def test_in_place_iterations_keep_the_fields():
    """New iteration counts and graph capture on the same buffer set: no
    rebuild, the state continues; two steps at (8, 8) then two at (12, 4)
    equal an oracle run handed the state after the first two."""
    dims = (64, 48, 40)
    print(f"reconfigure in place: {dims}", flush=True)
    g = solver(dims, (8, 8))
    g.step(DT, 2)
    a = oracle(dims, (8, 8), True)
    a.step(DT, 2)
    assert_equal(g, a, "before")
    assert g.reconfigure(pressure_iterations=12, correction_iterations=4, capture_graphs=1) is False
    g.step(DT, 2)
    b = oracle(dims, (12, 4), True)
    b.write(orc.FIELD_FLOW, a.read(orc.FIELD_FLOW))
    b.write(orc.FIELD_SPECIES, a.read(orc.FIELD_SPECIES))
    b.write(orc.FIELD_PRESSURE, a.read(orc.FIELD_PRESSURE))
    b.set_time(a.simulation_time, 2)
    b.step(DT, 2)
    assert_equal(g, b, "after in-place change")
    assert g.info["frame_index"] == 4
    g.close()


def test_rebuild_is_a_fresh_plume_with_the_parameters_kept():
    """A new grid rebuilds: time and frame restart, fields are the reset
    state of the new grid with the kept parameters, and stepping equals the
    oracle on that grid.  A guard below the estimated peak refuses and
    changes nothing."""
    dims, new = (48, 64, 40), (40, 56, 48)
    print(f"reconfigure rebuild: {dims} -> {new}", flush=True)
    g = solver(dims, (6, 6))
    g.step(DT, 3)
    before = {name: bits(g.read_field(name)).copy() for name, _ in FIELDS}
    with pytest.raises(FireXError, match="Transactional rebuild peak"):
        g.reconfigure(memory_guard_bytes=1 << 20, dims=new)
    assert g.dims == dims and g.info["frame_index"] == 3
    for name, _ in FIELDS:
        assert np.array_equal(bits(g.read_field(name)), before[name]), name
    assert g.reconfigure(memory_guard_bytes=1 << 30, dims=new) is True
    assert g.info["dims"] == new and g.info["frame_index"] == 0 and g.info["simulation_time"] == 0
    g.step(DT, 3)
    o = oracle(new, (6, 6), True)
    o.step(DT, 3)
    assert_equal(g, o, "after rebuild")
    g.close()
