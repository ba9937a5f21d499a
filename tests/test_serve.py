"""Serving: many resident plumes, admission by HBM budget, continuous
batching.  Each plume's result must be the same alone and in the mix (bit
for bit: no plume's kernels touch another plume's buffers)."""
import numpy as np
import pytest

from fire_amd import FireXHipSolver, device_count, seeded_emitters
from fire_amd import serve as sv
from fire_amd import solver as fs

DT = 1.0 / 60.0


def test_server_needs_a_device():
    if device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(fs.FireXError) as e:
        sv.FireXServer()
    assert e.value.code == fs.ERR_DEVICE


def test_poisson_stream_is_seeded_and_bounded():
    a = sv.poisson_stream(50, 20.0, seed=5)
    b = sv.poisson_stream(50, 20.0, seed=5)
    assert a == b
    times = [t for t, *_ in a]
    assert all(x < y for x, y in zip(times, times[1:]))
    assert 1.5 < times[-1] < 4.0  # mean gap 1/20 s
    assert all(60 <= s <= 600 for _, _, s, _ in a)


def _standalone(dims, steps, emitters, seed, storage="f16", iters=(8, 8)):
    with FireXHipSolver(dims, tier="uhd", pressure_iterations=iters[0], correction_iterations=iters[1],
                        species_storage=storage, plan="fused", heavy_mode="off") as g:
        g.set_emitters(emitters, jitter=0.1, velocity_noise=0.25, seed=seed)
        g.step(DT, steps)
        return {n: g.read_field(n) for n in ("flow", "species", "pressure", "reaction")}


MIX = [((32, 48, 32), 7), ((24, 40, 28), 12), ((40, 56, 36), 5), ((32, 48, 32), 9), ((16, 24, 16), 15),
       ((48, 64, 40), 4)]


@pytest.mark.gpu
def test_mixed_stream_results_equal_standalone_bitwise():
    if device_count() < 1:
        pytest.fail("no HIP device")
    with sv.FireXServer(max_active=3, streams=4, plan="fused", capture_graphs=True) as srv:
        ids = []
        for k, (dims, steps) in enumerate(MIX[:3]):
            em = seeded_emitters(dims, 2, radius_range=(3, 6), seed=100 + k)
            ids.append((srv.submit(sv.make_request(dims, steps, pressure_iterations=8, correction_iterations=8,
                                                   emitters=em, jitter=0.1, velocity_noise=0.25, seed=100 + k,
                                                   keep_result=True)), dims, steps, em, 100 + k))
        # requests that arrive while the first ones run join the batch
        for _ in range(3):
            srv.pump()
        for k, (dims, steps) in enumerate(MIX[3:], start=3):
            em = seeded_emitters(dims, 2, radius_range=(3, 6), seed=100 + k)
            ids.append((srv.submit(sv.make_request(dims, steps, pressure_iterations=8, correction_iterations=8,
                                                   emitters=em, jitter=0.1, velocity_noise=0.25, seed=100 + k,
                                                   keep_result=True)), dims, steps, em, 100 + k))
        srv.drain()
        stats = srv.stats()
        assert stats["done"] == len(MIX) and stats["failed"] == 0
        assert stats["steps_completed"] == sum(s for _, s in MIX)
        for rid, dims, steps, em, seed in ids:
            st = srv.status(rid)
            assert st["state_name"] == "done" and st["steps_done"] == steps
            alone = _standalone(dims, steps, em, seed)
            for name, ref in alone.items():
                got = srv.read_field(rid, name)
                np.testing.assert_array_equal(got.view(np.uint32), ref.view(np.uint32), err_msg=f"{rid} {name}")
            srv.release(rid)
        assert srv.stats()["resident_bytes"] == 0


@pytest.mark.gpu
def test_admission_respects_the_hbm_budget_and_slots():
    if device_count() < 1:
        pytest.fail("no HIP device")
    dims = (64, 96, 64)
    one = FireXHipSolver(dims, tier="uhd", species_storage="f16", heavy_mode="off", plan="fused")
    per_plume = one.info["device_bytes"]
    one.close()
    budget = int(per_plume * 2.5)  # room for two plumes at a time
    with sv.FireXServer(memory_budget_bytes=budget, max_active=8, streams=4) as srv:
        ids = [srv.submit(sv.make_request(dims, 6, pressure_iterations=4, correction_iterations=4, seed=i))
               for i in range(5)]
        peak_active, peak_bytes = 0, 0
        while srv.pump():
            st = srv.stats()
            peak_active = max(peak_active, st["active"])
            peak_bytes = max(peak_bytes, st["resident_bytes"])
            assert st["resident_bytes"] <= budget
        st = srv.stats()
        assert st["done"] == 5 and peak_active == 2 and peak_bytes <= budget
        admits = sorted(srv.status(i)["admit_time"] for i in ids)
        assert admits[2] > admits[1]  # the third waited for memory to free up
        assert st["latency_ms_p50"] > 0 and st["latency_ms_p99"] >= st["latency_ms_p50"]
    with pytest.raises(fs.FireXError) as e:
        with sv.FireXServer(memory_budget_bytes=per_plume // 2) as srv:
            srv.submit(sv.make_request(dims, 1))
    assert e.value.code == fs.ERR_BUDGET
