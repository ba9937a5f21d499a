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


# (dims, steps, (N, M), species storage): small grids the 16-wave ring plan runs
# (nx >= 64, ny >= 32, nz >= 6), a correction-free pair (M = 0 clears the
# residual fields inside the group step), iteration counts that end in a
# partial launch (10 = 4 + 4 + 2, 7 = 4 + 3) and a grid the ring kernel
# cannot run (ny < 32: it steps alone).
BATCH_MIX = [((64, 48, 40), 6, (10, 7), "f16"), ((96, 64, 48), 4, (10, 7), "f32"),
             ((72, 40, 36), 7, (10, 7), "bf16"), ((64, 32, 24), 5, (10, 7), "f16"),
             ((64, 40, 32), 5, (9, 0), "f16"), ((80, 36, 30), 4, (9, 0), "f32"),
             ((48, 24, 32), 4, (10, 7), "f16")]


@pytest.mark.gpu
def test_batched_groups_equal_standalone_bitwise():
    """batch_plumes on one stream: plumes with equal iteration counts step
    as groups whose Jacobi launches cover all of them; each plume's fields
    equal the same plume stepped alone bit for bit, and the first equals the
    FP32 oracle."""
    if device_count() < 1:
        pytest.fail("no HIP device")
    from oracle import oracle as orc
    # 96 x 64 x 48 is above the batch size limit: it steps alone on the
    # second stream with the other plumes grouped on the first
    with sv.FireXServer(max_active=8, streams=2, plan="fused", capture_graphs=True, batch_plumes=True,
                        batch_max_cells=200_000, batch_streams=1) as srv:
        ids = []
        for k, (dims, steps, iters, storage) in enumerate(BATCH_MIX):
            em = seeded_emitters(dims, 2, radius_range=(3, 6), seed=400 + k)
            rid = srv.submit(sv.make_request(dims, steps, pressure_iterations=iters[0], correction_iterations=iters[1],
                                             species_storage=storage, emitters=em, jitter=0.1, velocity_noise=0.25,
                                             seed=400 + k, keep_result=True))
            ids.append((rid, dims, steps, iters, storage, em, 400 + k))
            if k == 3:
                srv.pump()  # the later requests join groups already running
        srv.drain()
        stats = srv.stats()
        print(f"batched: {stats['group_steps']} group steps covering {stats['grouped_steps']} plume steps", flush=True)
        assert stats["done"] == len(BATCH_MIX) and stats["failed"] == 0
        assert stats["group_steps"] > 0 and stats["grouped_steps"] >= 2 * stats["group_steps"]
        assert stats["grouped_steps"] <= sum(s for d, s, _, _ in BATCH_MIX[:-1] if np.prod(d) <= 200_000)
        for i, (rid, dims, steps, iters, storage, em, seed) in enumerate(ids):
            print(f"  plume {dims} {steps} steps {iters} {storage}", flush=True)
            with FireXHipSolver(dims, tier="uhd", pressure_iterations=iters[0], correction_iterations=iters[1],
                                species_storage=storage, plan="fused", heavy_mode="off") as g:
                g.set_emitters(em, jitter=0.1, velocity_noise=0.25, seed=seed)
                g.step(DT, steps)
                for name in ("flow", "species", "pressure", "correction", "reaction", "divergence"):
                    a, b = srv.read_field(rid, name), g.read_field(name)
                    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f"{dims} {name}")
            if i == 0:
                o = orc.Oracle(dims, precision=32, pressure_iterations=iters[0], correction_iterations=iters[1])
                o.set_bool("cinematic", True)
                o.set_heavy_mode(orc.HEAVY_OFF)
                o.set_species_storage(orc.SPECIES_F16)
                o.reset()
                o.set_emitters(em, 0.1, 0.25, seed)
                o.step(DT, steps)
                for name, fid in (("flow", orc.FIELD_FLOW), ("species", orc.FIELD_SPECIES),
                                  ("pressure", orc.FIELD_PRESSURE), ("correction", orc.FIELD_CORRECTION)):
                    a = np.ascontiguousarray(srv.read_field(rid, name), dtype=np.float32)
                    b = np.ascontiguousarray(o.read(fid), dtype=np.float32)
                    np.testing.assert_array_equal(a.view(np.uint32), b.view(np.uint32), err_msg=f"oracle {name}")
                o.close()
            srv.release(rid)


def test_batch_options_are_checked():
    cfg = sv.ServerConfig()
    cfg.max_active, cfg.streams, cfg.batch_plumes, cfg.batch_streams = 4, 2, 1, 3
    h = sv.ctypes.c_void_p()
    rc = sv._lib().firex_server_create(sv.ctypes.byref(cfg), sv.ctypes.byref(h))
    assert rc == fs.ERR_INVALID
    cfg.batch_streams, cfg.batch_max_cells = 1, -5
    assert sv._lib().firex_server_create(sv.ctypes.byref(cfg), sv.ctypes.byref(h)) == fs.ERR_INVALID
