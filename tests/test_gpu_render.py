"""GPU parity of the volume ray marcher (volumeFragment,
qualityShaders.ts:374-575): the HIP kernel's rgba16f scene image against the
FP32 oracle's, bit for bit, on burning plumes after real solver steps —
every view mode, the cinematic and the standard path, water with droplets,
an HD-size grid — plus the device checksum of the image."""
import numpy as np
import pytest

from fire_amd import FireXHipSolver, checksum_bytes, device_count, seeded_emitters
from oracle import oracle as orc

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if device_count() < 1:
        pytest.fail("no HIP device visible to the render parity tests")


def _bits16(a):
    return np.ascontiguousarray(a, dtype=np.float32).astype(np.float16).view(np.uint16)


def assert_images_equal(g, o, label):
    a = g.read_field("scene")
    b = o
    if np.array_equal(_bits16(a), _bits16(b)):
        return
    bad = np.argwhere(_bits16(a) != _bits16(b))
    raise AssertionError(f"{label}: {len(bad)} of {a.size} values differ; first at {bad[0].tolist()} "
                         f"gpu={a[tuple(bad[0])]!r} oracle={b[tuple(bad[0])]!r}")


def make_pair(dims, steps, cinematic=True, params=None, particles=0, water=False, iters=(16, 16)):
    o = orc.Oracle(dims, precision=32, pressure_iterations=iters[0], correction_iterations=iters[1])
    o.set_bool("cinematic", cinematic)
    o.set_heavy_mode(orc.HEAVY_OFF)
    o.set_particle_count(particles)
    g = FireXHipSolver(dims, tier="hd", pressure_iterations=iters[0], correction_iterations=iters[1], plan="fused",
                       heavy_mode="off", particle_count=particles, cinematic=cinematic)
    em = seeded_emitters(dims, 2, radius_range=(3.0, 5.0), seed=7)
    o.set_emitters(em, 0.1, 0.0, 7)
    g.set_emitters(em, 0.1, 0.0, 7)
    for k, v in (params or {}).items():
        o.set_parameter(k, v)
        g.set_parameter(k, v)
    if water:
        o.set_bool("waterEnabled", True)
        g.set_parameter("waterEnabled", True)
    print(f"render pair {dims} {steps} steps cinematic {cinematic} water {water}", flush=True)
    g.step(DT, steps)
    o.step(DT, steps)
    return g, o


@pytest.mark.parametrize("cinematic", [True, False])
def test_every_view_mode_bitwise(cinematic):
    g, o = make_pair((48, 64, 48), 40, cinematic=cinematic)
    w, h = 192, 108
    for view in range(8):
        o.set_parameter("viewMode", view)
        g.set_parameter("viewMode", view)
        g.render_volume(w, h, 256)
        ref = o.render_volume(w, h, 256)
        assert_images_equal(g, ref, f"view {view} cinematic {cinematic}")
        if view in (0, 1, 2, 3):
            # the plume is drawn (the studio background alone stays below 0.06)
            assert ref[..., :3].max() > 0.15, f"view {view}: nothing drawn"
    raw = g.read_field_raw("scene")
    assert g.field_checksum("scene") == checksum_bytes(raw)
    g.close()
    o.close()


def test_camera_exposure_zoom_and_optical_detail_bitwise():
    params = {"cameraYaw": 0.9, "cameraPitch": 0.35, "cameraDistance": 3.2, "exposure": 1.7, "viewZoom": 1.3,
              "frontDefinition": 2.2, "opticalDetailTarget": 256}
    g, o = make_pair((48, 64, 48), 30, params=params)
    for w, h, steps in ((160, 200, 320), (97, 61, 80)):
        g.render_volume(w, h, steps)
        ref = o.render_volume(w, h, steps)
        assert_images_equal(g, ref, f"{w}x{h} {steps} steps")
    g.close()
    o.close()


def test_water_and_liquid_view_bitwise():
    params = {"waterFlow": 0.74, "sprayAngle": 30.0, "aimHeight": 0.14, "nozzleType": 1}
    g, o = make_pair((48, 64, 48), 30, params=params, particles=1024, water=True)
    for view in (0, 5, 6):
        o.set_parameter("viewMode", view)
        g.set_parameter("viewMode", view)
        g.render_volume(160, 90, 256)
        ref = o.render_volume(160, 90, 256)
        assert_images_equal(g, ref, f"water view {view}")
    g.close()
    o.close()


def test_hd_grid_render_bitwise():
    """The hd tier's grid (128 x 192 x 128, 3.1 M cells) after a few steps,
    rendered at 320 x 180 with the hd ray budget (256)."""
    g, o = make_pair((128, 192, 128), 3, iters=(8, 8))
    for view in (0, 7):
        o.set_parameter("viewMode", view)
        g.set_parameter("viewMode", view)
        g.render_volume(320, 180, 256)
        ref = o.render_volume(320, 180, 256)
        assert_images_equal(g, ref, f"hd grid view {view}")
    g.close()
    o.close()


@pytest.mark.parametrize("cinematic", [True, False])
def test_presented_frame_bitwise(cinematic):
    """firex_render_frame (ray march, quarter-size bloom, filmic composite and
    dither; or the standard presentation) against the oracle's frame, byte for
    byte, at even and odd sizes and several exposures (0.1 is clamped to 0.25);
    then the presentation alone over the GPU's own scene image."""
    g, o = make_pair((48, 64, 48), 40, cinematic=cinematic)
    for (w, h, steps), exposure in (((192, 108, 256), 1.0), ((97, 61, 128), 1.7), ((130, 203, 96), 0.1)):
        o.set_parameter("exposure", exposure)
        g.set_parameter("exposure", exposure)
        frame = g.render_frame(w, h, steps)
        ref = o.render_frame(w, h, steps)
        assert frame.shape == (h, w, 4)
        if not np.array_equal(frame, ref):
            bad = np.argwhere(frame != ref)
            raise AssertionError(f"{w}x{h} exposure {exposure}: {len(bad)} bytes differ; first at "
                                 f"{bad[0].tolist()} gpu={frame[tuple(bad[0])]} oracle={ref[tuple(bad[0])]}")
        assert frame[..., :3].max() > 40, "nothing drawn"
        want, _ = orc.present(g.read_field("scene"), exposure, cinematic)
        np.testing.assert_array_equal(frame, want)
        assert g.field_checksum("frame") == checksum_bytes(frame.tobytes())
    g.close()
    o.close()


def _frames_equal(frame, ref, label):
    if not np.array_equal(frame, ref):
        bad = np.argwhere(frame != ref)
        raise AssertionError(f"{label}: {len(bad)} bytes differ; first at {bad[0].tolist()} "
                             f"gpu={frame[tuple(bad[0])]} oracle={ref[tuple(bad[0])]}")


HOSE = {"waterFlow": 0.8, "sprayAngle": 24.0, "aimHeight": 0.30, "nozzleType": 1, "cameraYaw": 0.5}


@pytest.mark.parametrize("cinematic", [True, False])
def test_droplet_streaks_bitwise(cinematic):
    """The hose's droplets drawn over the scene (cinematic: into the rgba16f
    scene image before bloom) or into the rgba8 canvas (standard), in the
    beauty and liquid views, byte for byte; other views draw none."""
    g, o = make_pair((48, 64, 48), 24, cinematic=cinematic, params=HOSE, particles=2048, water=True)
    drawn = False
    for view, (w, h) in ((0, (320, 180)), (6, (257, 143)), (2, (160, 90))):
        o.set_parameter("viewMode", view)
        g.set_parameter("viewMode", view)
        frame = g.render_frame(w, h, 128)
        ref = o.render_frame(w, h, 128)
        _frames_equal(frame, ref, f"view {view} {w}x{h} cinematic {cinematic}")
        if cinematic:
            assert_images_equal(g, o.render_volume(w, h, 128), f"scene view {view}")
        if view in (0, 6):
            # the streaks changed the image: with the droplets massless it differs
            parts = g.read_field("particles")
            massless = parts.copy()
            massless[:, 7] = 0.0
            o.write_particle_buffers(massless, massless)
            bare = o.render_frame(w, h, 128)
            drawn |= int(np.count_nonzero(np.any(bare != ref, axis=-1))) > 20
            o.write_particle_buffers(parts, parts)
    assert drawn, "no droplet streak reached the image"
    g.close()
    o.close()


@pytest.mark.parametrize("w,h", [(4096, 2), (16384, 4)])
def test_droplet_streaks_extreme_aspect_bitwise(w, h):
    """The streak radius is 1.15 px of the height in NDC units, so on a very
    wide image each quad spans most of the width: the bin lists (capacity
    8 x 1024 + 4096 ids) overrun and the tiles read every droplet in order."""
    g, o = make_pair((48, 64, 48), 24, cinematic=True, params=HOSE, particles=1024, water=True)
    frame = g.render_frame(w, h, 64)
    ref = o.render_frame(w, h, 64)
    _frames_equal(frame, ref, f"{w}x{h}")
    assert_images_equal(g, o.render_volume(w, h, 64), f"scene {w}x{h}")
    g.close()
    o.close()
