"""Pins of the oracle's presentation passes (bloomFragment, brightExtract,
filmic, dither and compositeFragment, qualityShaders.ts:613-671) by what the
mathematics fixes:

* a uniform scene: every tap samples the one colour, so the bloom is
  brightExtract(colour) (zero below the knee) and the frame's mean over many
  pixels is the filmic curve of scene + 0.1 bloom times 255 (the dither has
  zero mean and at most half a step), in float64 here;
* a scene quadratic in x (red) and in y (green): at a size divisible by four
  the side taps land on texel centres and the centre tap half way between, so
  the nine-tap sum is Q(centre) + 18.25 q exactly — a wrong weight, sign,
  radius or axis moves it;
* the standard presentation stores rint(clamp(c, 0, 1) * 255) of the scene.
"""
import numpy as np
import pytest

from fire_amd import seeded_emitters
from oracle import oracle as orc

THRESHOLD, KNEE = 1.15, 0.35
LUMA = np.array([0.2126, 0.7152, 0.0722])


def bright_extract(c):
    """brightExtract of rgb rows (..., 3) in float64."""
    lum = c @ LUMA
    soft = np.clip((lum - THRESHOLD + KNEE) / (2.0 * KNEE), 0.0, 1.0)
    contribution = np.maximum(lum - THRESHOLD, 0.0) + soft * soft * KNEE
    return c * (contribution / np.maximum(lum, 1e-4))[..., None]


def filmic(c, exposure):
    x = np.maximum(c * max(exposure, 0.25), 0.0)
    m = np.clip(x * (2.51 * x + 0.03) / (x * (2.43 * x + 0.59) + 0.14), 0.0, 1.0)
    return m ** (1.0 / 2.2)


def uniform(w, h, rgb):
    img = np.empty((h, w, 4), dtype=np.float32)
    img[..., :3] = rgb
    img[..., 3] = 1.0
    return img


@pytest.mark.parametrize("rgb", [(0.05, 0.05, 0.05), (0.2, 0.35, 0.5), (0.9, 0.9, 0.9), (2.5, 1.2, 0.4),
                                 (6.0, 3.0, 1.0)])
@pytest.mark.parametrize("exposure", [0.1, 1.0, 1.7])
def test_uniform_scene_frame_mean_is_the_filmic_curve(rgb, exposure):
    w, h = 256, 192
    frame, bloom = orc.present(uniform(w, h, rgb), exposure, True)
    c = np.array(rgb, dtype=np.float64)
    b = bright_extract(c)
    assert bloom.shape == (h // 4, w // 4, 4)
    np.testing.assert_allclose(bloom[..., :3], np.broadcast_to(b, bloom[..., :3].shape), rtol=2e-3, atol=1e-6)
    if c @ LUMA < THRESHOLD - KNEE:
        assert not bloom[..., :3].any(), "no bloom below the knee"
    want = filmic(c + b * 0.10, exposure) * 255.0
    got = frame[..., :3].astype(np.float64)
    # the dither moves a pixel by at most half a step, rounding by another half
    assert np.abs(got - want).max() <= 1.0 + 1e-3
    # ... and has zero mean, so the mean is the curve itself
    np.testing.assert_allclose(got.reshape(-1, 3).mean(axis=0), want, atol=0.01)
    assert (frame[..., 3] == 255).all()


@pytest.mark.parametrize("w,h", [(200, 128), (96, 160)])
def test_bloom_taps_of_a_quadratic_scene(w, h):
    q = 1e-3
    x = np.arange(w, dtype=np.float64)
    y = np.arange(h, dtype=np.float64)
    img = np.empty((h, w, 4), dtype=np.float32)
    img[..., 0] = (1.5 + q * (x - w / 2) ** 2)[None, :]
    img[..., 1] = (1.5 + q * (y - h / 2) ** 2)[:, None]
    img[..., 2] = 1.5
    img[..., 3] = 1.0
    _, bloom = orc.present(img, 1.0, True)
    bw, bh = w // 4, h // 4
    tx = (np.arange(bw) + 0.5) / bw * w - 0.5
    ty = (np.arange(bh) + 0.5) / bh * h - 0.5
    pre = np.empty((bh, bw, 3))
    pre[..., 0] = (1.5 + q * (tx - w / 2) ** 2 + 18.25 * q)[None, :]
    pre[..., 1] = (1.5 + q * (ty - h / 2) ** 2 + 18.25 * q)[:, None]
    pre[..., 2] = 1.5
    want = bright_extract(pre)
    inside_x = (tx - 5.5 >= 0) & (tx + 5.5 <= w - 1)
    inside_y = (ty - 5.5 >= 0) & (ty + 5.5 <= h - 1)
    sel = np.ix_(inside_y, inside_x)
    assert inside_x.sum() > bw // 2 and inside_y.sum() > bh // 2
    np.testing.assert_allclose(bloom[..., :3][sel], want[sel], rtol=1e-3)
    # the 18.25 q term is what the side taps add: far above the tolerance
    assert 18.25 * q / 1.5 > 5e-3


def test_standard_presentation_is_the_scene_in_unorm8():
    rng = np.random.default_rng(3)
    w, h = 61, 37
    img = rng.uniform(-0.5, 1.5, size=(h, w, 4)).astype(np.float32)
    img[0, :3, :3] = [[0.0, 1.0, 2.0], [-1.0, 0.5 / 255.0, 1.5 / 255.0], [127.5 / 255.0, 128.5 / 255.0, 1.0]]
    frame, bloom = orc.present(img, 1.0, False)
    assert bloom is None
    assert (frame[..., 3] == 255).all()
    want = np.rint(np.clip(img[..., :3], 0.0, 1.0).astype(np.float64) * 255.0)
    # ties: c * 255 rounded in f32 first, then to even
    exact = np.rint(np.clip(img[..., :3], 0.0, 1.0) * np.float32(255.0))
    np.testing.assert_array_equal(frame[..., :3], exact.astype(np.uint8))
    assert np.abs(frame[..., :3].astype(np.float64) - want).max() <= 1.0
    assert frame[0, 0, :3].tolist() == [0, 255, 255]
    assert frame[0, 1, 0] == 0


def test_render_frame_presents_the_ray_marched_scene():
    o = orc.Oracle((24, 32, 24), precision=32, pressure_iterations=4, correction_iterations=4)
    o.set_heavy_mode(orc.HEAVY_OFF)
    o.set_emitters(seeded_emitters((24, 32, 24), 1, radius_range=(3.0, 4.0), seed=7), 0.1, 0.0, 7)
    o.step(1.0 / 60.0, 12)
    for cinematic in (True, False):
        o.set_bool("cinematic", cinematic)
        o.set_parameter("exposure", 1.3)
        scene = o.render_volume(64, 40, 96)
        frame = o.render_frame(64, 40, 96)
        want, _ = orc.present(scene, 1.3, cinematic)
        np.testing.assert_array_equal(frame, want)
    o.close()
