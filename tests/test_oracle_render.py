"""Pins of the oracle's volume ray marcher (volumeFragment,
qualityShaders.ts:374-575) and detail volume (createPeriodicDetailVolume,
FireXEngine.ts:311-355) by facts outside its own code:

  * detail volume: the same integer/float64 arithmetic run by the JS engine
    (node), bit for bit; periodicity (texel 0 equals texel size - 1 on every
    axis), alpha 255;
  * empty scene: nothing is drawn, so the image does not depend on the grid
    or the ray budget, and rays above the horizon show the vertical
    background gradient mix(c0, c1, uv.y) with uv.y = 1 at the top row;
  * uniform medium in a diagnostic view: along every ray the transmittance
    is exp(-8.5 a L) for the chord length L of the ray through the domain
    box (a numpy slab test from the camera), so the pixel is
    c (1 - e) 1.35 + e * background, a closed form in a;
  * tone mapping of the non-cinematic path: the top rows equal the ACES
    curve of the gradient at the set exposure."""
import json
import math
import os
import shutil
import subprocess

import numpy as np
import pytest

from oracle import oracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NODE = shutil.which("node")
F = np.float32


def _f16(x):
    return np.asarray(x, dtype=np.float32).astype(np.float16).astype(np.float64)


def _oracle(dims=(24, 32, 24), cinematic=True):
    o = orc.Oracle(dims, precision=32, pressure_iterations=4, correction_iterations=4)
    o.set_bool("cinematic", cinematic)
    return o


_JS_DETAIL = r"""
// createPeriodicDetailVolume's arithmetic (FireXEngine.ts:311-355) in the
// reference's language: Math.imul, <<, >>>, float64 and Math.round.
const size = 64, period = size - 1, data = new Uint8Array(size * size * size * 4);
const hash = (x, y, z, seed) => { let v = Math.imul(x, 73856093) ^ Math.imul(y, 19349663) ^ Math.imul(z, 83492791) ^ seed;
  v ^= v << 13; v ^= v >>> 17; v ^= v << 5; return (v >>> 24) & 255; };
const fade = (v) => v * v * (3 - 2 * v);
const mix = (a, b, t) => a + (b - a) * t;
const noise = (x, y, z, f, seed) => {
  const px = (x / period) * f, py = (y / period) * f, pz = (z / period) * f;
  const x0 = Math.floor(px) % f, y0 = Math.floor(py) % f, z0 = Math.floor(pz) % f;
  const x1 = (x0 + 1) % f, y1 = (y0 + 1) % f, z1 = (z0 + 1) % f;
  const tx = fade(px - Math.floor(px)), ty = fade(py - Math.floor(py)), tz = fade(pz - Math.floor(pz));
  const lo = mix(mix(hash(x0, y0, z0, seed), hash(x1, y0, z0, seed), tx), mix(hash(x0, y1, z0, seed), hash(x1, y1, z0, seed), tx), ty);
  const hi = mix(mix(hash(x0, y0, z1, seed), hash(x1, y0, z1, seed), tx), mix(hash(x0, y1, z1, seed), hash(x1, y1, z1, seed), tx), ty);
  return Math.round(mix(lo, hi, tz)); };
for (let z = 0; z < size; z += 1) for (let y = 0; y < size; y += 1) for (let x = 0; x < size; x += 1) {
  const o = (x + size * (y + size * z)) * 4;
  data[o] = noise(x, y, z, 4, 0x51f15e5d); data[o + 1] = noise(x, y, z, 8, 0x6d2b79f5);
  data[o + 2] = noise(x, y, z, 14, 0x1b873593); data[o + 3] = 255; }
process.stdout.write(Buffer.from(data.buffer).toString('base64'));
"""


def test_detail_volume_matches_javascript_and_is_periodic():
    d = orc.detail_volume()
    assert np.all(d[..., 3] == 255)
    for axis in range(3):
        np.testing.assert_array_equal(np.take(d, 0, axis=axis), np.take(d, 63, axis=axis))
    assert 20 < d[..., 0].std() and d[..., :3].min() >= 0 and d[..., :3].max() <= 255
    if NODE is None:
        pytest.skip("node is absent")
    import base64
    res = subprocess.run([NODE, "-e", _JS_DETAIL], capture_output=True, text=True, timeout=120, check=True)
    js = np.frombuffer(base64.b64decode(res.stdout), dtype=np.uint8).reshape(64, 64, 64, 4)
    np.testing.assert_array_equal(d, js)


def _background_top(uvy):
    c0 = np.array([0.020, 0.024, 0.030], dtype=F)
    c1 = np.array([0.0025, 0.004, 0.007], dtype=F)
    t = F(min(max(uvy, 0.0), 1.0))
    return c0 * (F(1) - t) + c1 * t


def test_empty_scene_draws_only_the_background():
    w, h = 64, 36
    imgs = []
    for dims, steps in (((24, 32, 24), 80), ((40, 48, 40), 384)):
        o = _oracle(dims)
        imgs.append(o.render_volume(w, h, steps))
        o.close()
    np.testing.assert_array_equal(imgs[0], imgs[1])
    img = imgs[0]
    for row in (0, 1, 5):  # rays above the horizon: the vertical gradient
        uvy = 1.0 - (row + 0.5) / h
        np.testing.assert_array_equal(img[row, :, :3], np.broadcast_to(_f16(_background_top(uvy)), (w, 3)))
    assert np.all(img[..., 3] == 1.0)


def _camera(width, height, extent=(1.15, 1.35, 1.15), yaw=-0.34, pitch=0.03, distance=4.9, zoom=2.1):
    e = np.array(extent, dtype=np.float64)
    target = np.array([e[0] * 0.5, e[1] * 0.40, e[2] * 0.5])
    orbit = np.array([math.sin(yaw) * math.cos(pitch), math.sin(pitch), math.cos(yaw) * math.cos(pitch)])
    eye = target + orbit * (0.90 + distance * 0.18) * e.max()
    fwd = (target - eye) / np.linalg.norm(target - eye)
    right = np.cross(fwd, [0, 1, 0]); right /= np.linalg.norm(right)
    up = np.cross(right, fwd); up /= np.linalg.norm(up)
    lens = 0.72 / min(max(zoom, 0.5), 3.0)
    return eye, fwd, right, up, lens, width / height, e


def _chord(eye, d, e):
    with np.errstate(divide="ignore"):
        inv = 1.0 / np.where(np.abs(d) > 1e-6, d, 1e-6)
    lo = np.minimum(-eye * inv, (e - eye) * inv).max()
    hi = np.maximum(-eye * inv, (e - eye) * inv).min()
    return (hi - max(lo, 0.0)) if hi > lo and hi > 0 else None


def test_uniform_medium_follows_beer_lambert():
    """Temperature view of a uniform 560 K field: strength a and colour c
    from the packed texel; each pixel is c 1.35 (1 - e) + e bg with
    e = exp(-8.5 a L), bg the same view of an empty domain."""
    dims = (24, 32, 24)
    w, h = 48, 27
    o = _oracle(dims)
    o.set_parameter("viewMode", 1)
    empty = o.render_volume(w, h, 384)  # T = 300 K everywhere: strength 0, the background alone
    nx, ny, nz = dims
    flow = np.zeros((nz, ny, nx, 4))
    flow[..., 3] = 560.0
    o.write(orc.FIELD_FLOW, flow)
    img = o.render_volume(w, h, 384)
    v = float(np.float16(np.float32((560.0 - 300.0) / 2600.0)))
    sm = lambda lo, hi, x: (lambda u: u * u * (3 - 2 * u))(min(max((x - lo) / (hi - lo), 0), 1))
    a = sm(0.015, 0.72, v)
    mix = lambda p, q, s: np.array(p) * (1 - s) + np.array(q) * s
    # thermalMap(v), qualityShaders.ts:311-319
    c = mix(mix([0.015, 0.025, 0.12], [0.02, 0.36, 1.0], sm(0.0, 0.30, v)),
            mix([1.0, 0.34, 0.01], [1.0, 0.96, 0.76], sm(0.62, 1.0, v)), sm(0.24, 0.72, v))
    eye, fwd, right, up, lens, aspect, e = _camera(w, h)
    checked = missed = 0
    for py in range(h):
        for px in range(w):
            sx = (px + 0.5) / w * 2 - 1
            sy = (1 - (py + 0.5) / h) * 2 - 1
            d = fwd + right * sx * aspect * lens + up * sy * lens
            d /= np.linalg.norm(d)
            L = _chord(eye, d, e)
            if L is None:
                np.testing.assert_array_equal(img[py, px], empty[py, px])  # misses the box: untouched
                missed += 1
                continue
            trans = math.exp(-8.5 * a * L)
            want = c * 1.35 * (1 - trans) + trans * empty[py, px, :3]
            np.testing.assert_allclose(img[py, px, :3], want, rtol=3e-3, atol=1e-4)
            checked += 1
    assert checked > 300 and missed > 50
    o.close()


def test_non_cinematic_path_tone_maps_the_background():
    o = _oracle(cinematic=False)
    o.set_parameter("exposure", 1.6)
    w, h = 32, 18
    img = o.render_volume(w, h, 80)
    for row in (0, 2):
        uvy = 1.0 - (row + 0.5) / h
        x = np.maximum(_background_top(uvy).astype(np.float64) * 1.6, 0)
        peak = x.max()
        curve = min(max((peak * (2.51 * peak + 0.03)) / (peak * (2.43 * peak + 0.59) + 0.14), 0), 1)
        want = np.clip(x * (curve / max(peak, 1e-6)), 0, 1) ** (1 / 2.2)
        np.testing.assert_allclose(img[row, 0, :3], want, rtol=2e-3)
    o.close()
