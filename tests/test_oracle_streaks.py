"""Pins of the oracle's droplet streaks (dropletVertex / dropletFragment,
qualityShaders.ts:156-232), drawn over the scene image (cinematic) or the
rgba8 canvas (standard), by what the geometry and the blend fix:

* with the camera on the z axis (yaw 0, pitch 0) a droplet at the camera
  target projects to the image centre, and a velocity along x or y draws a
  quad centred there, radius tall (across the motion) and radius +
  min(|streak|, 0.032) / 2 long (along it), |streak| = 0.018 |v| / (depth *
  aspect * lens) (x) or 0.018 |v| / (depth * lens) (y), in closed form; the
  pixels it changes are those whose capsule coverage clears 0.001;
* at the quad centre local = (0, 0): coverage 1, highlight 0.7^5, alpha
  clamp(0.13 + 0.15 mass) (x 1.38 in the liquid view), blended src-alpha
  over the volume pixel (stored as f16, or as unorm8 in the standard canvas);
* two droplets over one pixel blend in instance order;
* droplets that are massless, outside the unit box, or drawn in another
  view mode change nothing.
"""
import numpy as np
import pytest

from oracle import oracle as orc

W, H = 161, 91  # odd: the centre pixel's NDC centre is (0, 0) exactly


def make(count, cinematic=True, view=0):
    o = orc.Oracle((16, 24, 16), precision=32, pressure_iterations=2, correction_iterations=2)
    o.set_bool("cinematic", cinematic)
    o.set_heavy_mode(orc.HEAVY_OFF)
    o.set_particle_count(count)
    o.set_parameter("cameraYaw", 0.0)
    o.set_parameter("cameraPitch", 0.0)
    o.set_parameter("viewMode", view)
    return o


def extent(o):
    return np.array([o.get_parameter(k) for k in ("domainWidth", "domainHeight", "domainDepth")])


def droplets(o, rows):
    """rows of (normalized position, velocity, mass) -> the particle buffer."""
    a = np.zeros((o.particle_count, 8))
    a[:, 7] = 0.0  # the rest are massless
    for i, (pos, vel, mass) in enumerate(rows):
        a[i, 0:3] = pos
        a[i, 4:7] = vel
        a[i, 7] = mass
    return a


def render(o, rows, w=W, h=H, frame=False):
    buf = droplets(o, rows)
    o.write_particle_buffers(buf, buf)
    return o.render_frame(w, h, 24) if frame else o.render_volume(w, h, 24)


def camera(o, w, h):
    """depth of the camera target and lens, camera on +z looking down -z."""
    e32 = extent(o).astype(np.float32).astype(np.float64)
    frame = e32.max()
    depth = (0.90 + o.get_parameter("cameraDistance") * 0.18) * frame
    lens = 0.72 / np.clip(o.get_parameter("viewZoom"), 0.5, 3.0)
    return depth, lens, w / h


def target_fraction(o):
    return np.array([0.5, 0.4, 0.5])


def pixel_ndc(w, h):
    x = (np.arange(w) + 0.5) / w * 2.0 - 1.0
    y = (1.0 - (np.arange(h) + 0.5) / h) * 2.0 - 1.0
    return np.meshgrid(x, y)  # (h, w) each


def d_cut():
    """capsule distance where 1 - smoothstep(0.64, 1, d) = 0.001."""
    t = np.linspace(0.9, 1.0, 200001)
    s = t * t * (3.0 - 2.0 * t)
    return 0.64 + 0.36 * t[np.argmin(np.abs(s - 0.999))]


def pixel_radius(o):
    return 1.15 if o.get_parameter("nozzleType") >= 0.5 else 1.65


@pytest.mark.parametrize("axis,speed", [(0, 0.0), (0, 0.9), (0, 1.6), (0, 4.0), (1, 1.0), (1, -2.5)])
def test_streak_footprint_closed_form(axis, speed):
    w, h = 641, 361  # 0.005 NDC is 1.6 pixels across
    o = make(4)
    bg = render(o, [], w, h)
    vel = np.zeros(3)
    vel[axis] = speed
    img = render(o, [(target_fraction(o), vel, 1.2)], w, h)
    changed = np.any(img[..., :3] != bg[..., :3], axis=-1)
    depth, lens, aspect = camera(o, w, h)
    radius = pixel_radius(o) * 2.0 / h
    streak = 0.018 * abs(speed) / (depth * (aspect if axis == 0 else 1.0) * lens)
    half = radius + min(streak, 0.032) * 0.5
    px, py = pixel_ndc(w, h)
    along, across = (px, py) if axis == 0 else (py, px)
    lx, ly = across / radius, along / half
    inside = (np.abs(lx) < 1.0) & (np.abs(ly) < 1.0)
    dist = np.sqrt(lx ** 2 + np.maximum(np.abs(ly) - 0.48, 0.0) ** 2)
    want = inside & (dist < d_cut())
    edge = (np.abs(dist - d_cut()) < 0.02) | (np.abs(np.abs(lx) - 1.0) < 1e-3) | (np.abs(np.abs(ly) - 1.0) < 1e-3)
    assert want.sum() >= 2
    assert not np.any((changed != want) & ~edge), f"{np.argwhere(changed != want)[:4]}"
    if streak > 4.0 / w:
        # the streak is longer along the motion than across it (in NDC units)
        ys, xs = np.nonzero(changed)
        sx, sy = np.ptp(xs) * 2.0 / w, np.ptp(ys) * 2.0 / h
        assert (sx > sy) if axis == 0 else (sy > sx)
    o.close()


def fragment_at_centre(mass, view):
    highlight = np.float32(0.7) ** 5
    c0, c1 = np.array([0.27, 0.34, 0.38]), np.array([0.76, 0.86, 0.91])
    base = c0 + (c1 - c0) * highlight * 0.62
    alpha = np.clip(0.13 + mass * 0.15, 0.12, 0.42) * (1.38 if view == 6 else 1.0)
    return base * 0.72, alpha


def aces(c, exposure):
    x = np.maximum(c * max(exposure, 0.25), 0.0)
    peak = x.max()
    mapped = np.clip(peak * (2.51 * peak + 0.03) / (peak * (2.43 * peak + 0.59) + 0.14), 0.0, 1.0)
    return np.clip(x * (mapped / max(peak, 1e-6)), 0.0, 1.0) ** (1.0 / 2.2)


@pytest.mark.parametrize("mass,view", [(0.3, 0), (1.9, 0), (0.9, 6)])
def test_quad_centre_blend_closed_form(mass, view):
    o = make(4, view=view)
    bg = render(o, [])
    img = render(o, [(target_fraction(o), np.zeros(3), mass)])
    src, a = fragment_at_centre(mass, view)
    cy, cx = H // 2, W // 2
    want = src * a + bg[cy, cx, :3] * (1.0 - a)
    np.testing.assert_allclose(img[cy, cx, :3], want, rtol=2e-3)
    assert img[cy, cx, 3] == 1.0
    o.close()


def test_standard_canvas_blend_closed_form():
    o = make(4, cinematic=False)
    o.set_parameter("exposure", 1.4)
    bg = render(o, [], frame=True)
    img = render(o, [(target_fraction(o), np.zeros(3), 0.8)], frame=True)
    src, a = fragment_at_centre(0.8, 0)
    cy, cx = H // 2, W // 2
    want = (aces(src, 1.4) * a + bg[cy, cx, :3] / 255.0 * (1.0 - a)) * 255.0
    assert np.abs(img[cy, cx, :3].astype(np.float64) - want).max() <= 0.51
    assert img[cy, cx, 3] == 255
    # the scene image of the standard path carries no droplets
    np.testing.assert_array_equal(render(o, [(target_fraction(o), np.zeros(3), 0.8)]), render(o, []))
    o.close()


def test_instance_order():
    """Droplet B sits half a quad radius above A, so their colours differ at
    the centre pixel; the pixel follows blend(blend(bg, first), second)."""
    o = make(4)
    depth, lens, _ = camera(o, W, H)
    radius = pixel_radius(o) * 2.0 / H
    e = extent(o)
    shift = 0.5 * radius * depth * lens  # world y giving an NDC offset of radius / 2
    pa = target_fraction(o)
    pb = pa + np.array([0.0, shift / e[1], 0.0])
    bg = render(o, [])
    cy, cx = H // 2, W // 2
    ab = render(o, [(pa, np.zeros(3), 0.4), (pb, np.zeros(3), 2.0)])[cy, cx, :3]
    ba = render(o, [(pb, np.zeros(3), 2.0), (pa, np.zeros(3), 0.4)])[cy, cx, :3]

    def one(lx, mass):
        highlight = np.clip(1.0 - abs(lx + 0.30), 0.0, 1.0) ** 5
        c0, c1 = np.array([0.27, 0.34, 0.38]), np.array([0.76, 0.86, 0.91])
        return (c0 + (c1 - c0) * highlight * 0.62) * 0.72, np.clip(0.13 + mass * 0.15, 0.12, 0.42)

    sa, aa = one(0.0, 0.4)
    sb, ab_alpha = one(-0.5, 2.0)  # the pixel is half a radius below B's centre: local.x = -0.5
    f16 = lambda v: v.astype(np.float16).astype(np.float64)
    d = bg[cy, cx, :3]
    want_ab = f16(sb * ab_alpha + f16(sa * aa + d * (1 - aa)) * (1 - ab_alpha))
    want_ba = f16(sa * aa + f16(sb * ab_alpha + d * (1 - ab_alpha)) * (1 - aa))
    np.testing.assert_allclose(ab, want_ab, rtol=2e-3)
    np.testing.assert_allclose(ba, want_ba, rtol=2e-3)
    assert np.abs(want_ab - want_ba).max() > 2e-3  # several times the tolerance above
    o.close()


@pytest.mark.parametrize("case", ["massless", "outside", "view3", "view7"])
def test_hidden_droplets_change_nothing(case):
    view = {"view3": 3, "view7": 7}.get(case, 0)
    o = make(4, view=view)
    bg = render(o, [])
    pos, mass = target_fraction(o), 1.0
    if case == "massless":
        mass = 0.0001
    if case == "outside":
        pos = np.array([0.5, 1.0001, 0.5])
    img = render(o, [(pos, np.array([0.3, 0.2, 0.0]), mass)])
    np.testing.assert_array_equal(img, bg)
    o.close()
