"""Pins the oracle's restatement of the reference step with facts that do not
come from retyping its formulas: exact special cases of the interpolation,
closed forms that remain when most terms vanish, conservation and
stoichiometric invariants, the discrete operator identity D(G p) = L p, a
dense direct solve of the Jacobi system, boundary-condition values, and the
pass order of the step graph (FireXEngine.ts:2115-2276)."""
import numpy as np
import pytest

from oracle import oracle as orc

DT = 1.0 / 60.0
F32 = lambda v: float(np.float32(v))  # uniforms reach the kernels as f32


def uniform_state(o, flow4, spec4):
    nx, ny, nz = o.dims
    o.write(orc.FIELD_FLOW, np.broadcast_to(np.array(flow4, float), (nz, ny, nx, 4)))
    o.write(orc.FIELD_SPECIES, np.broadcast_to(np.array(spec4, float), (nz, ny, nx, 4)))


def interior_rows(o, lo=0.30, hi=0.80):
    nx, ny, nz = o.dims
    ys = [y for y in range(1, ny - 1) if lo <= (y + 0.5) / ny <= hi]
    return slice(1, nz - 1), ys, slice(1, nx - 1)


# ------------------------------------------------------------- interpolation
def test_trilinear_cell_centres_are_exact():
    o = orc.Oracle((7, 9, 6), precision=64)
    rng = np.random.default_rng(0)
    f = rng.normal(size=(6, 9, 7, 4))
    o.write(orc.FIELD_FLOW, f)
    for (x, y, z) in [(0, 0, 0), (3, 4, 2), (5, 7, 4), (2, 7, 1)]:
        np.testing.assert_array_equal(o.trilinear_flow(DT, (x + 0.5, y + 0.5, z + 0.5)), f[z, y, x])


def test_trilinear_reproduces_affine_fields_and_edge_clamp():
    o = orc.Oracle((8, 10, 6), precision=64)
    z, y, x = np.meshgrid(np.arange(6), np.arange(10), np.arange(8), indexing="ij")
    coef = np.array([[1.0, 2.0, 3.0, -0.5], [-1, 0.25, 0.5, 2], [0.3, -1, 1, 1], [2, 1, 1, 1]])
    f = np.stack([c[0] + c[1] * x + c[2] * y + c[3] * z for c in coef], axis=-1)
    o.write(orc.FIELD_FLOW, f)
    rng = np.random.default_rng(1)
    for _ in range(50):
        p = rng.uniform([0.5, 0.5, 0.5], [7.49, 9.49, 5.49])
        s = p - 0.5
        want = [c[0] + c[1] * s[0] + c[2] * s[1] + c[3] * s[2] for c in coef]
        np.testing.assert_allclose(o.trilinear_flow(DT, p), want, rtol=0, atol=1e-11)
    # beyond the last centre the sample is clamped to dims - 1.001
    # (shaders.ts:63): weights 0.001 / 0.999 on the last two cells
    v = o.trilinear_flow(DT, (20.0, 0.5, 0.5))
    want = 0.001 * f[0, 0, 6] + 0.999 * f[0, 0, 7]
    np.testing.assert_allclose(v, want, rtol=0, atol=1e-9)


# -------------------------------------------------------- source-pass closed forms
def test_quiescent_cell_closed_forms():
    """Zero velocity, uniform 500 K, no fuel, ambient O2: only buoyancy,
    radiation and the product/vapor decays act (away from walls, the
    burner and the outflow band)."""
    o = orc.Oracle((16, 32, 16), precision=64)
    T0, p0, v0 = 500.0, 0.5, 0.3
    uniform_state(o, (0, 0, 0, T0), (0, 1.0, p0, v0))
    o.kernel_simulate(DT)
    flow = o.read(orc.FIELD_FLOW_PREDICTED)
    spec = o.read(orc.FIELD_SPECIES)
    zs, ys, xs = interior_rows(o)
    dt = F32(DT)
    q = F32(1.45)
    rad = 5.67e-8 * (300.0 ** 4 - T0 ** 4) / (1005 * q)
    buoy = (T0 - 300) / 1500 * dt * 0.72 * F32(2.15)
    for y in ys:
        np.testing.assert_allclose(flow[zs, y, xs, 3], T0 + dt * rad, rtol=1e-12)
        np.testing.assert_allclose(flow[zs, y, xs, 1], buoy, rtol=1e-6)
        np.testing.assert_allclose(flow[zs, y, xs, 0], 0, atol=1e-9)
        np.testing.assert_allclose(spec[zs, y, xs, 2], p0 * np.exp(-dt * 0.075 / q), rtol=1e-12)
        np.testing.assert_allclose(spec[zs, y, xs, 3], v0 * np.exp(-dt * 0.14 / q), rtol=1e-12)
        np.testing.assert_allclose(spec[zs, y, xs, 1], 1.0, rtol=1e-12)
    assert np.all(o.read(orc.FIELD_REACTION)[zs, ys, xs] == 0.0)


def test_burning_cell_stoichiometry_and_heat_release():
    """Uniform hot mixture: fuel and O2 fall in the 1 : 4 ratio
    (shaders.ts:249-250), burn never exceeds min(fuel, O2/4), and products,
    vapor, heat release and the reaction diagnostic follow the burn."""
    o = orc.Oracle((16, 32, 16), precision=64)
    o.set_parameter("airEntrainment", 0.0)
    T0, f0, ox0 = 1400.0, 0.2, 0.6
    uniform_state(o, (0, 0, 0, T0), (f0, ox0, 0.0, 0.0))
    o.kernel_simulate(DT)
    flow = o.read(orc.FIELD_FLOW_PREDICTED)
    spec = o.read(orc.FIELD_SPECIES)
    react = o.read(orc.FIELD_REACTION)
    zs, ys, xs = interior_rows(o)
    dt, q, eta = F32(DT), F32(1.45), F32(0.86)
    for y in ys:
        burn = f0 - spec[zs, y, xs, 0]
        assert np.all(burn > 0)
        assert np.all(burn <= min(f0, ox0 / 4) + 1e-15)
        np.testing.assert_allclose(ox0 - spec[zs, y, xs, 1], 4 * burn, rtol=1e-10)
        rich = 1 - ox0 / (4 * f0)
        np.testing.assert_allclose(spec[zs, y, xs, 2], burn * (1 + 1.8 * rich) * np.exp(-dt * 0.075 / q), rtol=1e-10)
        np.testing.assert_allclose(spec[zs, y, xs, 3], burn * 0.62 * np.exp(-dt * 0.14 / q), rtol=1e-10)
        np.testing.assert_allclose(react[zs, y, xs], np.minimum(burn / dt * 0.22, 4.0), rtol=1e-10)
        tm = T0 + burn * (2150 + 1450 * eta)
        np.testing.assert_allclose(flow[zs, y, xs, 3], tm + dt * 5.67e-8 * (300.0 ** 4 - tm ** 4) / (1005 * q),
                                   rtol=1e-10)


def test_seeded_gaussian_source_shape():
    """With no jitter the injected fuel is fuelRate * firePower * Gaussian
    (shaders.ts:193 with the seeded source in place of emitter())."""
    dims = (24, 32, 24)
    def oxygen_free(jitter):
        # no oxygen anywhere: nothing burns, so the injected fuel stays
        o = orc.Oracle(dims, precision=64)
        o.set_parameter("oxygenRate", 0.0)
        o.set_parameter("airEntrainment", 0.0)
        o.set_emitters(em, jitter, 0.0, 5)
        uniform_state(o, (0, 0, 0, 300.0), (0, 0.0, 0, 0))
        return o

    em = [(12.0, 6.0, 12.0, 4.0 * np.sqrt(2), 1.0)]
    o = oxygen_free(0.0)
    o.kernel_simulate(DT)
    fuel = o.read(orc.FIELD_SPECIES)[..., 0]
    z, y, x = np.meshgrid(np.arange(24), np.arange(32), np.arange(24), indexing="ij")
    r = np.float32(em[0][3])
    g = np.exp(-(((x + 0.5) - 12) ** 2 + ((y + 0.5) - 6) ** 2 + ((z + 0.5) - 12) ** 2) / float(r) ** 2)
    inner = (slice(2, 22), slice(1, 20), slice(2, 22))
    np.testing.assert_allclose(fuel[inner], F32(0.3) * F32(2.5) * g[inner], rtol=1e-6, atol=1e-14)
    # with jitter the ratio stays within +-jitter and averages to one
    o2 = oxygen_free(0.1)
    o2.kernel_simulate(DT)
    fj = o2.read(orc.FIELD_SPECIES)[..., 0]
    mask = g[inner] > 1e-3
    ratio = fj[inner][mask] / fuel[inner][mask]
    assert ratio.min() >= 0.9 - 1e-6 and ratio.max() <= 1.1 + 1e-6
    assert abs(ratio.mean() - 1) < 0.01


def test_reference_emitter_vertical_separability_and_bounds():
    """emitter() = exp(-(radial + vertical)) * texture(x, z, t) * pulse(x, z, t):
    at fixed (x, z) the ratio of two heights is the Gaussian in y alone, with
    the burner at 0.075 m and thickness 0.0022 * clamp(sourceThickness)^2."""
    o = orc.Oracle((32, 48, 32), precision=64)
    th = F32(1.35)
    h = F32(1.35)  # domain height (m)
    for (nx_, nz_) in [(0.5, 0.5), (0.45, 0.6), (0.3, 0.52)]:
        y1, y2 = 0.04, 0.10
        a = o.emitter(DT, (nx_, y1, nz_))
        b = o.emitter(DT, (nx_, y2, nz_))
        # shader constants (0.075, 0.0022) are exact in the FP64 oracle;
        # uniforms (domain, thickness) arrive as f32
        v1 = (y1 * h - 0.075) ** 2 / (0.0022 * th * th)
        v2 = (y2 * h - 0.075) ** 2 / (0.0022 * th * th)
        assert b / a == pytest.approx(np.exp(-(v2 - v1)), rel=1e-9)
    vals = [o.emitter(DT, (x, y, z)) for x in np.linspace(0, 1, 9) for y in np.linspace(0, 0.3, 7)
            for z in np.linspace(0, 1, 9)]
    assert min(vals) >= 0 and max(vals) <= 1.18 * 1.0 + 1e-12


def _emitter_gaussian(X, Y, Z, W, H, D, t=0.0):
    """exp(-(radial + vertical)) of emitter() (shaders.ts:101-117) at time t,
    Inferno tuning (f32 uniforms), FP64 shader constants."""
    rb = np.sqrt(0.02024)
    rx = min(rb * min(max(F32(3.1), 0.35), 3.2), W * 0.42)
    rz = min(rb * min(max(F32(2.7), 0.35), 3.2), D * 0.42)
    th = min(max(F32(1.35), 0.35), 2.5)
    ox = X - (W * 0.5 + np.sin(t * 1.73) * 0.013)
    oy, oz = Y - 0.075, Z - (D * 0.5 + np.cos(t * 1.37) * 0.011)
    return np.exp(-(ox * ox / (rx * rx) + oz * oz / (rz * rz) + oy * oy / (0.0022 * th * th))), ox / rx, oz / rz


def test_reference_emitter_horizontal_texture_on_nodal_lines():
    """The horizontal factors of emitter() at time 0 (sin 0 = 0, cos 0 = 1):
    cellular = 0.76 + 0.24 sin(49 X) sin(43 Z) and pulse = 0.84 + 0.16 sin(27 X + 21 Z)
    (X, Z in metres).  On the pulse's nodal lines 27 X + 21 Z = m pi the pulse is
    exactly 0.84, so emitter / (0.84 * Gaussian) must be the cellular closed form
    along the line; where sin(49 X) = 0 as well it is exactly 0.76.  With the
    cinematic flag, on the burner's centre line (offset.x = 0) the lobe argument
    is 0 and the lobe factor is the closed form 0.48 + 0.70 * smoothstep(-0.42,
    0.48, 0).  A wrong frequency, phase, amplitude or radius fails these."""
    W, H, D = F32(1.15), F32(1.35), F32(1.15)
    o = orc.Oracle((32, 48, 32), precision=64)
    o.set_bool("cinematic", False)
    rng = np.random.default_rng(12)
    checked = 0
    for m in range(3, 12):
        for X in rng.uniform(0.25, 0.9, size=6):
            Z = (m * np.pi - 27 * X) / 21
            if not 0.25 < Z < 0.9:
                continue
            Y = rng.uniform(0.04, 0.11)
            g, _, _ = _emitter_gaussian(X, Y, Z, W, H, D)
            e = o.emitter(0.0, (X / W, Y / H, Z / D))
            cellular = 0.76 + 0.24 * np.sin(49 * X) * np.sin(43 * Z)
            assert e == pytest.approx(0.84 * cellular * g, rel=1e-9, abs=1e-300)
            checked += 1
    assert checked >= 10
    for k in range(13, 15):  # sin(49 X) = 0 at X = k pi / 49 (0.83..0.90 m)
        X = k * np.pi / 49
        Z = (4 * np.pi - 27 * X) / 21  # on the pulse's nodal line m = 4
        Z = Z if 0 < Z < D else (5 * np.pi - 27 * X) / 21
        g, _, _ = _emitter_gaussian(X, 0.075, Z, W, H, D)
        assert o.emitter(0.0, (X / W, 0.075 / H, Z / D)) == pytest.approx(0.76 * 0.84 * g, rel=1e-9)
    o.set_bool("cinematic", True)
    s = (0.42 / 0.90) ** 2 * (3 - 2 * (0.42 / 0.90))
    lobes = 0.48 + (1.18 - 0.48) * s
    for m in range(5, 9):
        X = W * 0.5
        Z = (m * np.pi - 27 * X) / 21
        if not 0.2 < Z < 0.95:
            continue
        g, ax, _ = _emitter_gaussian(X, 0.08, Z, W, H, D)
        assert abs(ax) < 1e-12
        cellular = 0.76 + 0.24 * np.sin(49 * X) * np.sin(43 * Z)
        assert o.emitter(0.0, (X / W, 0.08 / H, Z / D)) == pytest.approx(0.84 * cellular * lobes * g, rel=1e-9)
    # Anywhere, the cinematic / plain ratio cancels the Gaussian, cellular and
    # pulse factors and leaves the lobe factor alone.
    plain = orc.Oracle((32, 48, 32), precision=64)
    plain.set_bool("cinematic", False)
    for _ in range(40):
        X, Y, Z = rng.uniform(0.3, 0.85), rng.uniform(0.04, 0.11), rng.uniform(0.3, 0.85)
        _, ax, az = _emitter_gaussian(X, Y, Z, W, H, D)
        a = np.sin(ax * 6.2) * np.sin(az * 5.4)
        t = min(max((a + 0.42) / 0.90, 0.0), 1.0)
        want = 0.48 + (1.18 - 0.48) * t * t * (3 - 2 * t)
        pos = (X / W, Y / H, Z / D)
        assert o.emitter(0.0, pos) / plain.emitter(0.0, pos) == pytest.approx(want, rel=1e-9)
    # After one step the emitter reads time t = dt: the burner centre moves by
    # (0.013 sin 1.73t, 0, 0.011 cos 1.37t) and the lobes by +1.7t and -1.3t.
    o.step(DT, 1)
    plain.step(DT, 1)
    t = F32(DT)
    rb = np.sqrt(0.02024)
    rx = min(rb * F32(3.1), W * 0.42)
    rz = min(rb * F32(2.7), D * 0.42)
    for _ in range(40):
        X, Y, Z = rng.uniform(0.3, 0.85), rng.uniform(0.04, 0.11), rng.uniform(0.3, 0.85)
        ax = (X - (W * 0.5 + np.sin(t * 1.73) * 0.013)) / rx
        az = (Z - (D * 0.5 + np.cos(t * 1.37) * 0.011)) / rz
        a = np.sin(ax * 6.2 + t * 1.7) * np.sin(az * 5.4 - t * 1.3)
        u = min(max((a + 0.42) / 0.90, 0.0), 1.0)
        want = 0.48 + (1.18 - 0.48) * u * u * (3 - 2 * u)
        pos = (X / W, Y / H, Z / D)
        assert o.emitter(0.0, pos) / plain.emitter(0.0, pos) == pytest.approx(want, rel=1e-9)
    # ... and on the pulse's nodal lines 3.7t + 27X + 21Z = m pi the plain
    # emitter is 0.84 * cellular(t) * Gaussian(t).
    for m in range(4, 10):
        X = rng.uniform(0.3, 0.85)
        Z = (m * np.pi - 27 * X - 3.7 * t) / 21
        if not 0.25 < Z < 0.9:
            continue
        g, _, _ = _emitter_gaussian(X, 0.07, Z, W, H, D, t)
        cellular = 0.76 + 0.24 * np.sin(49 * X + 5.2 * t) * np.sin(43 * Z - 4.4 * t)
        assert plain.emitter(0.0, (X / W, 0.07 / H, Z / D)) == pytest.approx(0.84 * cellular * g, rel=1e-9)
    plain.close()


def test_clamps_and_boundary_values_after_simulate():
    dims = (12, 16, 10)
    o = orc.Oracle(dims, precision=64)
    rng = np.random.default_rng(4)
    flow = np.empty((10, 16, 12, 4))
    flow[..., :3] = rng.uniform(-5, 5, size=(10, 16, 12, 3))
    flow[..., 3] = rng.uniform(0, 5000, size=(10, 16, 12))
    o.write(orc.FIELD_FLOW, flow)
    o.write(orc.FIELD_SPECIES, rng.uniform(-1, 6, size=(10, 16, 12, 4)))
    o.kernel_simulate(DT)
    f = o.read(orc.FIELD_FLOW_PREDICTED)
    s = o.read(orc.FIELD_SPECIES)
    r = o.read(orc.FIELD_REACTION)
    assert np.abs(f[..., :3]).max() <= 1.4
    assert f[..., 3].min() >= 300 and f[..., 3].max() <= 2900
    assert s.min() >= 0 and s.max() <= 3
    assert np.all(f[:, :, 0, 0] == 0) and np.all(f[:, :, -1, 0] == 0)
    assert np.all(f[0, :, :, 2] == 0) and np.all(f[-1, :, :, 2] == 0)
    assert np.all(f[:, 0, :, 1] == 0) and np.all(f[:, -1, :, 1] >= 0)
    top = s[:, -1]
    assert np.all(top[..., 0] == 0) and np.all(top[..., 2] == 0) and np.all(top[..., 3] == 0)
    assert np.all(f[:, -1, :, 3] == 300) and np.all(r[:, -1] == 0)


# ------------------------------------------------------- projection / pressure
def test_divergence_of_gradient_is_the_jacobi_laplacian():
    dims = (10, 12, 9)
    o = orc.Oracle(dims, precision=64)
    wx, wy, wz = o.stencil_weights(DT)
    rng = np.random.default_rng(5)
    p = rng.normal(size=(9, 12, 10))
    p[:, -1] = 0.0
    o.write(orc.FIELD_FLOW, np.zeros((9, 12, 10, 4)))
    o.write(orc.FIELD_PRESSURE, p)
    o.kernel_project(DT, src=0)  # flow[1] = -grad p (with BCs)
    o.kernel_divergence(DT, from_predicted=True)
    d = o.read(orc.FIELD_DIVERGENCE)
    c = (slice(2, 7), slice(2, 9), slice(2, 8))  # interior (z, y, x)
    lap = (wx * (np.roll(p, -1, 2) - 2 * p + np.roll(p, 1, 2))
           + wy * (np.roll(p, -1, 1) - 2 * p + np.roll(p, 1, 1))
           + wz * (np.roll(p, -1, 0) - 2 * p + np.roll(p, 1, 0)))
    np.testing.assert_allclose(d[c], -lap[c], rtol=1e-10, atol=1e-9 * np.abs(lap[c]).max())


def test_projection_boundary_values():
    dims = (10, 12, 9)
    o = orc.Oracle(dims, precision=64)
    rng = np.random.default_rng(6)
    flow = rng.normal(size=(9, 12, 10, 4))
    p = rng.normal(size=(9, 12, 10))
    o.write(orc.FIELD_FLOW, flow)
    o.write(orc.FIELD_PRESSURE, p)
    o.kernel_project(DT, src=0)
    f = o.read(orc.FIELD_FLOW_PREDICTED)
    assert np.all(f[:, :, 0, 0] == 0) and np.all(f[:, :, -1, 0] == 0)
    assert np.all(f[0, :, :, 2] == 0) and np.all(f[-1, :, :, 2] == 0)
    assert np.all(f[:, 0, :, 1] == 0)
    # open top: zero-gradient copy of the projected cell below, u_y >= 0,
    # temperature kept (shaders.ts:481-487, 497-499)
    inner = (slice(1, -1), slice(1, -1))
    np.testing.assert_array_equal(f[1:-1, -1, 1:-1, 0], f[1:-1, -2, 1:-1, 0])
    np.testing.assert_array_equal(f[1:-1, -1, 1:-1, 2], f[1:-1, -2, 1:-1, 2])
    np.testing.assert_array_equal(f[:, -1, :, 1], np.maximum(0, f[:, -2, :, 1]))
    np.testing.assert_array_equal(f[..., 3], flow[..., 3])
    del inner


def _dense_solve(dims, w, div):
    nx, ny, nz = dims
    idx = {}
    for z in range(nz):
        for y in range(ny - 1):
            for x in range(nx):
                idx[(x, y, z)] = len(idx)
    n = len(idx)
    a = np.zeros((n, n))
    b = np.zeros(n)
    for (x, y, z), i in idx.items():
        a[i, i] += 2 * sum(w)
        b[i] = -div[z, y, x]
        for axis, (wa) in enumerate(w):
            for d in (-1, 1):
                c = [x, y, z]
                c[axis] += d
                if axis == 1 and c[1] >= ny - 1:
                    continue  # top row is the fixed zero (Dirichlet)
                c[0] = min(max(c[0], 0), nx - 1)
                c[1] = max(c[1], 0)
                c[2] = min(max(c[2], 0), nz - 1)
                a[i, idx[tuple(c)]] -= wa
    sol = np.linalg.solve(a, b)
    p = np.zeros((nz, ny, nx))
    for (x, y, z), i in idx.items():
        p[z, y, x] = sol[i]
    return p


def test_jacobi_fixed_point_matches_dense_direct_solve():
    dims = (5, 7, 4)
    o = orc.Oracle(dims, precision=64)
    w = o.stencil_weights(DT)
    rng = np.random.default_rng(7)
    div = rng.normal(size=(4, 7, 5)).astype(np.float32).astype(np.float64)
    o.write(orc.FIELD_DIVERGENCE, div)
    for _ in range(6000):
        o.kernel_jacobi(DT)
    p = o.read(orc.FIELD_PRESSURE)
    want = _dense_solve(dims, w, div)
    np.testing.assert_allclose(p, want, rtol=1e-8, atol=1e-10 * np.abs(want).max())
    assert np.all(p[:, -1] == 0)


def test_zero_is_a_fixed_point_and_f16_storage_rounds_like_numpy():
    dims = (8, 10, 6)
    o = orc.Oracle(dims, precision=32)
    o.kernel_jacobi(DT)
    assert np.all(o.read(orc.FIELD_PRESSURE) == 0)
    rng = np.random.default_rng(8)
    div = rng.normal(scale=50, size=(6, 10, 8)).astype(np.float32).astype(np.float64)
    p0 = rng.normal(size=(6, 10, 8)).astype(np.float16).astype(np.float64)
    o.write(orc.FIELD_DIVERGENCE, div)
    o.write(orc.FIELD_PRESSURE, p0)
    o.kernel_jacobi(DT, f16=False)
    p32 = o.read(orc.FIELD_PRESSURE).astype(np.float32)
    o2 = orc.Oracle(dims, precision=32)
    o2.write(orc.FIELD_DIVERGENCE, div)
    o2.write(orc.FIELD_PRESSURE, p0)
    o2.kernel_jacobi(DT, f16=True)
    p16 = o2.read(orc.FIELD_PRESSURE).astype(np.float32)
    np.testing.assert_array_equal(p16, p32.astype(np.float16).astype(np.float32))


def _fma32(a, b, c):
    """f32 fused multiply-add from float64: a*b of two f32 values is exact in
    float64 (48 significant bits), the add rounds once to float64 and the cast
    once more to f32.  The second rounding can only differ from a single one
    when the float64 sum lands on an f32 rounding midpoint (about 2^-29 per
    value); the seeded data below has no such case."""
    return (a.astype(np.float64) * b.astype(np.float64) + c.astype(np.float64)).astype(np.float32)


def test_jacobi_f32_sweep_is_the_contracted_form():
    """One f32 sweep equals an independent numpy restatement of solvePressureGlobal
    (shaders.ts:375-403): replicated walls and floor, zero above the open top, top
    row forced to 0, the weighted sum contracted as fma(wz, sz, fma(wy, sy, wx sx)),
    then (nb - div) times the f32 reciprocal of 2 (wx + wy + wz).  The uncontracted
    sum and the other contraction order both differ from it, so either mistake fails."""
    nx, ny, nz = 8, 10, 6
    o = orc.Oracle((nx, ny, nz), precision=32)
    rng = np.random.default_rng(11)
    p = rng.normal(size=(nz, ny, nx)).astype(np.float32)
    div = rng.normal(scale=50, size=(nz, ny, nx)).astype(np.float32)
    o.write(orc.FIELD_DIVERGENCE, div.astype(np.float64))
    o.write(orc.FIELD_PRESSURE, p.astype(np.float64))
    o.kernel_jacobi(DT, f16=False)
    got = o.read(orc.FIELD_PRESSURE).astype(np.float32)

    f = np.float32
    mx, my, mz = f(nx) / f(1.15), f(ny) / f(1.35), f(nz) / f(1.15)
    wx, wy, wz = mx * mx, my * my, mz * mz
    inv = f(1) / (f(2) * ((wx + wy) + wz))
    pe = np.pad(p, ((1, 1), (1, 1), (1, 1)), mode="edge")  # clamped reads
    pe[:, -1, :] = 0  # y = ny reads 0 (open top)
    c = (slice(1, -1),) * 3
    sx = pe[1:-1, 1:-1, :-2] + pe[1:-1, 1:-1, 2:]
    sy = pe[1:-1, :-2, 1:-1] + pe[1:-1, 2:, 1:-1]
    sz = pe[:-2, 1:-1, 1:-1] + pe[2:, 1:-1, 1:-1]
    assert sx.dtype == np.float32 and pe[c].shape == p.shape

    def finish(nb):
        v = (nb - div) * inv
        v[:, ny - 1, :] = 0
        return v.astype(np.float32)

    full = np.full_like(sx, 1)
    want = finish(_fma32(wz * full, sz, _fma32(wy * full, sy, wx * sx)))
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))
    plain = finish(wx * sx + wy * sy + wz * sz)
    other = finish(_fma32(wz * full, sz, _fma32(wx * full, sx, wy * sy)))
    assert (plain.view(np.uint32) != want.view(np.uint32)).sum() > 0
    assert (other.view(np.uint32) != want.view(np.uint32)).sum() > 0


# ---------------------------------------------------------------- step graph
def test_pass_order_and_ping_pong_indices():
    o = orc.Oracle((12, 16, 12), precision=32, pressure_iterations=5, correction_iterations=3)
    o.step(DT)
    assert o.pass_log() == ("simulate;divergence;jacobix5;project;divergence;clear_corr;jacobix3;project;"
                            "copy_flow;")
    assert (o.species_index, o.pressure_index) == (1, 1)
    o.step(DT)
    assert (o.species_index, o.pressure_index) == (0, 0)
    assert o.simulation_time == pytest.approx(2 * DT)
    h = orc.Oracle((12, 16, 12), precision=32, pressure_iterations=4, correction_iterations=4)
    h.set_heavy_mode(orc.HEAVY_F16)
    h.step(DT)
    assert h.pass_log() == ("simulate;divergence;jacobi_f16x4;project;divergence;clear_corr;jacobi_f16x4;project;"
                            "copy_flow;")
    c = orc.Oracle((12, 16, 12), precision=32, pressure_iterations=4, correction_iterations=4)
    c.set_heavy_mode(orc.HEAVY_F32)
    c.step(DT)
    assert c.pass_log().startswith("simulate;divergence;clear_pressure;jacobix4;project;divergence;clear_pressure;")


def test_primary_solve_is_warm_started():
    """The primary pressure is never cleared below 64 Mi cells
    (FireXEngine.ts:2248-2255): zeroing it between steps changes step 2."""
    def run(clear):
        o = orc.Oracle((16, 24, 16), precision=64, pressure_iterations=6, correction_iterations=6)
        o.step(DT)
        if clear:
            o.write(orc.FIELD_PRESSURE, np.zeros((16, 24, 16)))
        o.step(DT)
        return o.read(orc.FIELD_FLOW)
    assert not np.array_equal(run(False), run(True))
    # the compact heavy f32 path clears before each solve: no warm start
    def run_heavy(clear):
        o = orc.Oracle((16, 24, 16), precision=64, pressure_iterations=6, correction_iterations=6)
        o.set_heavy_mode(orc.HEAVY_F32)
        o.step(DT)
        if clear:
            o.write(orc.FIELD_PRESSURE, np.zeros((16, 24, 16)))
        o.step(DT)
        return o.read(orc.FIELD_FLOW)
    np.testing.assert_array_equal(run_heavy(False), run_heavy(True))


def test_fp32_step_tracks_fp64():
    dims = (32, 48, 32)
    a = orc.Oracle(dims, precision=32, pressure_iterations=16, correction_iterations=16)
    b = orc.Oracle(dims, precision=64, pressure_iterations=16, correction_iterations=16)
    a.step(DT, 10)
    b.step(DT, 10)
    for fid in (orc.FIELD_FLOW, orc.FIELD_SPECIES, orc.FIELD_PRESSURE):
        x, y = a.read(fid), b.read(fid)
        rel = np.linalg.norm(x - y) / np.linalg.norm(y)
        assert rel < 1e-3, (fid, rel)
