"""Pins each transport and force term of the oracle's simulate pass
(shaders.ts:151-342) with an expected value derived by hand from the
reference text, one term at a time.

Every case builds an input state in which the other terms of the pass vanish
or take a closed form: a zero-amplitude seeded emitter removes the source,
uniform or affine fields make advection and the six-neighbour blends exact,
temperatures below 520 K switch the turbulence off (smoothstep(520, 1650, T)
= 0) and below 690 K the reaction (smoothstep(690, 930, T) = 0), an oxygen
field at the ambient value removes the entrainment.  The expected values are
computed here in float64 numpy from the reference's constants, with the
uniforms (dt, scene parameters, time) rounded to f32 as the uniform block
carries them; the FP64 oracle is compared with them.

Terms and the reference lines they follow:
  six-neighbour diffusion blends            shaders.ts:177-187
  fuel / heat injection, swirl and lift     shaders.ts:189-201
  ambient-wall relaxation                   shaders.ts:203-208
  liquid cooling, evaporation, spray        shaders.ts:210-215, 290-297
  oxygen entrainment                        shaders.ts:217-232
  Arrhenius burn and ignition edges         shaders.ts:234-256
  vorticity-like force cross(grad T, curl)  shaders.ts:258-272
  lateral turbulence and height envelope    shaders.ts:274-285
  buoyancy                                  shaders.ts:287-288
  outflow band and top row                  shaders.ts:306-321
"""
import numpy as np
import pytest

from oracle import oracle as orc


def f32(v):
    """A uniform as the kernels see it."""
    return float(np.float32(v))


DT = f32(1.0 / 60.0)
# Inferno scene (FireXEngine.ts:422-451, scenes.ts:45-50), as f32 uniforms
FUEL_RATE, OXYGEN, HEAT_EFF, VORT = f32(0.30), f32(1.0), f32(0.86), f32(3.6)
BURNER_POWER, LIFT, BUOY, PERSIST = f32(2.5), f32(1.6), f32(2.15), f32(1.45)
ENTRAIN = f32(1.65)
W, H, D = f32(1.15), f32(1.35), f32(1.15)
AIM = f32(0.42)
AMBIENT = 300.0


# WGSL built-ins (WGSL spec, section 17.5), float64
def smoothstep(lo, hi, x):
    t = np.clip((np.asarray(x, float) - lo) / (hi - lo), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def mix(a, b, t):
    return a * (1.0 - t) + b * t


def radiated(T, dt=DT):
    """shaders.ts:299-302: Stefan-Boltzmann loss at the persistence scale."""
    q = min(max(PERSIST, 0.4), 3.0)
    return np.maximum(AMBIENT, T + dt * 5.67e-8 * (AMBIENT ** 4 - np.maximum(T, 0.0) ** 4) / (1005.0 * q))


def buoyancy(T, dt=DT):
    """shaders.ts:287-288."""
    return np.clip((np.asarray(T, float) - AMBIENT) / 1500.0, -0.1, 1.2) * dt * 0.72 * min(max(BUOY, 0.0), 4.0)


Q = min(max(PERSIST, 0.4), 3.0)
PRODUCT_DECAY = np.exp(-DT * 0.075 / Q)
VAPOR_DECAY = np.exp(-DT * 0.14 / Q)


def make(dims=(16, 32, 16), **params):
    """FP64 oracle with the source switched off: one seeded emitter of
    amplitude 0 replaces emitter() and injects exactly nothing."""
    o = orc.Oracle(dims, precision=64)
    o.set_emitters([(0.5, 0.5, 0.5, 1.0, 0.0)])
    for k, v in params.items():
        o.set_parameter(k, v)
    return o


def put(o, flow, species, liquid=None):
    nx, ny, nz = o.dims
    o.write(orc.FIELD_FLOW, np.broadcast_to(np.asarray(flow, float), (nz, ny, nx, 4)))
    o.write(orc.FIELD_SPECIES, np.broadcast_to(np.asarray(species, float), (nz, ny, nx, 4)))
    if liquid is not None:
        o.write(orc.FIELD_LIQUID, np.broadcast_to(np.asarray(liquid, float), (nz, ny, nx)))


def simulate(o, dt=DT):
    o.kernel_simulate(dt)
    return o.read(orc.FIELD_FLOW_PREDICTED), o.read(orc.FIELD_SPECIES), o.read(orc.FIELD_REACTION)


def grid(o):
    """Normalised cell centres npx, npy, npz as [z, y, x] arrays (shaders.ts:160-161)."""
    nx, ny, nz = o.dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return (x + 0.5) / f32(nx), (y + 0.5) / f32(ny), (z + 0.5) / f32(nz)


INTERIOR = (slice(1, -1), slice(1, 25), slice(1, -1))  # ny = 32: below the outflow band (npy < 0.78)


# ------------------------------------------------------------- diffusion blends
@pytest.mark.parametrize("dt", [DT, 0.75])
def test_six_neighbour_blends_from_a_single_cell_impulse(dt):
    """shaders.ts:177-187.  An impulse (u_x, T, products) in one cell at rest:
    every x neighbour advects its own value (zero velocity: the backtrace is
    the cell centre) and then mixes in the six-neighbour mean, one sixth of
    the impulse, with weights clamp(0.32 dt, 0, 0.12) for velocity,
    clamp(0.18 dt, 0, 0.08) for temperature and clamp(0.12 dt, 0, 0.06) for
    species.  dt = 0.75 reaches all three caps.  The x neighbours of a pure
    u_x impulse have no curl, so no other term moves their u_x."""
    o = make((16, 32, 20))
    nx, ny, nz = o.dims
    cx, cy, cz = 8, 12, 8
    U, Tc, P = 0.6, 480.0, 0.9
    flow = np.zeros((nz, ny, nx, 4))
    flow[..., 3] = AMBIENT
    flow[cz, cy, cx] = (U, 0.0, 0.0, Tc)
    spec = np.zeros((nz, ny, nx, 4))
    spec[..., 1] = 1.0
    spec[cz, cy, cx, 2] = P
    o.write(orc.FIELD_FLOW, flow)
    o.write(orc.FIELD_SPECIES, spec)
    fl, sp, _ = simulate(o, dt)
    wv = min(max(dt * 0.32, 0.0), 0.12)
    wt = min(max(dt * 0.18, 0.0), 0.08)
    ws = min(max(dt * 0.12, 0.0), 0.06)
    t_nb = radiated(mix(AMBIENT, (5 * AMBIENT + Tc) / 6.0, wt), dt)
    p_nb = mix(0.0, P / 6.0, ws) * np.exp(-dt * 0.075 / Q)
    for dx in (-1, 1):
        n = (cz, cy, cx + dx)
        assert fl[n][0] == pytest.approx(U / 6.0 * wv, rel=1e-12)
        assert fl[n][3] == pytest.approx(t_nb, rel=1e-12)
    for n in [(cz, cy, cx - 1), (cz, cy, cx + 1), (cz, cy - 1, cx), (cz, cy + 1, cx), (cz - 1, cy, cx),
              (cz + 1, cy, cx)]:
        assert sp[n][2] == pytest.approx(p_nb, rel=1e-12)
        assert fl[n][3] == pytest.approx(t_nb, rel=1e-12)
    # two cells away nothing arrives
    assert sp[cz, cy, cx + 2, 2] == 0.0 and fl[cz, cy, cx + 2, 0] == 0.0


@pytest.mark.parametrize("dt", [DT, 0.75])
def test_impulse_neighbours_across_the_flow(dt):
    """The y and z neighbours of a u_x impulse at 300 K: the blend brings
    U/6 clamp(0.32 dt, 0, 0.12) of u_x from every side, and the impulse is a
    shear there, so the vorticity-like force acts too (shaders.ts:258-272).
    With a uniform temperature the gradient direction is normalize(1e-5 (1,
    1, 1)) = (1, 1, 1)/sqrt 3; the curl is (0, 0.5 (F.x - K.x) mz, 0) at the
    z neighbours and (0, 0, -0.5 (Tp.x - B.x) my) at the y neighbours.  The
    grid's x and z metrics differ, so a wrong metric shows."""
    o = make((16, 32, 20))
    nx, ny, nz = o.dims
    cx, cy, cz = 8, 12, 8
    U = 0.6
    flow = np.zeros((nz, ny, nx, 4))
    flow[..., 3] = AMBIENT
    flow[cz, cy, cx, 0] = U
    o.write(orc.FIELD_FLOW, flow)
    o.write(orc.FIELD_SPECIES, np.broadcast_to(np.array([0.0, 1.0, 0.0, 0.0]), (nz, ny, nx, 4)))
    fl, _, _ = simulate(o, dt)
    wv = min(max(dt * 0.32, 0.0), 0.12)
    vk = 0.00135 * VORT * dt
    mx, my, mz = nx / W, ny / H, nz / D
    a = 1.0 / np.sqrt(3.0)
    blend = U / 6.0 * wv
    for dz, sign in ((-1, 1.0), (1, -1.0)):  # F = c below the impulse, K = c above
        curl_y = 0.5 * sign * U * mz
        want = np.array([blend - a * curl_y * vk, 0.0, a * curl_y * vk])
        np.testing.assert_allclose(fl[cz + dz, cy, cx, :3], want, rtol=1e-10, atol=1e-17)
    for dy, sign in ((-1, 1.0), (1, -1.0)):  # Tp = c below, B = c above
        curl_z = -0.5 * sign * U * my
        want = np.array([blend + a * curl_z * vk, -a * curl_z * vk, 0.0])
        np.testing.assert_allclose(fl[cz, cy + dy, cx, :3], want, rtol=1e-10, atol=1e-17)


# ------------------------------------------------------------ source injection
def test_source_injection_swirl_and_lift():
    """shaders.ts:189-201 with a known source s (a seeded Gaussian of radius
    1e4 cells, amplitude 0.1, whose value the test computes): fuel =
    s fuelRate P, T = 300 + 1180 s min(r, 1 + 0.22 (r - 1)) with r = sqrt(P),
    swirl (sin(2.3 t + 22 z_n), 0, cos(1.9 t + 25 x_n)) s dt 0.11 vort r and
    lift s dt 0.52 clamp(sourceLift, 0, 4) r on u_y, at t = 0.375.  The
    source heats to about 430 K: no burn, no turbulence; u_y also carries the
    buoyancy of that temperature."""
    o = make()
    o.set_emitters([(8.0, 3.0, 8.0, 1.0e4, 0.1)])
    t0 = 0.375
    o.set_time(t0, 7)
    put(o, (0, 0, 0, AMBIENT), (0, 1.0, 0, 0))
    fl, sp, react = simulate(o)
    npx, npy, npz = grid(o)
    nx, ny, nz = o.dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    d2 = (x + 0.5 - 8.0) ** 2 + (y + 0.5 - 3.0) ** 2 + (z + 0.5 - 8.0) ** 2
    s = f32(0.1) * np.exp(-(d2 / 1.0e8))  # emitter parameters are stored as f32
    P = min(max(BURNER_POWER, 0.1), 4.0)
    r = np.sqrt(P)
    thermal = min(r, 1.0 + 0.22 * (r - 1.0))
    T_src = AMBIENT + s * 1180.0 * thermal
    k = s * DT * 0.11 * VORT * r
    i = INTERIOR
    np.testing.assert_allclose(sp[..., 0][i], (s * FUEL_RATE * P)[i], rtol=1e-12)
    np.testing.assert_allclose(fl[..., 3][i], radiated(T_src)[i], rtol=1e-12)
    np.testing.assert_allclose(fl[..., 0][i], (np.sin(t0 * 2.3 + npz * 22.0) * k)[i], rtol=1e-9, atol=1e-16)
    np.testing.assert_allclose(fl[..., 2][i], (np.cos(t0 * 1.9 + npx * 25.0) * k)[i], rtol=1e-9, atol=1e-16)
    lift = s * DT * 0.52 * min(max(LIFT, 0.0), 4.0) * r
    np.testing.assert_allclose(fl[..., 1][i], (lift + buoyancy(T_src))[i], rtol=1e-12)
    assert np.all(react[i] == 0.0)


# --------------------------------------------------------------- ambient walls
def test_ambient_walls_relax_temperature_and_restore_oxygen():
    """shaders.ts:203-208: on x = 0, x = nx - 1, z = 0, z = nz - 1 and the top
    row (not the floor), O2 = max(O2, ambient) and T = mix(T, 300, 0.035),
    once per cell (a corner is relaxed once).  Uniform 1000 K, O2 0.5, no
    fuel: walls hold 975.5 K before radiation, the interior and the floor
    1000 K; every wall cell and the top row end with O2 = 1 exactly."""
    o = make()
    put(o, (0, 0, 0, 1000.0), (0, 0.5, 0, 0))
    fl, sp, _ = simulate(o)
    nx, ny, nz = o.dims
    T_wall = radiated(mix(1000.0, AMBIENT, 0.035))
    T_in = radiated(1000.0)
    ys = slice(1, 25)
    for cells in [(slice(1, -1), ys, 0), (slice(1, -1), ys, nx - 1), (0, ys, slice(1, -1)), (nz - 1, ys, slice(1, -1)),
                  (0, ys, 0), (nz - 1, ys, nx - 1)]:
        np.testing.assert_allclose(fl[..., 3][cells], T_wall, rtol=1e-12)
        np.testing.assert_array_equal(sp[..., 1][cells], 1.0)
    np.testing.assert_allclose(fl[1:-1, 1:25, 1:-1, 3], T_in, rtol=1e-12)
    np.testing.assert_allclose(fl[1:-1, 0, 1:-1, 3], T_in, rtol=1e-12)  # the floor is not an ambient boundary
    assert np.all(sp[1:-1, 1:25, 1:-1, 1] < 1.0)
    np.testing.assert_array_equal(sp[:, ny - 1, :, 1], 1.0)
    np.testing.assert_array_equal(fl[:, ny - 1, :, 3], AMBIENT)


# ---------------------------------------------------------------------- liquid
@pytest.mark.parametrize("liquid_units", [1229, 4096 * 2])
def test_liquid_cooling_evaporation_and_spray(liquid_units):
    """shaders.ts:210-215 and 290-297 with a uniform deposited liquid field
    (u32, 4096 per unit; 2 units reach the water cap of 5): evaporation =
    w smoothstep(335, 1040, T) dt 13.5, T = max(300, T - 2080 evap -
    260 w dt), vapor += 2.15 evap, products *= exp(-4.8 w dt), and the spray
    momentum normalize((0.52, aim, 0.5) - (0.035, 0.5, 0.5)) * extent) w dt
    0.82 on the velocity.  500 K with no fuel: nothing burns and nothing
    is turbulent."""
    o = make()
    p0 = 0.5
    put(o, (0, 0, 0, 500.0), (0, 1.0, p0, 0), liquid=liquid_units)
    fl, sp, _ = simulate(o)
    w = min(max(liquid_units / 4096.0 * 3.4, 0.0), 5.0)
    evap = w * smoothstep(335.0, 1040.0, 500.0) * DT * 13.5
    T1 = max(AMBIENT, 500.0 - evap * 2080.0 - w * DT * 260.0)
    vapor = min(max(evap * 2.15, 0.0), 3.0) * VAPOR_DECAY
    prod = p0 * np.exp(-w * DT * 4.8) * PRODUCT_DECAY
    sdir = np.array([(0.52 - 0.035) * W, (AIM - 0.50) * H, 0.0])
    sdir /= np.linalg.norm(sdir)
    i = INTERIOR
    np.testing.assert_allclose(fl[..., 3][i], radiated(T1), rtol=1e-12)
    np.testing.assert_allclose(sp[..., 3][i], vapor, rtol=1e-12)
    np.testing.assert_allclose(sp[..., 2][i], prod, rtol=1e-12)
    np.testing.assert_allclose(fl[..., 0][i], sdir[0] * w * DT * 0.82, rtol=1e-12)
    np.testing.assert_allclose(fl[..., 1][i], sdir[1] * w * DT * 0.82 + buoyancy(T1), rtol=1e-10, atol=1e-16)
    np.testing.assert_allclose(fl[..., 2][i], 0.0, atol=1e-15)


# ---------------------------------------------------------------- entrainment
@pytest.mark.parametrize("dt,velocity,ambient_o2", [(DT, (0.3, 0.4, 0.12), 1.0), (DT, (0.0, 0.0, 0.0), 1.0),
                                                   (DT, (0.3, 0.4, 0.12), 0.8), (0.5, (0.3, -0.4, 0.12), 1.0)])
def test_oxygen_entrainment(dt, velocity, ambient_o2):
    """shaders.ts:217-232: O2 = mix(O2, ambient, m) with m = clamp(dt ent
    (0.12 + 0.28 |u| + 0.22 smoothstep(400, 1500, T)) deficit mix(1, 0.18,
    smoothstep(0.30, 0.78, fuel)), 0, 0.08) and deficit = (ambient - O2) /
    ambient.  Uniform flow at 500 K (no ignition, no turbulence) with fuel
    0.5 and O2 0.4; dt = 0.5 reaches the 0.08 cap."""
    o = make(oxygenRate=ambient_o2)
    fuel, ox = 0.5, 0.4
    put(o, (*velocity, 500.0), (fuel, ox, 0, 0))
    _, sp, _ = simulate(o, dt)
    amb = max(f32(ambient_o2), 0.001)
    speed = float(np.linalg.norm(velocity))
    deficit = min(max((amb - ox) / amb, 0.0), 1.0)
    weight = deficit * mix(1.0, 0.18, smoothstep(0.30, 0.78, fuel))
    m = min(max(dt * ENTRAIN * (0.12 + 0.28 * speed + 0.22 * smoothstep(400.0, 1500.0, 500.0)) * weight, 0.0), 0.08)
    if dt == 0.5:
        assert m == 0.08
    np.testing.assert_allclose(sp[..., 1][INTERIOR], mix(ox, amb, m), rtol=1e-12)
    np.testing.assert_allclose(sp[..., 0][INTERIOR], fuel, rtol=1e-12)


# ------------------------------------------------------------------- reaction
def expected_burn(T, fuel, ox, dt=DT, water=0.0):
    """shaders.ts:236-247 at the temperature the reaction sees."""
    avail = min(fuel, ox * 0.25)
    k = 2.5e6 * np.exp(-15000.0 / max(T, AMBIENT)) * max(fuel, 1e-5) ** -0.30 * max(ox, 1e-5) ** 1.30
    k = min(max(k, 0.0), 80.0)
    ign = smoothstep(690.0, 930.0, T)
    return min(max(avail * (1.0 - np.exp(-k * ign * dt)) * np.exp(-water * 18.0), 0.0), avail), k


@pytest.mark.parametrize("T0", [650.0, 689.5, 692.0, 695.0, 700.0, 760.0, 850.0, 925.0, 935.0, 1200.0, 1800.0, 2600.0])
@pytest.mark.parametrize("fuel,ox", [(0.2, 0.6), (0.05, 1.0), (0.5, 0.3)])
def test_burn_follows_arrhenius_and_ignition_edges(T0, fuel, ox):
    """The burn of a uniform mixture at rest (entrainment off, so O2 is the
    reaction's input): zero below 690 K, avail (1 - exp(-k ign dt)) with
    ign = smoothstep(690, 930, T) and the capped Arrhenius rate above.  At
    2600 K the rate is capped at 80, so burn = avail (1 - exp(-80 dt))."""
    o = make(airEntrainment=0.0)
    put(o, (0, 0, 0, T0), (fuel, ox, 0, 0))
    _, sp, react = simulate(o)
    burn = fuel - sp[..., 0][INTERIOR]
    want, k = expected_burn(T0, fuel, ox)
    if T0 < 690.0:
        assert np.all(burn == 0.0)
    else:
        assert want > 0
    # 1 - exp(-x) loses relative precision for the tiny rates just above
    # 690 K (in the oracle as in the shader): an absolute bound there
    np.testing.assert_allclose(burn, want, rtol=1e-9, atol=1e-16)
    np.testing.assert_allclose(react[INTERIOR], min(want / DT * 0.22, 4.0), rtol=1e-9, atol=1e-14)
    if T0 == 2600.0:
        assert k == 80.0
        np.testing.assert_allclose(burn, min(fuel, ox / 4) * (1 - np.exp(-80.0 * DT)), rtol=1e-12)


def test_liquid_suppresses_the_burn():
    """shaders.ts:245-246: the burn carries exp(-18 w), and the reaction sees
    the temperature after the liquid cooling."""
    o = make(airEntrainment=0.0)
    units, T0, fuel, ox = 300, 1300.0, 0.2, 0.8
    put(o, (0, 0, 0, T0), (fuel, ox, 0, 0), liquid=units)
    _, sp, _ = simulate(o)
    w = min(max(units / 4096.0 * 3.4, 0.0), 5.0)
    evap = w * smoothstep(335.0, 1040.0, T0) * DT * 13.5
    T1 = max(AMBIENT, T0 - evap * 2080.0 - w * DT * 260.0)
    want, _ = expected_burn(T1, fuel, ox, water=w)
    np.testing.assert_allclose(fuel - sp[..., 0][INTERIOR], want, rtol=1e-9)


# ------------------------------------------------------------ vorticity force
def test_vorticity_force_is_cross_of_gradient_direction_and_curl():
    """shaders.ts:258-272 on affine fields: u = (a y + b z, c x + d z,
    e x + f y), T = 350 + gx x + gy y + gz z (cell units, below 520 K).
    Central differences are exact: curl = (f my - d mz, b mz - e mx,
    c mx - a my) and grad T = (gx mx, gy my, gz mz); the force is
    cross(normalize(grad T + 1e-5), curl) 0.00135 vort dt.  It is the only
    term that depends on the vorticity parameter in this state (no source,
    no turbulence), so two runs, vorticity 3.6 and 0, differ by exactly the
    force."""
    dims = (16, 32, 20)
    nx, ny, nz = dims
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    a, b, c, d, e, f = 0.011, -0.007, 0.013, 0.005, -0.009, 0.004
    gx, gy, gz = 2.0, 3.0, 1.5
    flow = np.stack([a * y + b * z, c * x + d * z, e * x + f * y, 350.0 + gx * x + gy * y + gz * z], axis=-1)
    spec = np.broadcast_to(np.array([0.0, 1.0, 0.0, 0.0]), (nz, ny, nx, 4))
    runs = []
    for vort in (3.6, 0.0):
        o = make(dims, vorticity=vort)
        o.write(orc.FIELD_FLOW, flow)
        o.write(orc.FIELD_SPECIES, spec)
        runs.append(simulate(o)[0])
    mx, my, mz = nx / W, ny / H, nz / D
    curl = np.array([f * my - d * mz, b * mz - e * mx, c * mx - a * my])
    g = np.array([gx * mx, gy * my, gz * mz]) + 1e-5
    force = np.cross(g / np.linalg.norm(g), curl) * 0.00135 * VORT * DT
    diff = (runs[0] - runs[1])[2:-2, 2:-2, 2:-2, :3]
    np.testing.assert_allclose(diff, np.broadcast_to(force, diff.shape), rtol=1e-7, atol=1e-15)
    assert np.all(np.abs(force) > 1e-7)


# ----------------------------------------------------------------- turbulence
@pytest.mark.parametrize("T0", [510.0, 1000.0, 1400.0, 1700.0, 2600.0])
def test_lateral_turbulence_and_height_envelope(T0):
    """shaders.ts:274-288 in a uniform hot gas at rest: u += (sin(31 y_n + 19 z_n + 3.1 t) + 0.45 sin(67 y_n -
    5.3 t), 0.12 sin(29 x_n + 23 z_n - 4.1 t), cos(27 y_n - 21 x_n + 2.7 t) +
    0.45 cos(59 y_n + 4.7 t)) (0.075 + 0.16 vort) envelope(y_n) dt with the
    envelope smoothstep(0.045, 0.15, y_n) (1 - smoothstep(0.80, 1, y_n)), at
    t = 1.25, all scaled by hot = smoothstep(520, 1650, T) (0 at 510 K, 1 at
    1700 K); u_y also has the buoyancy, capped at 1.2 at 2600 K."""
    o = make()
    t0 = 1.25
    o.set_time(t0, 75)
    put(o, (0, 0, 0, T0), (0, 1.0, 0, 0))
    fl, _, _ = simulate(o)
    npx, npy, npz = grid(o)
    env = smoothstep(0.045, 0.15, npy) * (1.0 - smoothstep(0.80, 1.0, npy))
    tk = (0.075 + 0.16 * VORT) * smoothstep(520.0, 1650.0, T0) * env * DT
    lt = [np.sin(npy * 31.0 + npz * 19.0 + t0 * 3.1) + 0.45 * np.sin(npy * 67.0 - t0 * 5.3),
          0.12 * np.sin(npx * 29.0 + npz * 23.0 - t0 * 4.1),
          np.cos(npy * 27.0 - npx * 21.0 + t0 * 2.7) + 0.45 * np.cos(npy * 59.0 + t0 * 4.7)]
    i = (slice(1, -1), slice(1, -1), slice(1, -1))
    np.testing.assert_allclose(fl[..., 0][i], (lt[0] * tk)[i], rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(fl[..., 2][i], (lt[2] * tk)[i], rtol=1e-9, atol=1e-15)
    np.testing.assert_allclose(fl[..., 1][i], (lt[1] * tk + buoyancy(T0))[i], rtol=1e-9, atol=1e-15)
    if T0 == 2600.0:
        assert buoyancy(T0) == pytest.approx(1.2 * DT * 0.72 * BUOY)
    if T0 == 510.0:
        assert np.all(fl[..., 0] == 0.0) and np.all(fl[..., 2] == 0.0)
    # walls and floor zero their normal component
    nx, ny, nz = o.dims
    assert np.all(fl[:, :, 0, 0] == 0) and np.all(fl[:, :, nx - 1, 0] == 0)
    assert np.all(fl[0, :, :, 2] == 0) and np.all(fl[nz - 1, :, :, 2] == 0)
    assert np.all(fl[:, 0, :, 1] == 0)


# -------------------------------------------------------------------- outflow
def test_outflow_band_and_top_row():
    """shaders.ts:306-321: with o = smoothstep(0.84, 0.995, y_n), fuel *=
    1 - o, products and vapor *= exp(-15 o dt / q), T = mix(T, 300, 0.16 o
    / q); the top row holds no fuel, products or vapor and 300 K.  Uniform
    500 K gas at rest with all species, no burn; ny = 32 puts rows 27-30 in
    the band."""
    o = make()
    f0, ox, p0, v0 = 0.4, 1.0, 0.5, 0.3
    put(o, (0, 0, 0, 500.0), (f0, ox, p0, v0))
    fl, sp, _ = simulate(o)
    nx, ny, nz = o.dims
    inner = (slice(1, -1), slice(1, -1))
    for yy in range(20, ny - 1):
        out = float(smoothstep(0.84, 0.995, (yy + 0.5) / ny))
        if yy >= 27:
            assert out > 0
        decay = np.exp(-out * DT * 15.0 / Q)
        np.testing.assert_allclose(sp[inner[0], yy, inner[1], 0], f0 * (1.0 - out), rtol=1e-12)
        np.testing.assert_allclose(sp[inner[0], yy, inner[1], 2], p0 * PRODUCT_DECAY * decay, rtol=1e-12)
        np.testing.assert_allclose(sp[inner[0], yy, inner[1], 3], v0 * VAPOR_DECAY * decay, rtol=1e-12)
        np.testing.assert_allclose(fl[inner[0], yy, inner[1], 3], mix(radiated(500.0), AMBIENT, out * 0.16 / Q),
                                   rtol=1e-12)
    top = sp[:, ny - 1]
    assert np.all(top[..., 0] == 0) and np.all(top[..., 2] == 0) and np.all(top[..., 3] == 0)
    assert np.all(fl[:, ny - 1, :, 3] == AMBIENT)


def test_outflow_scales_the_reaction_diagnostic():
    """shaders.ts:311, 315: in the band the burn diagnostic is multiplied by
    1 - o after the fuel is (fuel after the burn, then * (1 - o))."""
    o = make(airEntrainment=0.0)
    f0, ox, T0 = 0.2, 0.6, 1400.0
    put(o, (0, 0, 0, T0), (f0, ox, 0, 0))
    _, sp, react = simulate(o)
    ny = o.dims[1]
    want, _ = expected_burn(T0, f0, ox)
    for yy in (27, 28, 29, 30):
        out = float(smoothstep(0.84, 0.995, (yy + 0.5) / ny))
        np.testing.assert_allclose(sp[1:-1, yy, 1:-1, 0], (f0 - want) * (1.0 - out), rtol=1e-9)
        np.testing.assert_allclose(react[1:-1, yy, 1:-1], min(want / DT * 0.22, 4.0) * (1.0 - out), rtol=1e-9)
    assert np.all(react[:, ny - 1] == 0.0)


def test_temperature_cap():
    """shaders.ts:338: T is clamped to [300, 2900] last.  A 2895 K gas at rest
    radiates to about 2849 K, below the cap; a burning 2880 K mixture would
    rise above 3300 K and is held at 2900 K."""
    o = make()
    put(o, (0, 0, 0, 2895.0), (0, 1.0, 0, 0))
    fl, _, _ = simulate(o)
    assert 2800.0 < radiated(2895.0) < 2890.0
    np.testing.assert_allclose(fl[..., 3][INTERIOR], radiated(2895.0), rtol=1e-12)
    o = make(airEntrainment=0.0)
    put(o, (0, 0, 0, 2880.0), (0.2, 0.8, 0, 0))
    fl, _, _ = simulate(o)
    np.testing.assert_array_equal(fl[..., 3][INTERIOR], 2900.0)
