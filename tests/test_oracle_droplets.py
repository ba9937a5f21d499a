"""Pins of the oracle's SPH droplet passes (shaders.ts:514-825) and host
particle state (FireXEngine.ts:1976-2034, 2158-2171) by facts that do not
come from its own code:

  * initial state: unit launch directions, the speed band, mass, the
    nozzle-disk offset perpendicular to the aim direction, the inactive state;
  * density: pair values q^2 and q^3 in closed form, exclusions (self,
    coincident, inactive, beyond the radius), the radius switch by nozzle;
  * integration: the closed-form update of an isolated droplet in still and
    hot gas, the floor and wall rules, the velocity clamp, repulsion of a
    close pair in closed form, and momentum balance of a cluster (pair
    forces cancel: the sum of velocity changes is N * gravity * dt);
  * deposition: trilinear weights at known fractions, the normalised 3^3
    kernel at a cell centre in closed form, the three-sample sweep, nothing
    outside the domain, conservation up to the u32 truncation per corner;
  * the water branch of the step graph: pass log, liquid clear and the
    one-time reset after water is switched off.
The FP32 oracle is the one the HIP kernels are compared with bit for bit;
closed forms are checked on the FP64 oracle and, within f32 rounding, on the
FP32 oracle."""
import math

import numpy as np
import pytest

from oracle import oracle as orc

DIMS = (32, 32, 32)
# default domain (metres) as the f32 uniform block carries it
EXTENT = np.array([1.15, 1.35, 1.15], dtype=np.float32).astype(np.float64)
DT = 1.0 / 60.0
DTF = float(np.float32(DT))  # dt as the f32 uniform block carries it
LAMINAR_RADIUS, SPRAY_RADIUS = 0.038, 0.060


def _oracle(precision=64, count=16, water=False, flow=0.8, nozzle=0):
    o = orc.Oracle(DIMS, precision=precision, pressure_iterations=2, correction_iterations=2)
    o.set_parameter("waterFlow", flow)
    o.set_parameter("nozzleType", nozzle)
    o.set_bool("waterEnabled", water)
    o.set_particle_count(count)
    return o


def _particles(count, records):
    """A particle array with the given (index -> 8 values) records; all
    other particles inactive far apart (mass 0)."""
    p = np.zeros((count, 8))
    p[:, 0] = np.linspace(0.1, 0.9, count)
    p[:, 1] = 0.9
    p[:, 2] = 0.9
    p[:, 3] = -0.5
    for i, rec in records.items():
        p[i] = rec
    return p


def _pair_terms(d, radius):
    q = 1.0 - d / radius
    return q * q, q * q * q


# ---------------------------------------------------------------- initial state
@pytest.mark.parametrize("nozzle", [0, 1])
def test_initial_state_directions_speeds_and_mass(nozzle):
    wf = 0.8
    o = _oracle(count=256, water=True, flow=wf, nozzle=nozzle)
    p = o.read(orc.FIELD_PARTICLES)
    age = np.arange(256) / 256 * 0.72
    np.testing.assert_allclose(p[:, 3], age.astype(np.float32), rtol=0, atol=0)
    launch = p[:, 4:7].copy()
    launch[:, 1] += 0.58 * age  # remove the gravity term of the seeded velocity
    speed = np.linalg.norm(launch, axis=1)
    variation = 0.24 if nozzle else 0.05
    base = (0.82 + 0.82 * wf)
    assert np.all(speed >= base * (1 - variation / 2) * (1 - 1e-6))
    assert np.all(speed <= base * (1 + variation / 2) * (1 + 1e-6))
    np.testing.assert_array_equal(p[:, 7], np.float32(0.52 + 0.88 * wf))
    # particle 0 has age 0: its offset from the nozzle lies in the disk
    # perpendicular to the aim direction, within the origin radius
    aim = np.array([(0.52 - 0.035) * 1.15, (0.42 - 0.5) * 1.35, 0.0])
    aim /= np.linalg.norm(aim)
    offset = (p[0, :3] - [0.035, 0.5, 0.5]) * EXTENT
    assert abs(offset @ aim) < 1e-7
    assert np.linalg.norm(offset) <= (0.006 if nozzle else 0.0028) * (1 + 1e-5)
    o.close()


def test_zero_spray_angle_launches_along_the_aim():
    o = _oracle(count=64, water=True)
    o.set_parameter("sprayAngle", 0.0)
    o.write_particle_initial_state(True)
    p = o.read(orc.FIELD_PARTICLES)
    age = np.arange(64) / 64 * 0.72
    launch = p[:, 4:7].copy()
    launch[:, 1] += 0.58 * age
    aim = np.array([(0.52 - 0.035) * 1.15, (0.42 - 0.5) * 1.35, 0.0])
    aim /= np.linalg.norm(aim)
    unit = launch / np.linalg.norm(launch, axis=1, keepdims=True)
    np.testing.assert_allclose(unit, np.broadcast_to(aim, unit.shape), atol=2e-7)
    assert np.all(p[:, 6] == 0.0)
    o.close()


def test_inactive_state_has_no_velocity_or_mass():
    o = _oracle(count=100, water=False)
    p = o.read(orc.FIELD_PARTICLES)
    assert np.all(p[:, 4:8] == 0.0)
    np.testing.assert_array_equal(p[:, 3], (-(np.arange(100) / 100) * 0.72).astype(np.float32))
    assert np.all(p[:, 1] >= np.float32(0.012))
    o.close()


# ---------------------------------------------------------------------- density
@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("nozzle", [0, 1])
def test_density_pair_closed_form_and_exclusions(precision, nozzle):
    radius = SPRAY_RADIUS if nozzle else LAMINAR_RADIUS
    o = _oracle(precision=precision, count=8, nozzle=nozzle)
    d = 0.4 * radius
    x0 = 0.5
    x1 = x0 + d / EXTENT[0]
    recs = {
        0: [x0, 0.5, 0.5, 0.1, 0, 0, 0, 1.0],
        1: [x1, 0.5, 0.5, 0.1, 0, 0, 0, 1.0],
        2: [x0, 0.5, 0.5 + 0.5 * radius / EXTENT[2], 0.1, 0, 0, 0, 0.0],   # inactive: ignored
        3: [x0, 0.5 + 1.5 * radius / EXTENT[1], 0.5, 0.1, 0, 0, 0, 1.0],  # beyond the radius
        4: [0.2, 0.2, 0.2, 0.1, 0, 0, 0, 1.0],
        5: [0.2, 0.2, 0.2, 0.1, 0, 0, 0, 1.0],                             # coincident with 4: ignored
    }
    o.write(orc.FIELD_PARTICLES, _particles(8, recs))
    o.kernel_droplets(DT, orc.DROPLET_DENSITY)
    dens = o.read(orc.FIELD_PARTICLE_DENSITY)
    dd = np.linalg.norm(([x1, 0.5, 0.5] - np.array([x0, 0.5, 0.5])) * EXTENT) if precision == 64 else d
    q2, q3 = _pair_terms(dd, radius)
    tol = dict(rtol=1e-12) if precision == 64 else dict(rtol=2e-5, atol=1e-6)
    for i in (0, 1):
        np.testing.assert_allclose(dens[i], [q2, q3, 0.0, q3 * 0.065], **tol)
    assert np.all(dens[2] == 0.0)   # inactive particles write zero
    assert np.all(dens[3][:2] == 0.0) and np.all(dens[4][:2] == 0.0) and np.all(dens[5][:2] == 0.0)
    o.close()


def test_density_pressure_above_rest_density():
    """Seven particles packed within a small ball: density above the rest
    density 2.7 + 0.9 * waterFlow gives pressure (density - rest) * 0.055."""
    wf = 0.5
    o = _oracle(count=7, nozzle=1, flow=wf)
    rng = np.random.default_rng(1)
    pos = 0.5 + rng.uniform(-1, 1, (7, 3)) * 0.004 / EXTENT
    recs = {i: [*pos[i], 0.1, 0, 0, 0, 1.0] for i in range(7)}
    o.write(orc.FIELD_PARTICLES, _particles(7, recs))
    o.kernel_droplets(DT, orc.DROPLET_DENSITY)
    dens = o.read(orc.FIELD_PARTICLE_DENSITY)
    m = pos * EXTENT
    for i in range(7):
        dist = np.linalg.norm(m[i] - np.delete(m, i, axis=0), axis=1)
        q = 1 - dist / SPRAY_RADIUS
        rho, near = np.sum(q * q), np.sum(q ** 3)
        np.testing.assert_allclose(dens[i], [rho, near, max(rho - (2.7 + 0.9 * wf), 0) * 0.055, near * 0.065],
                                   rtol=1e-12)
        assert dens[i][2] > 0
    o.close()


# ------------------------------------------------------------------ integration
def _integrate(o, particles, dt=DT):
    o.write(orc.FIELD_PARTICLES, particles)
    o.kernel_droplets(dt, orc.DROPLET_DENSITY)
    o.kernel_droplets(dt, orc.DROPLET_INTEGRATE)
    return o.read(orc.FIELD_PARTICLES)


def _gas(o, velocity=(0.0, 0.0, 0.0), temperature=300.0):
    nx, ny, nz = DIMS
    f = np.zeros((nz, ny, nx, 4))
    f[..., :3] = velocity
    f[..., 3] = temperature
    o.write(orc.FIELD_FLOW, f)


@pytest.mark.parametrize("temperature", [300.0, 1500.0])
def test_isolated_droplet_closed_form(temperature):
    o = _oracle(count=4)
    gas_v = np.array([0.3, 0.5, -0.2])
    _gas(o, gas_v, temperature)
    v0 = np.array([0.4, -0.1, 0.2])
    p0 = np.array([0.5, 0.5, 0.5])
    out = _integrate(o, _particles(4, {1: [*p0, 0.1, *v0, 0.9]}))[1]
    heat = 0.0 if temperature < 380 else 1.0
    acc = np.array([0.0, -0.58, 0.0]) + (gas_v - v0) * (0.30 + 0.18 * heat)
    v1 = v0 + acc * DTF
    np.testing.assert_allclose(out[4:7], v1, rtol=1e-14)
    np.testing.assert_allclose(out[:3], p0 + v1 * DTF / EXTENT, rtol=1e-14)
    assert out[3] == pytest.approx(0.1 + DTF, rel=1e-15)
    assert out[7] == pytest.approx(0.9 * math.exp(-heat * DTF * 0.34), rel=1e-15)
    o.close()


def test_floor_wall_and_clamp_rules():
    o = _oracle(count=4)
    _gas(o)
    recs = {
        0: [0.5, 0.0121, 0.5, 0.1, 0.3, -1.0, 0.2, 1.0],   # through the floor
        1: [0.9915, 0.5, 0.5, 0.1, 1.5, 0.0, 0.0, 1.0],    # through the x = 0.992 wall
        2: [0.5, 0.5, 0.5, 0.1, 0.0, 0.0, 3.0, 1.0],       # z speed clamped to 2
    }
    out = _integrate(o, _particles(4, recs))
    g = np.array([0.0, -0.58, 0.0])
    v = np.array(recs[0][4:7]); v = v + (g - v * 0.30) * DTF
    assert out[0][1] == pytest.approx(0.012, abs=1e-15)
    np.testing.assert_allclose(out[0][4:7], [v[0] * 0.72, abs(v[1]) * 0.18, v[2] * 0.72], rtol=1e-14)
    v = np.array(recs[1][4:7]); v = v + (g - v * 0.30) * DTF
    assert out[1][0] == pytest.approx(0.992, abs=1e-15)
    assert out[1][4] == pytest.approx(v[0] * -0.24, rel=1e-14)
    assert out[2][6] == 2.0
    o.close()


def test_close_pair_repels_in_closed_form():
    """Two droplets 0.5 r apart along x in still air, both at rest: each is
    pushed away from the other by dir * ((p_i + p_j) q + (n_i + n_j) q^2)
    times the laminar dispersion 0.055."""
    o = _oracle(count=2)
    _gas(o)
    d = 0.5 * LAMINAR_RADIUS
    x0, x1 = 0.5, 0.5 + d / EXTENT[0]
    out = _integrate(o, np.array([[x0, 0.5, 0.5, 0.1, 0, 0, 0, 1.0], [x1, 0.5, 0.5, 0.1, 0, 0, 0, 1.0]]))
    dd = (x1 - x0) * EXTENT[0]
    q = 1 - dd / LAMINAR_RADIUS
    press, near = 0.0, q ** 3 * 0.065  # density q^2 < rest: zero pressure
    push = (2 * press) * q + (2 * near) * q * q
    ax = push * 0.055
    assert out[0][4] == pytest.approx(-ax * DTF, rel=1e-12)
    assert out[1][4] == pytest.approx(ax * DTF, rel=1e-12)
    assert out[0][4] < 0 < out[1][4]
    o.close()


@pytest.mark.parametrize("nozzle", [0, 1])
def test_cluster_momentum_balance(nozzle):
    """Pair forces (pressure and viscosity) cancel in sum: starting from rest
    in still air, the velocity changes add up to N * gravity * dt."""
    n = 40
    o = _oracle(count=n, nozzle=nozzle)
    _gas(o)
    rng = np.random.default_rng(7)
    pos = 0.5 + rng.uniform(-1, 1, (n, 3)) * 0.03 / EXTENT
    vel = rng.uniform(-0.05, 0.05, (n, 3))
    recs = {i: [*pos[i], 0.1, *vel[i], 1.0] for i in range(n)}
    before = _particles(n, recs)
    out = _integrate(o, before)
    dv = out[:, 4:7] - before[:, 4:7]
    expected = n * np.array([0.0, -0.58, 0.0]) * DTF + np.sum(-before[:, 4:7] * 0.30 * DTF, axis=0)
    np.testing.assert_allclose(dv.sum(axis=0), expected, atol=1e-12)
    assert np.abs(dv - (np.array([0.0, -0.58, 0.0]) - before[:, 4:7] * 0.30) * DTF).max() > 1e-4  # forces act
    o.close()


def test_expiry_and_inactive_rules():
    o = _oracle(count=4, water=False)
    _gas(o)
    recs = {
        0: [0.5, 0.5, 0.5, 5.0, 0.1, 0.1, 0.1, 1.0],    # older than any lifetime (<= 1.4 s)
        1: [0.5, 0.5, 1.05, 0.1, 0.1, 0.1, 0.1, 1.0],   # outside the domain
        2: [0.3, 0.3, 0.3, -0.5, 0.0, 0.0, 0.0, 0.0],   # inactive, waiting
    }
    out = _integrate(o, _particles(4, recs))
    for i in (0, 1):
        np.testing.assert_array_equal(out[i][[0, 1, 2]], np.array([0.035, 0.5, 0.5]))
        assert -0.72 < out[i][3] <= 0.0 and np.all(out[i][4:] == 0.0)
    np.testing.assert_array_equal(out[2][:3], [0.035, 0.5, 0.5])
    assert out[2][3] == pytest.approx(-0.5 + DTF, rel=1e-15)
    o.close()


@pytest.mark.parametrize("nozzle", [0, 1])
def test_respawn_launches_a_fresh_droplet(nozzle):
    wf = 0.6
    o = _oracle(count=4, water=True, flow=wf, nozzle=nozzle)
    _gas(o)
    recs = {1: [0.3, 0.3, 0.3, -DT / 2, 0.0, 0.0, 0.0, 0.0]}   # wait reaches >= 0 this step
    out = _integrate(o, _particles(4, recs))[1]
    speed = np.linalg.norm(out[4:7])
    variation = 0.24 if nozzle else 0.05
    base = 0.82 + 0.82 * wf
    assert base * (1 - variation / 2) * (1 - 1e-12) <= speed <= base * (1 + variation / 2) * (1 + 1e-12)
    assert out[3] == 0.0 and out[7] == pytest.approx(0.52 + 0.88 * wf)
    offset = (out[:3] - [0.035, 0.5, 0.5]) * EXTENT
    assert np.linalg.norm(offset) <= (0.006 if nozzle else 0.0028) * (1 + 1e-12)
    o.close()


# ------------------------------------------------------------------- deposition
def _run_deposit(nozzle, current, previous):
    """depositParticles with `current` as the new particle state and
    `previous` as the one before the integration."""
    o = _oracle(count=current.shape[0], nozzle=nozzle)
    o.write_particle_buffers(current, previous)
    nx, ny, nz = DIMS
    o.write(orc.FIELD_LIQUID, np.zeros((nz, ny, nx)))
    o.kernel_droplets(0.0, orc.DROPLET_DEPOSIT)
    liq = o.read(orc.FIELD_LIQUID)
    o.close()
    return liq


def _cell_center(i, n):
    return (i + 0.5) / n


def test_laminar_deposit_at_a_cell_centre_and_known_fractions():
    n = DIMS[0]
    mass = 0.9
    cur = np.zeros((2, 8))
    cur[0] = [_cell_center(5, n), _cell_center(6, n), _cell_center(7, n), 0.1, 0, 0, 0, mass]
    fx, fy, fz = 0.25, 0.5, 0.75
    cur[1] = [(20 + 0.5 + fx) / n, (10 + 0.5 + fy) / n, (12 + 0.5 + fz) / n, 0.1, 0, 0, 0, mass]
    liq = _run_deposit(0, cur, cur * [1, 1, 1, 1, 1, 1, 1, 0])  # previous inactive: no sweep
    fixed = lambda w: math.floor(w * mass * 0.2 * 4096)
    assert liq[7, 6, 5] == fixed(1.0)
    expect = {}
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (fx if dx else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                expect[(12 + dz, 10 + dy, 20 + dx)] = fixed(w)
    for k, v in expect.items():
        assert liq[k] == v, (k, liq[k], v)
    assert liq.sum() == fixed(1.0) + sum(expect.values())


def test_spray_deposit_kernel_at_a_cell_centre():
    n = DIMS[0]
    mass = 1.2
    cur = np.zeros((1, 8))
    cur[0] = [_cell_center(9, n), _cell_center(11, n), _cell_center(13, n), 0.1, 0, 0, 0, mass]
    liq = _run_deposit(1, cur, np.zeros((1, 8)))
    w = {0: 1.0}
    for k, d in ((1, 1.0), (2, math.sqrt(2)), (3, math.sqrt(3))):
        w[k] = max(0.0, 1 - 0.52 * d) ** 2
    counts = {0: 1, 1: 6, 2: 12, 3: 8}
    total = sum(w[k] * counts[k] for k in w)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                k = abs(dx) + abs(dy) + abs(dz)
                want = math.floor(w[k] / total * mass * 4096)
                assert abs(int(liq[13 + dz, 11 + dy, 9 + dx]) - want) <= 1
    assert mass * 4096 - 27 <= liq.sum() <= mass * 4096


def test_sweep_deposits_three_samples_and_outside_deposits_nothing():
    n = DIMS[0]
    mass = 0.9
    prev = np.zeros((2, 8))
    cur = np.zeros((2, 8))
    prev[0] = [_cell_center(4, n), _cell_center(8, n), _cell_center(8, n), 0.1, 0, 0, 0, mass]
    cur[0] = [_cell_center(8, n), _cell_center(8, n), _cell_center(8, n), 0.2, 0, 0, 0, mass]
    cur[1] = [1.01, 0.5, 0.5, 0.2, 0, 0, 0, mass]   # outside: nothing
    prev[1] = cur[1]
    liq = _run_deposit(0, cur, prev)
    third = math.floor(mass / 3 * 0.2 * 4096)
    for x in (4, 6, 8):
        assert liq[8, 8, x] == third
    assert liq.sum() == 3 * third


# --------------------------------------------------------------- step graph
def test_water_branch_of_the_step_graph():
    o = _oracle(precision=32, count=64, water=False)
    o.step(DT, 1)
    assert "sph" not in o.pass_log() and "clear_liquid" not in o.pass_log()
    o.set_bool("waterEnabled", True)
    o.step(DT, 3)
    assert o.pass_log().startswith("clear_liquid;sph_density;sph_integrate;liquid_deposit;simulate;")
    assert o.read(orc.FIELD_LIQUID).sum() > 0
    p = o.read(orc.FIELD_PARTICLES)
    assert np.all(p[:, 7] > 0)
    o.set_bool("waterEnabled", False)
    o.step(DT, 1)
    assert o.pass_log().startswith("clear_liquid;particle_reset;simulate;")
    assert o.read(orc.FIELD_LIQUID).sum() == 0
    assert np.all(o.read(orc.FIELD_PARTICLES)[:, 4:] == 0)
    o.step(DT, 1)
    assert o.pass_log().startswith("simulate;")
    o.close()


def test_fp32_and_fp64_droplets_agree_without_respawn():
    """A burning plume with droplets in it for a few steps: the FP32 pass
    results stay within f32 rounding of the FP64 ones (no droplet expires,
    so no hash-seeded respawn enters)."""
    res = {}
    for prec in (64, 32):
        o = _oracle(precision=prec, count=128, water=True, flow=0.7, nozzle=1)
        o.set_bool("cinematic", True)
        o.step(DT, 4)
        res[prec] = (o.read(orc.FIELD_PARTICLES), o.read(orc.FIELD_PARTICLE_DENSITY))
        o.close()
    p64, d64 = res[64]
    p32, d32 = res[32]
    alive = p64[:, 7] > 1e-4
    assert alive.sum() > 64
    np.testing.assert_allclose(p32[alive], p64[alive], rtol=2e-4, atol=2e-5)
    np.testing.assert_allclose(d32[alive], d64[alive], rtol=2e-3, atol=2e-4)


# ------------------------------------------------- JS arithmetic of the host state
_JS_INITIAL_STATE = r"""
// writeParticleInitialState's arithmetic (FireXEngine.ts:1976-2034) in the
// reference's language, for Math.sin / Math.tan / Math.hypot and the
// Float32Array store as the JS engine evaluates them.
const [count, active, waterFlow, sprayAngle, aimHeight, nozzleType] = JSON.parse(process.argv[1]);
const W = 1.15, H = 1.35, D = 1.15;
const aimX = (0.52 - 0.035) * W, aimY = (aimHeight - 0.5) * H, aimLength = Math.hypot(aimX, aimY);
const base = [aimX / aimLength, aimY / aimLength, 0], tangent = [base[1], -base[0], 0];
const hash = (v) => { const h = Math.sin(v * 12.9898 + 1.317) * 43758.5453; return h - Math.floor(h); };
const out = new Float32Array(count * 8);
for (let i = 0; i < count; i += 1) {
  const o = i * 8, phase = i / count;
  const angle = hash(i * 3.771 + 4.91) * Math.PI * 2, radial = Math.sqrt(hash(i * 3.771 + 0.17));
  const spread = nozzleType >= 0.5 ? 0.72 : 0.1;
  const cone = Math.tan((sprayAngle * Math.PI) / 180) * radial * spread;
  const dx0 = tangent[0] * Math.cos(angle), dy0 = tangent[1] * Math.cos(angle), dz0 = -Math.sin(angle);
  let x = base[0] + dx0 * cone, y = base[1] + dy0 * cone, z = dz0 * cone;
  const len = Math.hypot(x, y, z); x /= len; y /= len; z /= len;
  const variation = nozzleType >= 0.5 ? 0.24 : 0.05;
  const speed = (0.82 + waterFlow * 0.82) * (1 - variation * 0.5 + hash(i * 3.771 + 9.37) * variation);
  const age = phase * 0.72, r0 = nozzleType >= 0.5 ? 0.006 : 0.0028;
  out[o] = 0.035 + (dx0 * radial * r0 + x * speed * age) / W;
  out[o + 1] = Math.max(0.012, 0.5 + (dy0 * radial * r0 + y * speed * age - 0.29 * age * age) / H);
  out[o + 2] = 0.5 + (dz0 * radial * r0 + z * speed * age) / D;
  out[o + 3] = active ? age : -phase * 0.72;
  out[o + 4] = active ? x * speed : 0;
  out[o + 5] = active ? y * speed - 0.58 * age : 0;
  out[o + 6] = active ? z * speed : 0;
  out[o + 7] = active ? 0.52 + waterFlow * 0.88 : 0;
}
process.stdout.write(Buffer.from(out.buffer).toString('base64'));
"""


@pytest.mark.parametrize("case", [(4096, 1, 0.8, 10.0, 0.42, 0), (1536, 1, 0.35, 37.0, 0.7, 1),
                                  (640, 0, 0.0, 10.0, 0.42, 0)])
def test_initial_state_matches_javascript_arithmetic(case):
    """The host-side state equals the same arithmetic run by the JS engine
    (Math.hypot, Float32Array rounding, Math.sin/cos/tan).  ECMAScript leaves
    the trigonometric functions implementation-approximated: the host uses
    the C library's, node 12's V8 an fdlibm port, and the two differ by one
    double ulp on rare arguments, which the hash's * 43758.5453 can carry to
    an adjacent f32.  So: bit-identical except for at most 0.1 % of values,
    each of those within one f32 ulp."""
    import base64
    import json
    import shutil
    import subprocess

    node = shutil.which("node")
    if node is None:
        pytest.skip("node is absent")
    count, active, wf, angle, aim, nozzle = case
    res = subprocess.run([node, "-e", _JS_INITIAL_STATE, json.dumps(list(case))], capture_output=True, text=True,
                         timeout=60, check=True)
    js = np.frombuffer(base64.b64decode(res.stdout), dtype=np.float32).reshape(count, 8)
    o = _oracle(precision=32, count=count, water=bool(active), flow=wf, nozzle=nozzle)
    o.set_parameter("sprayAngle", angle)
    o.set_parameter("aimHeight", aim)
    o.write_particle_initial_state(bool(active))
    ours = o.read(orc.FIELD_PARTICLES).astype(np.float32)
    o.close()
    a, b = ours.view(np.int32).astype(np.int64), js.view(np.int32).astype(np.int64)
    differ = a != b
    assert differ.mean() <= 1e-3, f"{int(differ.sum())} of {a.size} values differ"
    assert np.all(np.abs(a - b)[differ] <= 1), "a value differs by more than one f32 ulp"
