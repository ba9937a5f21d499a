"""Pins of the oracle's packFields restatement (qualityShaders.ts:37-73 and
the presentation divergence refresh, FireXEngine.ts:2590-2597) by facts that
do not come from its own code: endpoints and clamps of the temperature
normalisation, soot closed forms at known richness, the liquid scaling,
numpy float16 rounding of the texel components, and, in the divergence view,
a numpy backward-difference divergence of the written flow field."""
import numpy as np

from oracle import oracle as orc

DIMS = (12, 10, 8)
DT = 1.0 / 60.0


def _oracle(precision=32):
    o = orc.Oracle(DIMS, precision=precision, pressure_iterations=4, correction_iterations=4)
    o.step(DT, 1)  # uniforms written, metric known
    return o


def _smoothstep(e0, e1, x):
    t = np.clip((x - e0) / (e1 - e0), 0.0, 1.0)
    return t * t * (3.0 - 2.0 * t)


def _write_state(o, flow, species, reaction, liquid):
    o.write(orc.FIELD_FLOW, flow)
    o.write(orc.FIELD_SPECIES, species)
    o.write(orc.FIELD_REACTION, reaction)
    o.write(orc.FIELD_LIQUID, liquid)


def _blank():
    nx, ny, nz = DIMS
    flow = np.zeros((nz, ny, nx, 4))
    flow[..., 3] = 300.0
    species = np.zeros((nz, ny, nx, 4))
    return flow, species, np.zeros((nz, ny, nx)), np.zeros((nz, ny, nx))


def test_quiescent_state_packs_to_ambient_values():
    o = orc.Oracle(DIMS, precision=32, pressure_iterations=4, correction_iterations=4)
    pf, pc, pm = o.pack_fields()
    assert np.all(pf == 0.0)  # zero velocity, (300 - 300) / 2600 = 0
    oxygen = o.read(orc.FIELD_SPECIES)[..., 1]
    np.testing.assert_array_equal(pc[..., 1], oxygen.astype(np.float16).astype(np.float64))
    assert np.all(pc[..., [0, 2, 3]] == 0.0)
    assert np.all(pm == 0.0)  # no reaction, soot, liquid; smoothstep(420, 700, 300) = 0
    o.close()


def test_temperature_normalisation_endpoints_and_clamps():
    o = _oracle()
    flow, species, reaction, liquid = _blank()
    temps = [0.0, 300.0, 1600.0, 2900.0, 4000.0]
    for i, t in enumerate(temps):
        flow[0, 0, i, 3] = t
    _write_state(o, flow, species, reaction, liquid)
    pf, _, _ = o.pack_fields()
    np.testing.assert_array_equal(pf[0, 0, :5, 3], [0.0, 0.0, 0.5, 1.0, 1.0])
    o.close()


def test_soot_richness_closed_forms():
    o = _oracle()
    flow, species, reaction, liquid = _blank()
    # (fuel, O2, products): richness = fuel / (fuel + O2 / 4)
    cases = [(0.6, 0.0, 0.8),    # richness 1: soot = products
             (0.0, 0.5, 0.8),    # richness 0: no soot
             (0.1, 0.4, 0.8),    # richness 1/2: soot = products * smoothstep(0.25, 0.82, 0.5)
             (-0.3, 0.2, 0.5)]   # negative fuel reads as 0
    for i, (fu, ox, pr) in enumerate(cases):
        species[1, 2, i, :3] = (fu, ox, pr)
    _write_state(o, flow, species, reaction, liquid)
    _, pc, pm = o.pack_fields()
    soot = pm[1, 2, :4, 1]
    expect = [0.8, 0.0, 0.8 * _smoothstep(0.25, 0.82, 0.5), 0.0]
    np.testing.assert_allclose(soot, expect, rtol=2e-3, atol=0)
    assert pc[1, 2, 3, 0] == 0.0  # max(species, 0)
    o.close()


def test_liquid_reaction_and_display_value():
    o = _oracle()
    flow, species, reaction, liquid = _blank()
    liquid[2, 3, 4] = 4096          # 3.4 after scaling
    reaction[2, 3, 5] = 7.5         # clamped to 4
    reaction[2, 3, 6] = 0.3
    species[2, 3, 6, 3] = 0.9       # vapor above the reaction: display = vapor
    flow[2, 3, 7, 3] = 1000.0       # smoothstep(420, 700, 1000) = 1 dominates
    _write_state(o, flow, species, reaction, liquid)
    _, _, pm = o.pack_fields()
    assert pm[2, 3, 4, 2] == np.float64(np.float16(3.4))
    assert pm[2, 3, 4, 3] == pm[2, 3, 4, 2]
    assert pm[2, 3, 5, 0] == 4.0 and pm[2, 3, 5, 3] == 4.0
    assert pm[2, 3, 6, 3] == np.float64(np.float16(np.float32(0.9)))
    assert pm[2, 3, 7, 3] == 1.0
    o.close()


def test_texels_are_f16_rounded_species():
    rng = np.random.default_rng(3)
    o = _oracle()
    flow, species, reaction, liquid = _blank()
    species[...] = rng.uniform(-0.5, 3.0, species.shape)
    _write_state(o, flow, species, reaction, liquid)
    _, pc, _ = o.pack_fields()
    stored = o.read(orc.FIELD_SPECIES).astype(np.float32)  # f32 state the pass reads
    np.testing.assert_array_equal(pc, np.maximum(stored, 0.0).astype(np.float16).astype(np.float64))
    o.close()


def test_divergence_view_shows_the_final_flow_divergence():
    rng = np.random.default_rng(5)
    o = _oracle()
    flow, species, reaction, liquid = _blank()
    flow[..., :3] = rng.uniform(-1.0, 1.0, flow[..., :3].shape)
    _write_state(o, flow, species, reaction, liquid)
    o.set_parameter("viewMode", 7)
    o.step(0.0, 0)
    _, _, pm = o.pack_fields()
    f = o.read(orc.FIELD_FLOW).astype(np.float32).astype(np.float64)
    metric = np.array(DIMS, dtype=np.float64) / np.array([1.15, 1.35, 1.15])
    # backward differences with the clamped (zero-difference) first cell
    dx = np.diff(f[..., 0], axis=2, prepend=f[:, :, :1, 0])
    dy = np.diff(f[..., 1], axis=1, prepend=f[:, :1, :, 1])
    dz = np.diff(f[..., 2], axis=0, prepend=f[:1, :, :, 2])
    div = dx * metric[0] + dy * metric[1] + dz * metric[2]
    inner = pm[1:-1, 1:-1, 1:-1, 3]
    np.testing.assert_allclose(inner, div[1:-1, 1:-1, 1:-1], rtol=2e-3, atol=2e-3)
    shell = np.ones(pm.shape[:3], bool)
    shell[1:-1, 1:-1, 1:-1] = False
    assert np.all(pm[shell][:, 3] == 0.0)
    np.testing.assert_array_equal(o.read(orc.FIELD_DIVERGENCE)[1:-1, 1:-1, 1:-1].astype(np.float16).astype(np.float64),
                                  inner)
    o.close()
