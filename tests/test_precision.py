"""Per-field error of the f32 step and of reduced-precision species storage
against the FP64 oracle.

The HIP kernels equal the FP32 oracle bit for bit (test_gpu_parity), so the
error of the GPU path is the error of the FP32 oracle with the same storage.
Measured here on the CPU: relative L2 (||a - b|| / ||b||) and max-abs error
per field and channel after 40 steps of a seeded-emitter plume at 48x72x48
with 48 + 48 Jacobi sweeps, against the FP64 oracle (libm transcendentals,
true division, same discrete problem).

Bounds (written per storage below):
  f32   : every field and channel <= 2e-4 relative L2;
  f16   : every state field (flow, species, pressure) <= 1e-3; the reaction
          diagnostic (burn rate, shaders.ts:259-279; an output for rendering,
          not state) <= 3e-3 -- it amplifies the fuel / oxygen rounding through
          the Arrhenius term and is above the 1e-3 bar;
  bf16  : species channels between 1e-3 and 3e-2 -- bf16 does not meet the
          1e-3 bar.
"""
import numpy as np
import pytest

from fire_amd.solver import seeded_emitters
from oracle import oracle as orc

DIMS = (48, 72, 48)
STEPS = 40
ITERS = 48
DT = 1.0 / 60.0
FIELDS = (("flow", orc.FIELD_FLOW), ("species", orc.FIELD_SPECIES), ("pressure", orc.FIELD_PRESSURE),
          ("reaction", orc.FIELD_REACTION))


def _run(precision, storage):
    em = seeded_emitters(DIMS, 1, radius_range=(4 * 2 ** 0.5, 4 * 2 ** 0.5), seed=2026)
    o = orc.Oracle(DIMS, precision=precision, pressure_iterations=ITERS, correction_iterations=ITERS)
    o.set_bool("cinematic", True)
    o.set_heavy_mode(orc.HEAVY_OFF)
    if storage is not None:
        o.set_species_storage(storage)
        o.reset()
    o.set_emitters(em, 0.1, 0.0, 2026)
    o.step(DT, STEPS)
    out = {name: o.read(fid).astype(np.float64) for name, fid in FIELDS}
    o.close()
    return out


def _errors(a, b):
    """{field[.channel]: (relative L2, max abs)}"""
    out = {}
    for name, _ in FIELDS:
        x, y = a[name], b[name]
        chans = [(f"{name}.{'xyzw'[c]}", x[..., c], y[..., c]) for c in range(4)] if x.ndim == 4 else [(name, x, y)]
        for key, u, v in chans:
            out[key] = (float(np.linalg.norm(u - v) / max(np.linalg.norm(v), 1e-30)), float(np.max(np.abs(u - v))))
    return out


@pytest.fixture(scope="module")
def fp64():
    return _run(64, None)


@pytest.fixture(scope="module")
def errors(fp64):
    return {name: _errors(_run(32, st), fp64) for name, st in (
        ("f32", None), ("f16", orc.SPECIES_F16), ("bf16", orc.SPECIES_BF16))}


def _table(errs):
    return "\n".join(f"  {k:12s} rel L2 {r:.2e}  max abs {m:.2e}" for k, (r, m) in errs.items())


def test_the_plume_is_burning(fp64):
    """The comparison runs on a plume with fuel, burning and flow (not a
    quiescent field that any storage would reproduce)."""
    assert fp64["reaction"].max() > 0.05
    assert fp64["species"][..., 0].max() > 0.1
    assert np.abs(fp64["flow"][..., 1]).max() > 0.01


def test_f32_step_against_fp64(errors):
    errs = errors["f32"]
    print("\nf32 vs FP64\n" + _table(errs))
    assert all(r <= 2e-4 for r, _ in errs.values()), _table(errs)


def test_f16_species_against_fp64(errors):
    errs = errors["f16"]
    print("\nf16 species vs FP64\n" + _table(errs))
    state = {k: v for k, v in errs.items() if not k.startswith("reaction")}
    assert all(r <= 1e-3 for r, _ in state.values()), _table(errs)
    assert errs["reaction"][0] <= 3e-3, _table(errs)


def test_bf16_species_against_fp64(errors):
    errs = errors["bf16"]
    print("\nbf16 species vs FP64\n" + _table(errs))
    species = [r for k, (r, _) in errs.items() if k.startswith("species")]
    assert max(species) > 1e-3  # above the 1e-3 bar: stated, not hidden
    assert all(r <= 3e-2 for r, _ in errs.values()), _table(errs)
