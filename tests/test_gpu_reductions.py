"""The FP64 reductions on the device (firex_reduce) against float64 numpy over
the fields read back, and the residual norms' behaviour across Jacobi sweeps.

The reductions are the quantities the reference's notes ask a solver to
report (docs/implementation-notes.md:223-235: residual norms, mass and heat
budgets); the reference itself computes none.  Each device result must equal
numpy's float64 value of the same definition to 1e-12 relative: the per-cell
terms are the same float64 operations, only the summation order differs.
"""
import numpy as np
import pytest

from fire_amd import FireXError, FireXHipSolver, REDUCTIONS, device_count, seeded_emitters

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
DOMAIN = (1.15, 1.35, 1.15)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if device_count() < 1:
        pytest.fail("no HIP device visible to the reduction tests")


def metric(dims):
    """cellMetric in f32 (shaders.ts:39-41) and the Jacobi weights metric^2."""
    m = np.float32(dims) / np.float32(DOMAIN)
    return m.astype(np.float64), (m * m).astype(np.float32).astype(np.float64)


def clamp_index(n):
    return np.r_[0, np.arange(n - 1)], np.r_[np.arange(1, n), n - 1]


def residual_l2(p, div, w):
    """||w (neighbour sums) - 2 (wx + wy + wz) p - div|| over the non-top cells,
    neighbours clamped at the walls and 0 above the top (shaders.ts:366-403)."""
    p = p.astype(np.float64)
    div = div.astype(np.float64)
    nz, ny, nx = p.shape
    xm, xp = clamp_index(nx)
    ym, _ = clamp_index(ny)
    zm, zp = clamp_index(nz)
    above = np.concatenate([p[:, 1:, :], np.zeros((nz, 1, nx))], axis=1)
    wx, wy, wz = w
    nb = wx * (p[:, :, xm] + p[:, :, xp]) + wy * (p[:, ym, :] + above) + wz * (p[zm] + p[zp])
    r = nb - 2.0 * (wx + wy + wz) * p - div
    r[:, ny - 1, :] = 0.0
    return float(np.sqrt(np.sum(r * r)))


def flow_divergence_l2(flow, m):
    """||computeDivergence(flow)|| in float64 (shaders.ts:344-364)."""
    f = flow.astype(np.float64)
    nz, ny, nx = f.shape[:3]
    xm, _ = clamp_index(nx)
    ym, _ = clamp_index(ny)
    zm, _ = clamp_index(nz)
    u, v, w = f[..., 0], f[..., 1], f[..., 2]
    d = (u - u[:, :, xm]) * m[0] + (v - v[:, ym, :]) * m[1] + (w - w[zm]) * m[2]
    return float(np.sqrt(np.sum(d * d)))


def burning_solver(dims, steps, storage="f16", heavy="off", iters=(96, 96)):
    g = FireXHipSolver(dims, tier="uhd", pressure_iterations=iters[0], correction_iterations=iters[1],
                       species_storage=storage, heavy_mode=heavy, plan="fused")
    em = seeded_emitters(dims, 3, radius_range=(6.0, 14.0), seed=2026)
    g.set_emitters(em, jitter=0.1, velocity_noise=0.0, seed=2026)
    g.step(DT, steps)
    return g


@pytest.mark.parametrize("heavy,storage", [("off", "f16"), ("f16", "bf16")])
def test_reductions_match_numpy_fp64_at_uhd(heavy, storage):
    """All eight FIREX_REDUCE_* outputs at 192 x 288 x 192 after 40 steps of the
    bench workload (f32 pressure) and on the f16-pressure path."""
    dims = (192, 288, 192)
    print(f"reductions: {dims} heavy={heavy} species={storage}", flush=True)
    g = burning_solver(dims, 40, storage=storage, heavy=heavy)
    assert g.info["f16_pressure"] == (heavy == "f16")
    flow = g.read_field("flow")
    species = g.read_field("species")
    m, w = metric(dims)
    want = {
        "sum_fuel": np.sum(species[..., 0], dtype=np.float64),
        "sum_oxygen": np.sum(species[..., 1], dtype=np.float64),
        "sum_products": np.sum(species[..., 2], dtype=np.float64),
        "sum_vapor": np.sum(species[..., 3], dtype=np.float64),
        "sum_temperature": np.sum(flow[..., 3], dtype=np.float64),
        "sum_reaction": np.sum(g.read_field("reaction"), dtype=np.float64),
        "residual_l2": residual_l2(g.read_field("correction"), g.read_field("divergence"), w),
        "flow_divergence_l2": flow_divergence_l2(flow, m),
    }
    assert want["sum_reaction"] > 0 and want["sum_products"] > 0 and want["residual_l2"] > 0
    for name in REDUCTIONS:
        got = g.reduce(name)
        assert got == pytest.approx(want[name], rel=1e-12), name
    g.close()


def _step_once(dims, iters):
    g = burning_solver(dims, 1, storage="f32", iters=iters)
    r = g.reduce("residual_l2")
    g.close()
    return r


def test_primary_residual_decreases_with_sweeps():
    """With no residual sweeps the correction field is 0, so the residual
    norm is ||div(primary-projected flow)||, which equals the primary solve's
    residual ||div u* - L p|| on the non-top cells (D(G p) = L p).  From the
    same state it must fall with every added primary sweep."""
    dims = (64, 96, 64)
    sweeps = [0, 1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96]
    res = [_step_once(dims, (n, 0)) for n in sweeps]
    print("primary residual:", dict(zip(sweeps, res)), flush=True)
    assert res[0] > 0
    assert all(b < a for a, b in zip(res, res[1:])), res


def test_correction_residual_decreases_with_sweeps():
    """The residual solve starts from 0 on the same right-hand side whatever
    its sweep count (the primary solve is fixed): its norm falls with every
    added sweep (the Jacobi residual map N / (2 sum w) has norm below 1)."""
    dims = (64, 96, 64)
    sweeps = [0, 1, 2, 3, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96]
    res = [_step_once(dims, (24, m)) for m in sweeps]
    print("correction residual:", dict(zip(sweeps, res)), flush=True)
    assert res[0] > 0
    assert all(b < a for a, b in zip(res, res[1:])), res


def test_residual_refused_after_the_divergence_view():
    """The divergence view's refresh (firex_pack_fields with viewMode 7)
    replaces the residual's right-hand side; the residual norm is refused
    until the next step, and the other reductions still work."""
    g = burning_solver((64, 96, 64), 3, iters=(24, 24))
    before = g.reduce("residual_l2")
    g.set_parameter("viewMode", 7)
    g.pack_fields()
    with pytest.raises(FireXError, match="right-hand side"):
        g.reduce("residual_l2")
    assert g.reduce("flow_divergence_l2") > 0
    g.step(DT, 1)
    assert g.reduce("residual_l2") > 0 and before > 0
    g.close()
