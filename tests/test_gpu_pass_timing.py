"""Per-pass device timing (firex_set_pass_timing / firex_get_pass_timing), the
figures the benchmark's per-pass rates and roofline are computed from.

For a known plan every pass kind must report exactly the launches the step
encodes, a positive device time for each kind that ran, and nothing for the
kinds that did not; switching timing on or off clears the totals, and steps
taken with timing off add nothing to them.  The step sequence counted is
encodeSimulationStep's (FireXEngine.ts:2172-2275): simulate, divergence, the
pressure solve, projection, divergence of the projected flow, the residual
solve, projection (and, on the reference plan, the residual field's two
clears and copyFlow).
"""
import pytest

from fire_amd import FireXHipSolver, device_count

pytestmark = pytest.mark.gpu

DT = 1.0 / 60.0
DIMS = (64, 48, 32)
ITERS = 12
STEPS = 3


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if device_count() < 1:
        pytest.fail("no HIP device visible to the pass-timing tests")


def timed_steps(solver, steps):
    solver.set_pass_timing(True)
    solver.step(DT, steps)
    t = solver.pass_timing()
    solver.set_pass_timing(False)
    return t


def expected_launches(plan, sweeps):
    jacobi = 2 * (-(-ITERS // sweeps)) if plan == "fused" else 2 * ITERS
    fill_copy = 0 if plan == "fused" else 3
    per_step = {"simulate": 1, "divergence": 2, "jacobi": jacobi, "project": 2, "fill_copy": fill_copy}
    return {k: v * STEPS for k, v in per_step.items()}


@pytest.mark.parametrize("plan,sweeps,graphs", [("fused", 3, False), ("fused", 5, True), ("reference", 1, False)])
def test_pass_launch_counts_and_times(plan, sweeps, graphs):
    print(f"pass timing: {DIMS} plan {plan}, {sweeps} sweeps per launch, graphs {graphs}", flush=True)
    kw = {"jacobi_sweeps_per_launch": sweeps} if plan == "fused" else {}
    with FireXHipSolver(DIMS, tier="hd", pressure_iterations=ITERS, correction_iterations=ITERS, plan=plan,
                        capture_graphs=graphs, **kw) as s:
        s.step(DT, 2)  # warm start and (with graphs) a captured step before timing
        t = timed_steps(s, STEPS)
        want = expected_launches(plan, sweeps)
        for name, n in want.items():
            assert t[name]["launches"] == n, (name, t[name], n)
            if n:
                assert t[name]["ms"] > 0.0, (name, t[name])
            else:
                assert t[name]["ms"] == 0.0, (name, t[name])
        for name in ("project_divergence", "pack", "droplets", "render", "present", "streaks"):
            assert t[name] == {"ms": 0.0, "launches": 0}, (name, t[name])
        assert t["steps"] == STEPS
        assert t["jacobi_sweeps"] == 2 * ITERS * STEPS

        # switching timing off clears the totals, and untimed steps add nothing
        s.step(DT, 2)
        off = s.pass_timing()
        assert off["steps"] == 0 and off["jacobi_sweeps"] == 0
        assert all(off[name] == {"ms": 0.0, "launches": 0} for name in want), off
        # timing on again: the totals count from zero
        again = timed_steps(s, 1)
        assert again["steps"] == 1
        assert again["simulate"]["launches"] == 1
        assert again["jacobi"]["launches"] == want["jacobi"] // STEPS
