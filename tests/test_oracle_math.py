"""Pins the f32 WGSL built-in layer (fire_amd/csrc/fx_math.h) used by the FP32
oracle and the gfx950 kernels: error bounds against float64 numpy and exact
special values.  The f16/bf16 storage rounding of the oracle is checked
against numpy's float16 and torch's bfloat16 conversions."""
import numpy as np
import pytest

from oracle import oracle as orc


def ulp32(v):
    v = np.abs(np.asarray(v, dtype=np.float32))
    return np.spacing(np.maximum(v, np.float32(np.finfo(np.float32).tiny)))


def test_exp_error_bound_and_specials():
    x = np.linspace(-85.9, 88.7, 400_001, dtype=np.float32)
    got = orc.f32_eval("exp", x).astype(np.float64)
    want = np.exp(x.astype(np.float64))
    err = np.abs(got - want) / ulp32(want)
    assert err.max() <= 1.5
    s = orc.f32_eval("exp", np.array([0.0, -0.0, 100.0, -100.0, np.nan], dtype=np.float32))
    assert s[0] == 1.0 and s[1] == 1.0 and np.isinf(s[2]) and s[3] == 0.0 and np.isnan(s[4])


def test_exp2_error_bound_and_integers():
    x = np.linspace(-124.9, 127.9, 400_001, dtype=np.float32)
    got = orc.f32_eval("exp2", x).astype(np.float64)
    want = np.exp2(x.astype(np.float64))
    assert (np.abs(got - want) / ulp32(want)).max() <= 1.5
    k = np.arange(-125, 128, dtype=np.float32)
    np.testing.assert_array_equal(orc.f32_eval("exp2", k), np.exp2(k.astype(np.float64)).astype(np.float32))


def test_log2_error_bound_and_powers_of_two():
    x = np.exp(np.linspace(np.log(1e-30), np.log(1e30), 400_001)).astype(np.float32)
    got = orc.f32_eval("log2", x).astype(np.float64)
    want = np.log2(x.astype(np.float64))
    assert np.abs(got - want).max() <= 4e-6
    k = np.arange(-120, 120, dtype=np.float64)
    np.testing.assert_array_equal(orc.f32_eval("log2", np.exp2(k).astype(np.float32)), k.astype(np.float32))
    s = orc.f32_eval("log2", np.array([0.0, -1.0], dtype=np.float32))
    assert np.isneginf(s[0]) and np.isnan(s[1])


@pytest.mark.parametrize("fn", ["sin", "cos"])
def test_sin_cos_error_bound(fn):
    x = np.linspace(-1500.0, 1500.0, 600_001, dtype=np.float32)
    got = orc.f32_eval(fn, x).astype(np.float64)
    want = getattr(np, fn)(x.astype(np.float64))
    assert np.abs(got - want).max() <= 1.5e-7


def test_sin_cos_identities():
    x = np.linspace(-50, 50, 100_001, dtype=np.float32)
    s = orc.f32_eval("sin", x).astype(np.float64)
    c = orc.f32_eval("cos", x).astype(np.float64)
    assert np.abs(s * s + c * c - 1).max() < 5e-7
    z = orc.f32_eval("sin", np.array([0.0, -0.0], dtype=np.float32))
    assert z[0] == 0.0 and z[1] == 0.0  # the sign of a zero result is not specified by WGSL
    assert orc.f32_eval("cos", np.array([0.0], dtype=np.float32))[0] == 1.0


@pytest.mark.parametrize("lo,hi,y", [(1e-5, 3.0, -0.3), (1e-5, 3.0, 1.3), (300.0, 2900.0, 4.0)])
def test_pow_relative_error(lo, hi, y):
    # WGSL: pow(x, y) inherits the error of exp2(y * log2(x)).
    x = np.exp(np.linspace(np.log(lo), np.log(hi), 200_001)).astype(np.float32)
    yy = np.full_like(x, y, dtype=np.float32)
    got = orc.f32_eval("pow", x, yy).astype(np.float64)
    want = np.power(x.astype(np.float64), np.float64(np.float32(y)))
    assert (np.abs(got - want) / want).max() <= 4e-6


def test_smoothstep_shape():
    x = np.linspace(-1, 2, 30001, dtype=np.float32)
    y = np.tile(np.array([0.0, 1.0], dtype=np.float32), x.size)
    s = orc.f32_eval("smoothstep", x, y)
    assert s[x <= 0].max() == 0.0 and s[x >= 1].min() == 1.0
    assert np.all(np.diff(s.astype(np.float64)) >= 0)
    mid = orc.f32_eval("smoothstep", np.array([0.5], np.float32), np.array([0.0, 1.0], np.float32))
    assert mid[0] == 0.5


def test_f16_rounding_matches_numpy():
    rng = np.random.default_rng(0)
    v = np.concatenate([
        rng.uniform(-3, 3, 50_000), rng.uniform(-1e-4, 1e-4, 20_000), rng.uniform(-70000, 70000, 5000),
        np.array([0.0, -0.0, 65504.0, 65519.0, 65520.0, 6.1e-5, 5.96e-8, 2.98e-8, 1e-9]),
    ]).astype(np.float32)
    lib = orc.load()
    got = np.array([lib.orc_round_f16(float(t)) for t in v], dtype=np.float32)
    want = v.astype(np.float16).astype(np.float32)
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(1)
    v = np.concatenate([rng.uniform(-3, 3, 50_000), rng.uniform(-1e-30, 1e-30, 5000),
                        np.array([0.0, -0.0, 3.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8])]).astype(np.float32)
    lib = orc.load()
    got = np.array([lib.orc_round_bf16(float(t)) for t in v], dtype=np.float32)
    want = torch.from_numpy(v).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(got.view(np.uint32), want.view(np.uint32))


def test_seeded_uniform_statistics():
    lib = orc.load()
    u = np.array([lib.orc_seeded_uniform(99, f, c, s) for f in range(4) for c in range(2500) for s in range(4)])
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01
    assert abs(u.var() - 1 / 12) < 0.005
    # streams and frames give different values
    a = np.array([lib.orc_seeded_uniform(99, 0, c, 0) for c in range(1000)])
    b = np.array([lib.orc_seeded_uniform(99, 1, c, 0) for c in range(1000)])
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.1
