"""Accuracy of the fixed-sequence sin/cos/pow/log/exp/atan2 (oracle/fmath.h; the kernels
evaluate the same sequences, so renders agree bit for bit).

Each function is compared with the correctly rounded binary32 value of the
exact result (float64 libm, then rounded): the error bound is a few ulp over
the ranges the renderer uses, and the special values are exact."""
import numpy as np
import pytest


def ulp_err(got, exact64):
    want = exact64.astype(np.float32)
    spacing = np.spacing(np.abs(want)).astype(np.float64)
    return np.abs(got.astype(np.float64) - exact64) / spacing


@pytest.mark.parametrize("name,fn", [("sin", np.sin), ("cos", np.cos)])
def test_sincos_ulp(oracle_mod, name, fn):
    rng = np.random.default_rng(1)
    # the renderer's range: 2 pi u for u in [0, 1), plus wider arguments
    x = np.concatenate([np.linspace(0, 2 * np.pi, 200001), rng.uniform(-1000, 1000, 200000),
                        rng.uniform(-1e-3, 1e-3, 1000)]).astype(np.float32)
    got = oracle_mod.fmath(name, x)
    exact = fn(x.astype(np.float64))
    # absolute error near the zeros, relative (ulp) elsewhere
    err = np.abs(got.astype(np.float64) - exact)
    ok = (ulp_err(got, exact) <= 4) | (err <= 1.2e-7)
    assert ok.all(), (x[~ok][:5], got[~ok][:5], exact[~ok][:5])
    assert (np.abs(got) <= 1.0).all()


def test_sincos_special_values(oracle_mod):
    assert oracle_mod.fmath("sin", [0.0])[0] == 0.0
    assert oracle_mod.fmath("cos", [0.0])[0] == 1.0
    s = oracle_mod.fmath("sin", np.float32([np.pi / 2, np.pi, 3 * np.pi / 2]))
    assert abs(s[0] - 1) < 1e-7 and abs(s[1]) < 2e-7 and abs(s[2] + 1) < 1e-7
    # sin^2 + cos^2 = 1 to rounding
    x = np.linspace(0, 6.3, 5001, dtype=np.float32)
    s, c = oracle_mod.fmath("sin", x), oracle_mod.fmath("cos", x)
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 3e-7


def test_log2_exp2_ulp(oracle_mod):
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.uniform(1e-6, 1, 100000), rng.uniform(1, 1e6, 100000)]).astype(np.float32)
    got = oracle_mod.fmath("log2", x)
    exact = np.log2(x.astype(np.float64))
    ok = (ulp_err(got, exact) <= 4) | (np.abs(got - exact) <= 2e-7)
    assert ok.all()
    assert oracle_mod.fmath("log2", np.float32([1.0, 2.0, 0.5, 1024.0])).tolist() == [0.0, 1.0, -1.0, 10.0]
    v = rng.uniform(-100, 100, 100000).astype(np.float32)
    got = oracle_mod.fmath("exp2", v)
    assert (ulp_err(got, np.exp2(v.astype(np.float64))) <= 4).all()
    assert oracle_mod.fmath("exp2", np.float32([0.0, 1.0, -3.0])).tolist() == [1.0, 2.0, 0.125]


@pytest.mark.parametrize("y", [0.0, 1.0, 2.0, 5.0, 7.0, -2.0, 0.5, 1.7, 3.3, 4.999])
def test_pow_ulp(oracle_mod, y):
    x = np.linspace(0, 1, 100001, dtype=np.float32)
    if y < 0:
        x = x[1:]
    got = oracle_mod.fmath("pow", x, np.float32(y))
    exact = np.power(x.astype(np.float64), y)
    tiny = np.abs(exact) < 1e-30
    # exp2(y log2 x) in binary32: the rounding of y log2 x scales the error by
    # |y log2 x| (integral y is exact square-and-multiply: a few ulp)
    with np.errstate(divide="ignore", invalid="ignore"):
        amp = 1.0 + np.nan_to_num(np.abs(y * np.log2(x.astype(np.float64))), posinf=0.0)
    ok = (ulp_err(got, exact) <= 4 * amp) | tiny
    assert ok.all(), (x[~ok][:3], got[~ok][:3], exact[~ok][:3])
    if y == 0.0:
        assert (got == 1.0).all()
    if y == 1.0:
        assert np.array_equal(got, x)


def test_log_exp_ulp(oracle_mod):
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(1e-6, 1, 100000), rng.uniform(1, 1e4, 100000)]).astype(np.float32)
    got = oracle_mod.fmath("log", x)
    exact = np.log(x.astype(np.float64))
    ok = (ulp_err(got, exact) <= 6) | (np.abs(got - exact) <= 3e-7)
    assert ok.all(), (x[~ok][:3], got[~ok][:3], exact[~ok][:3])
    v = rng.uniform(-60, 60, 100000).astype(np.float32)
    got = oracle_mod.fmath("exp", v)
    # the product x * log2(e) in binary32 scales the error by |x log2 e|
    amp = 1.0 + np.abs(v.astype(np.float64)) * 1.4427
    assert (ulp_err(got, np.exp(v.astype(np.float64))) <= 4 * amp).all()
    assert oracle_mod.fmath("exp", [0.0])[0] == 1.0 and oracle_mod.fmath("log", [1.0])[0] == 0.0


def test_atan2(oracle_mod):
    rng = np.random.default_rng(4)
    y = np.concatenate([rng.normal(size=200000), rng.uniform(-1e-3, 1e-3, 1000)]).astype(np.float32)
    x = np.concatenate([rng.normal(size=200000), rng.uniform(-1e-3, 1e-3, 1000)]).astype(np.float32)
    got = oracle_mod.fmath("atan2", y, x)
    exact = np.arctan2(y.astype(np.float64), x.astype(np.float64))
    ok = (ulp_err(got, exact) <= 4) | (np.abs(got - exact) <= 2.5e-7)
    assert ok.all(), (y[~ok][:3], x[~ok][:3], got[~ok][:3], exact[~ok][:3])
    # IEEE special cases of the signed zeros and the axes
    zs = np.float32([0.0, -0.0, 0.0, -0.0, 1.0, -1.0, 0.0, 0.0])
    xs = np.float32([0.0, 0.0, -0.0, -0.0, 0.0, 0.0, 1.0, -1.0])
    want = np.arctan2(zs.astype(np.float64), xs.astype(np.float64)).astype(np.float32)
    got = oracle_mod.fmath("atan2", zs, xs)
    assert np.array_equal(np.signbit(got), np.signbit(want)) and np.allclose(got, want, atol=0, rtol=1e-7)


@pytest.mark.parametrize("name,fn", [("asin", np.arcsin), ("acos", np.arccos)])
def test_asin_acos_ulp(oracle_mod, name, fn):
    """asin and acos through atan2 of the triangle's legs: within 4 ulp (or
    2e-7 absolute near asin's zero) over [-1, 1], exact at the ends, and the
    argument clamped to [-1, 1] (a normalised direction may exceed 1 by an
    ulp)."""
    rng = np.random.default_rng(7)
    x = np.concatenate([np.linspace(-1, 1, 200001), rng.uniform(-1, 1, 100000),
                        1 - rng.uniform(0, 1e-3, 10000), -1 + rng.uniform(0, 1e-3, 10000)]).astype(np.float32)
    got = oracle_mod.fmath(name, x)
    exact = fn(x.astype(np.float64))
    ok = (ulp_err(got, exact) <= 4) | (np.abs(got - exact) <= 2e-7)
    assert ok.all(), (x[~ok][:5], got[~ok][:5], exact[~ok][:5])
    ends = oracle_mod.fmath(name, np.float32([-1.0, 0.0, 1.0, 1.0000001, -1.0000001]))
    want = np.float32(fn(np.float64([-1.0, 0.0, 1.0, 1.0, -1.0])))
    assert np.abs(ends - want).max() <= 2.4e-7
