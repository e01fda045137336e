"""Device arithmetic helpers (GPU): the refined hardware reciprocal, square root, reciprocal square
root and inverse cube root that the face and viscous kernels use for positive operands, against
the exact values (numpy in 64-bit-mantissa long double), in ulps of the result, on 2^20
log-uniform operands over the range each form accepts.

The double forms are one Newton-Raphson step on the hardware reciprocal (relative error e0 of the
estimate goes to about e0^2, so below 4e-15 when e0 < 6e-8), one Goldschmidt step and a residual
correction on the hardware reciprocal square root (square root: within half an ulp plus the
residual correction's rounding), a Newton step on that (reciprocal square root), and a float
log2 / exp2 estimate with one float and two double Newton steps (inverse cube root). The float forms of the mixed precision are the hardware
instructions themselves (one ulp) and one float Newton step (inverse cube root)."""
import numpy as np
import pytest

import cfdx

pytestmark = pytest.mark.gpu

LD = np.longdouble


@pytest.fixture(scope="module")
def ctx():
    if cfdx.device_count() < 1:
        pytest.skip("no HIP device")
    return cfdx.Ctx(0)


def operands(lo, hi, n, seed):
    """log-uniform positive operands in [lo, hi], with powers of two and values next to 1"""
    rng = np.random.default_rng(seed)
    x = np.exp(rng.uniform(np.log(lo), np.log(hi), n))
    special = [1.0, 2.0, 4.0, 0.5, 0.25, 3.0, 10.0, np.nextafter(1.0, 2.0), np.nextafter(1.0, 0.0),
               np.nextafter(2.0, 0.0), 1.4, 0.4, 1e-7, 1e7]
    special = [s for s in special if lo <= s <= hi]
    return np.concatenate([x, np.array(special)])


EXACT = {
    "pinv": lambda x: LD(1) / x,
    "psqrt": lambda x: np.sqrt(x),
    "prsq": lambda x: LD(1) / np.sqrt(x),
    "prcbrt": lambda x: LD(1) / np.cbrt(x),
}


def ulps(y, exact, dtype):
    """|y - exact| in ulps of exact rounded to dtype"""
    ref = exact.astype(dtype)
    return (np.abs(y.astype(LD) - exact) / np.spacing(np.abs(ref)).astype(LD)).astype(np.float64)


@pytest.mark.parametrize("fn,lo,hi,max_ulp", [
    ("pinv", 1e-300, 1e300, 20.0),
    ("psqrt", 1e-300, 1e300, 0.51),
    ("prsq", 1e-300, 1e300, 2.0),
    ("prcbrt", 1e-36, 1e36, 1.5),
])
def test_double_helpers(ctx, fn, lo, hi, max_ulp):
    x = operands(lo, hi, 1 << 20, 11)
    y = ctx.math_eval(fn, x, "fp64")
    exact = EXACT[fn](x.astype(LD))
    u = ulps(y, exact, np.float64)
    rel = np.max(np.abs(y.astype(LD) - exact) / exact)
    print(f"{fn} fp64: max {u.max():.3f} ulp, mean {u.mean():.3f} ulp, max rel {float(rel):.3e}, "
          f"correctly rounded {np.mean(y == exact.astype(np.float64)):.6f}")
    assert np.all(np.isfinite(y))
    assert u.max() <= max_ulp
    assert u.mean() <= 1.0
    if fn == "pinv":
        # one Newton step from an estimate of relative error below 6e-8
        assert rel <= 4e-15
    if fn == "psqrt":
        # correctly rounded for nearly every operand: a wrong last bit only where the exact value
        # lies next to the midpoint between two doubles
        assert np.mean(y == exact.astype(np.float64)) >= 0.999


@pytest.mark.parametrize("fn,max_ulp", [("pinv", 1.0), ("psqrt", 1.0), ("prsq", 1.0), ("prcbrt", 2.0)])
def test_float_helpers(ctx, fn, max_ulp):
    x = operands(1e-36, 1e36, 1 << 20, 12)
    xf = x.astype(np.float32)
    y = ctx.math_eval(fn, x, "fp32")
    assert np.all(y == y.astype(np.float32).astype(np.float64))   # results are floats
    exact = EXACT[fn](xf.astype(LD))
    u = ulps(y, exact, np.float32)
    print(f"{fn} fp32: max {u.max():.3f} ulp, mean {u.mean():.3f} ulp")
    assert np.all(np.isfinite(y))
    assert u.max() <= max_ulp


def test_math_eval_arguments(ctx):
    assert ctx.math_eval("pinv", np.zeros(0)).size == 0
    with pytest.raises(ValueError):
        ctx.math_eval("exp", np.ones(4))
    with pytest.raises(cfdx.CfdxError):
        cfdx._check(cfdx._lib.cfdx_math_eval(ctx.h, 9, None, None, 4))
