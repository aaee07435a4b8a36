"""Accuracy of the column physics' transcendental routines on the device
(exo_math_eval), the counterparts of the reference's accurate f32 math
(shared/gpu/math/wgsl/precise.wgsl), against float64 numpy over the
argument ranges the physics gives them.

Bounds, in units in the last place of the float32 result:
* log2 on v_log_f32: <= 2 ulp where |log2 a| >= 1; its absolute error is
  <= 2^-22 near a = 1, where the result's ulp shrinks;
* ln and log10 (log2 times a constant): one more rounding;
* pow(a, b) = exp2(b log2 a): the product b log2 a carries the log's
  error, which exp2 scales by ln 2 |b log2 a|, so the bound is
  2 + 2^-21 ln 2 |b log2 a| / 2^-24 ulp;
* e^x for x <= 0 as exp2(x log2 e): the rounding of x log2 e scales with
  |x|, bound 2 + 2 ln 2 |x log2 e| ulp (the routine is used for
  arguments above -40);
* the device library's expf, logf and powf: <= 3 ulp (the reference's
  precise_* functions aim at correct rounding; measured here: 2.3 ulp for
  logf at binade edges).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ulps(got, want):
    want32 = want.astype(np.float32)
    return np.abs(got.astype(np.float64) - want) / np.spacing(np.abs(want32)).astype(np.float64)


@pytest.fixture(scope="module")
def gpu(exo):
    e = exo.Engine(0, 1)
    yield e
    e.close()


def test_logarithms(gpu):
    rng = np.random.default_rng(1)
    a = np.float32(10.0) ** rng.uniform(-30, 30, 200000).astype(np.float32)
    a = np.concatenate([a, np.float32(1) + rng.uniform(-0.4, 0.4, 20000).astype(np.float32)])
    a64 = a.astype(np.float64)
    cases = [(0, np.log2(a64), 2), (3, np.log(a64), 3), (2, np.log10(a64), 3), (6, np.log(a64), 3)]
    for fn, want, bound in cases:
        got = gpu.math_eval(fn, a)
        far = np.abs(np.log2(a64)) >= 1
        u = _ulps(got[far], want[far])
        near = np.abs(got[~far] - want[~far]).max()
        print(f"fn {fn}: max {u.max():.2f} ulp for |log2 a| >= 1, max abs error {near:.2e} near 1")
        assert u.max() <= bound, fn
        assert near <= 2.0 ** -21, fn


@pytest.mark.parametrize("b", [0.243, 0.26, 0.421, 0.436, 0.5, 1.5])
def test_powers_of_absorber_paths(gpu, b):
    """The longwave fits' powers (exponents 0.243, 0.26, 0.421, 0.436) of
    column amounts plus offsets, and other powers on the same path."""
    rng = np.random.default_rng(2)
    a = np.float32(10.0) ** rng.uniform(-5, 4, 200000).astype(np.float32)
    want = np.power(a.astype(np.float64), np.float64(np.float32(b)))
    got = gpu.math_eval(1, a, np.float32(b))
    mag = np.abs(np.float64(np.float32(b)) * np.log2(a.astype(np.float64)))
    u = _ulps(got, want)
    bound = 2 + 2.0 ** -21 * np.log(2) * mag / 2.0 ** -24
    print(f"pow b={b}: max {u.max():.2f} ulp, max over bound {np.max(u / bound):.2f}")
    assert np.all(u <= bound)
    lib = gpu.math_eval(7, a, np.float32(b))
    ul = _ulps(lib, want)
    print(f"powf b={b}: max {ul.max():.2f} ulp")
    assert ul.max() <= 3


def test_exponentials(gpu):
    rng = np.random.default_rng(3)
    x = -rng.uniform(0, 40, 200000).astype(np.float32)
    want = np.exp(x.astype(np.float64))
    got = gpu.math_eval(4, x)
    u = _ulps(got, want)
    bound = 2 + 2 * np.log(2) * np.abs(x.astype(np.float64) * np.log2(np.e))
    print(f"exp_neg: max {u.max():.2f} ulp, max over bound {np.max(u / bound):.2f}")
    assert np.all(u <= bound)
    xs = rng.uniform(-80, 80, 200000).astype(np.float32)
    lib = gpu.math_eval(5, xs)
    ul = _ulps(lib, np.exp(xs.astype(np.float64)))
    print(f"expf: max {ul.max():.2f} ulp")
    assert ul.max() <= 3
