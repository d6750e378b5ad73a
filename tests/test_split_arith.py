"""CPU model of the two split-precision digit schemes of the GPU kernels,
checked against exact arithmetic.

The kernels (tensortoolkit_amd/csrc/gemm_split.hip) emulate an FP64 product
from digit products that are exact on the MFMA units.  What makes that exact
enough is arithmetic, not code, so it is pinned here with Python integers and
fractions:

* int8 form: Y = rint(y 2^46) for |y| < 1 is written in balanced base 255 with
  six digits in [-127, 127] by the kernels' FP64 steps (q = rint(Y / 255),
  d = Y - 255 q); the digits must reconstruct Y exactly, stay in range, and
  the 21 kept digit pairs (sa + sb <= 5) must give an inner product within
  the error budget; an int32 level stays below 2^31 for 8192 complex k.
* BF16 form: five slices c_s with |c_s| <= 256 (so every slice is exact in
  BF16), y = 2^-8 sum_s c_s 2^(-9(s-1)) + O(2^-44), and the 15 kept slice
  pairs give an inner product within the budget.
"""
from fractions import Fraction

import numpy as np
import pytest

TWO46 = 2.0 ** 46


def i8_digits(y: np.ndarray) -> np.ndarray:
    """Six base-255 digits (most significant first) of rint(y 2^46), by the
    same FP64 operations the pack kernels use."""
    Y = np.rint(y * TWO46)
    d = np.zeros((6,) + y.shape, dtype=np.int64)
    for s in range(5, -1, -1):
        q = np.rint(Y * (1.0 / 255.0))
        r = Y - 255.0 * q          # exact: both are integers below 2^53
        d[s] = r.astype(np.int64)
        Y = q
    assert np.all(Y == 0)
    return d


def i8_value(d: np.ndarray) -> np.ndarray:
    return sum(d[s].astype(object) * 255 ** (5 - s) for s in range(6))


def bf16_slices(y: np.ndarray) -> np.ndarray:
    """Five integer slices c_s of y (|y| < 1): c_1 = rint(y 2^8), then
    c_s = rint(r 2^(8 + 9(s-1))) of the remainder, as in the pack kernels."""
    c = np.zeros((5,) + y.shape, dtype=np.int64)
    r = y.copy()
    for s in range(5):
        scale = 2.0 ** (8 + 9 * s)
        cs = np.rint(r * scale)
        r = r - cs / scale          # exact (power-of-two scaling)
        c[s] = cs.astype(np.int64)
    return c


def _unit_rows(rng, n):
    """Values scaled like a packed row: max |y| in [1/2, 1)."""
    y = rng.uniform(-1.0, 1.0, n)
    y /= np.max(np.abs(y))
    y *= 0.999
    return y


@pytest.mark.parametrize("seed", range(4))
def test_i8_digits_reconstruct_exactly_and_stay_in_range(seed):
    rng = np.random.default_rng(seed)
    y = np.concatenate([rng.uniform(-1, 1, 20000), rng.uniform(-1, 1, 2000) * 2.0 ** -30,
                        np.array([0.0, 1 - 2 ** -53, -(1 - 2 ** -53), 0.5, -0.5, 2 ** -46, -(2 ** -47)])])
    d = i8_digits(y)
    assert d.min() >= -127 and d.max() <= 127
    exact = [int(v) for v in np.rint(y * TWO46)]
    assert list(i8_value(d)) == exact
    # the leading digit of |Y| <= 2^46 is small: 2^46 / 255^5 < 66
    assert np.abs(d[0]).max() <= 65


def test_i8_digit_set_is_closed_under_negation():
    """The complex real part uses -Ai: the digits of -Y are the negated digits."""
    rng = np.random.default_rng(7)
    y = rng.uniform(-1, 1, 5000)
    assert np.array_equal(i8_digits(-y), -i8_digits(y))


def test_i8_level_sums_fit_int32_for_8192_k():
    """Level L holds at most 6 digit pairs; a complex k adds two products per
    pair to each output component, each at most 127^2 in magnitude."""
    per_k = 2 * 6 * 127 * 127
    assert per_k * 8192 < 2 ** 31
    assert per_k * 11096 >= 2 ** 31 > per_k * 11095   # the exact limit


def test_bf16_slices_are_exact_in_bf16_and_reconstruct():
    rng = np.random.default_rng(3)
    y = rng.uniform(-1, 1, 20000) * 0.999
    c = bf16_slices(y)
    assert np.abs(c).max() <= 256          # <= 8 significant bits: exact in BF16
    for i in range(0, 20000, 997):
        v = sum(Fraction(int(c[s][i]), 2 ** (8 + 9 * s)) for s in range(5))
        assert abs(v - Fraction(y[i])) <= Fraction(1, 2 ** 44)


def _exact_dot(a, b):
    return sum(Fraction(x) * Fraction(z) for x, z in zip(a, b))


def _sqrt_k_budget(per_term: float, k: int) -> Fraction:
    """Error budget for random data, whose per-term errors add like a random
    walk (sqrt(k)); the worst case grows like k."""
    return Fraction(per_term) * Fraction(np.sqrt(k))


@pytest.mark.parametrize("k", [1, 7, 64, 300])
def test_i8_truncated_products_stay_inside_the_budget(k):
    """sum over kept pairs sa + sb <= 5 of (sum_k d_sa d_sb) 255^(10-sa-sb) /
    2^92 against the exact inner product of the unrounded values.  Per term
    (max|a| = max|b| < 1) the error is the quantisation of each value, 2^-47,
    plus the dropped pairs, about 2^-46: below 1e-13 in the worst case, and
    about 4e-14 sqrt(k) in total for random data."""
    rng = np.random.default_rng(100 + k)
    a, b = _unit_rows(rng, k), _unit_rows(rng, k)
    da, db = i8_digits(a), i8_digits(b)
    acc = 0
    for sa in range(6):
        for sb in range(6 - sa):
            lvl = int(np.dot(da[sa], db[sb]))
            acc += lvl * 255 ** (10 - sa - sb)
    got = Fraction(acc, 2 ** 92)
    exact = _exact_dot(a, b)
    assert abs(got - exact) <= Fraction(1, 10 ** 13) * k
    assert abs(got - exact) <= _sqrt_k_budget(2e-13, k)
    # a plausible mistake - keeping only 15 pairs (sa + sb <= 4) - fails the
    # budget (largest error over 16 draws: one draw can cancel by chance)
    worst = Fraction(0)
    for _ in range(16):
        a, b = _unit_rows(rng, k), _unit_rows(rng, k)
        da, db = i8_digits(a), i8_digits(b)
        acc15 = sum(int(np.dot(da[sa], db[sb])) * 255 ** (10 - sa - sb) for sa in range(6) for sb in range(5 - sa))
        worst = max(worst, abs(Fraction(acc15, 2 ** 92) - _exact_dot(a, b)))
    assert worst > _sqrt_k_budget(2e-13, k)


@pytest.mark.parametrize("k", [1, 64, 300])
def test_bf16_truncated_products_stay_inside_the_budget(k):
    """15 kept slice pairs (sa + sb <= 4, 0-based) of the BF16 form: per term
    the quantisation 2^-45 of each value plus the dropped pairs, about 2^-45."""
    rng = np.random.default_rng(200 + k)
    a, b = _unit_rows(rng, k), _unit_rows(rng, k)
    ca, cb = bf16_slices(a), bf16_slices(b)
    got = Fraction(0)
    for sa in range(5):
        for sb in range(5 - sa):
            got += Fraction(int(np.dot(ca[sa], cb[sb])), 2 ** (16 + 9 * (sa + sb)))
    exact = _exact_dot(a, b)
    assert abs(got - exact) <= Fraction(1, 10 ** 12) * k
    assert abs(got - exact) <= _sqrt_k_budget(5e-13, k)
    # dropping the last kept level (10 pairs) fails the budget (largest error
    # over 16 draws)
    worst = Fraction(0)
    for _ in range(16):
        a, b = _unit_rows(rng, k), _unit_rows(rng, k)
        ca, cb = bf16_slices(a), bf16_slices(b)
        got10 = sum(Fraction(int(np.dot(ca[sa], cb[sb])), 2 ** (16 + 9 * (sa + sb)))
                    for sa in range(5) for sb in range(4 - sa))
        worst = max(worst, abs(got10 - _exact_dot(a, b)))
    assert worst > _sqrt_k_budget(5e-13, k)
