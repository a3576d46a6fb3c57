"""The gradient accumulators' FP32 -> 64-bit fixed-point conversion (csrc/mmff_hip.hip to_fx),
restated with numpy's IEEE float64 and checked against its definition: round(v * 2^28) with
ties to even, exact negation symmetry (so +g and -g of a pair term cancel exactly), and an
exact, order-independent sum. The device code performs the same IEEE operations (a clamp,
FP32 -> FP64 conversion, one FP64 add in round-to-nearest-even, a 64-bit integer subtract)."""
from __future__ import annotations

import numpy as np

FX_MAGIC = 25165824.0  # 1.5 * 2^24
FX_CAP = np.float32(8388000.0)
SCALE = 2.0 ** 28


def to_fx(v: np.ndarray) -> np.ndarray:
    v = np.asarray(v, np.float32)
    # fmaxf / fminf return the non-NaN operand: NaN -> -FX_CAP
    c = np.where(v == v, np.clip(v, -FX_CAP, FX_CAP), -FX_CAP).astype(np.float32)
    y = c.astype(np.float64) + FX_MAGIC
    return y.view(np.int64) - np.array([FX_MAGIC]).view(np.int64)[0]


def _samples(rng, n):
    mags = 10.0 ** rng.uniform(-12, 6.9, n)
    v = (mags * rng.choice([-1.0, 1.0], n)).astype(np.float32)
    # exact ties: odd multiples of 2^-29
    ties = (rng.integers(-2 ** 20, 2 ** 20, n // 10) * 2 + 1).astype(np.float64) * 2.0 ** -29
    return np.concatenate([v, ties.astype(np.float32), np.float32([0.0, -0.0, 1.0, -1.0, 8388000.0, -8388000.0])])


def test_conversion_is_round_half_even_of_the_scaled_value():
    rng = np.random.default_rng(7)
    v = _samples(rng, 200000)
    want = np.rint(v.astype(np.float64) * SCALE).astype(np.int64)  # exact product, ties to even
    assert np.array_equal(to_fx(v), want)


def test_negation_is_exact_and_cap_and_nan():
    rng = np.random.default_rng(8)
    v = _samples(rng, 100000)
    assert np.array_equal(to_fx(-v), -to_fx(v))
    big = np.float32([1e9, -1e9, np.inf, -np.inf])
    assert np.array_equal(to_fx(big), np.rint(np.float64([8388000.0, -8388000.0] * 2) * SCALE).astype(np.int64))
    assert to_fx(np.float32([np.nan]))[0] == -int(np.rint(8388000.0 * SCALE))


def test_sum_is_exact_and_order_independent():
    rng = np.random.default_rng(9)
    v = (rng.normal(0, 30.0, 5000)).astype(np.float32)
    fx = to_fx(v)
    s1 = int(np.sum(fx))
    s2 = int(np.sum(fx[rng.permutation(len(fx))]))
    assert s1 == s2
    # the fixed-point sum is within 2^-29 per contribution of the exact sum of the FP32 values
    exact = float(np.sum(v.astype(np.float64)))
    assert abs(s1 / SCALE - exact) <= len(v) * 2.0 ** -29 + 1e-9
