"""Exhaustive checks of the exact arithmetic shortcuts the kernels rely on
(host IEEE arithmetic; the device executes the same correctly rounded ops)."""
import numpy as np


def test_reciprocal_normalisation_is_exact_for_all_operands():
    # XY::normalize (random_tree_image.h:82-87) computes (int)((float)v / d) with
    # d = mm / 1000.0f.  The kernels compute (int)(float)(v * (1.0 / (double)d)).
    mm = np.arange(1, 50001, dtype=np.int64)
    d = (mm.astype(np.float32) / np.float32(1000.0)).astype(np.float32)
    rinv = 1.0 / d.astype(np.float64)
    for v in range(-127, 128):
        exact = np.trunc(np.float32(v) / d).astype(np.int64)
        fast = np.trunc((np.float64(v) * rinv).astype(np.float32)).astype(np.int64)
        assert np.array_equal(exact, fast), v


def _round_up_to_float(r):
    f = np.float32(r)
    if float(f) < r:
        f = np.nextafter(f, np.float32(np.inf))
    return f


def test_round_up_key_preserves_double_vs_float_comparison():
    # r <= (double)t  <=>  roundUpToFloat(r) <= t  for every float threshold t
    rng = np.random.default_rng(0)
    thr = np.concatenate([rng.normal(0, 100, 2000).astype(np.float32),
                          np.array([0.0, -0.0, 1e-30, -1e-30, 3.4e38, -3.4e38], np.float32)])
    for t in thr[:200]:
        base = float(t)
        cands = [base, np.nextafter(base, np.inf), np.nextafter(base, -np.inf), base + 1e-9, base - 1e-9,
                 float(np.nextafter(t, np.float32(np.inf))), float(np.nextafter(t, np.float32(-np.inf)))]
        for r in cands:
            q = _round_up_to_float(r)
            assert (r <= float(t)) == (q <= t)
    r = rng.normal(0, 100, 20000)
    q = np.array([_round_up_to_float(x) for x in r], np.float32)
    for t in thr:
        assert np.array_equal(r <= float(t), q <= t)
