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


def _specials():
    f32 = np.float32
    tiny = np.finfo(f32).tiny
    den = np.nextafter(f32(0), f32(1))
    big = np.finfo(f32).max
    v = [0.0, -0.0, den, -den, 2 * den, tiny, -tiny, np.nextafter(tiny, f32(0)), 1.0, -1.0, big, -big,
         np.nextafter(big, f32(0)), 1e-30, -1e-30, 0.5, 3.0, -3.0]
    v = np.array(v, f32)
    with np.errstate(over="ignore"):
        return np.concatenate([v, np.nextafter(v, f32(np.inf)), np.nextafter(v, f32(-np.inf))]).astype(f32)


def test_sign_bit_bin_count_equals_comparisons():
    # k_eval_bins counts #{t : th[t] < q} as the sum of the sign bits of th[t] - (q + 0) over
    # thresholds stored with NaN -> +inf and -0 -> +0 (k_bin_tables), for finite q (NaN and
    # infinite q take other paths).  Host float32 arithmetic keeps denormals, as the kernels
    # are compiled to.
    rng = np.random.default_rng(7)
    sp = _specials()
    thr_raw = np.concatenate([sp, np.array([np.nan, -np.nan, np.inf, -np.inf], np.float32),
                              rng.normal(0, 50, 400).astype(np.float32),
                              (rng.normal(0, 1, 100) * 1e-40).astype(np.float32)])
    with np.errstate(invalid="ignore", over="ignore"):
        thr = np.where(np.isnan(thr_raw), np.float32(np.inf), thr_raw + np.float32(0)).astype(np.float32)
        assert not np.any(np.signbit(thr) & (thr == 0))
        q = np.concatenate([sp, rng.normal(0, 50, 3000).astype(np.float32),
                            thr[np.isfinite(thr)], (rng.normal(0, 1, 300) * 1e-40).astype(np.float32)])
        q = q[np.isfinite(q)].astype(np.float32)
        qc = (q + np.float32(0)).astype(np.float32)
        diff = (thr[None, :] - qc[:, None]).astype(np.float32)
        sign = (diff.view(np.uint32) >> 31).astype(np.int64)
        # the canonical thresholds compare as the stored ones do
        assert np.array_equal(thr[None, :] < q[:, None], thr_raw[None, :] < q[:, None])
        assert np.array_equal(sign.astype(bool), thr[None, :] < q[:, None])
        assert np.array_equal(sign.sum(1), (thr[None, :] < q[:, None]).sum(1))


def test_branch_free_next_float_up():
    # roundUpToFloat's next float above f: +-0 -> smallest denormal, negative -> toward zero,
    # positive -> away from zero, i.e. nextafter(f, +inf) for every f below +inf
    rng = np.random.default_rng(3)
    f = np.concatenate([_specials(), np.float32([-np.inf]), rng.normal(0, 1e3, 5000).astype(np.float32)])
    f = f[f < np.inf]
    u = f.view(np.uint32).astype(np.uint64)
    nxt = np.where(f == 0, 1, np.where(u & 0x80000000, u - 1, u + 1)).astype(np.uint32).view(np.float32)
    with np.errstate(over="ignore"):   # the largest float steps to +inf
        ref = np.nextafter(f, np.float32(np.inf))
    assert np.array_equal(nxt.view(np.uint32), ref.view(np.uint32))


def test_division_by_1000_as_corrected_product_is_exact_for_every_int32(tmp_path):
    # averageRegionDepth's (double)sum / 1000.0 (random_tree_image.h:288-316): the kernels compute it as
    # the product with RN(1/1000) plus one fma-corrected step; every 32-bit sum gives the same double
    import subprocess
    import os
    src = os.path.join(os.path.dirname(__file__), "native", "check_div1000.c")
    exe = str(tmp_path / "check_div1000")
    subprocess.run(["gcc", "-O2", "-march=native", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.strip()
    assert out == "0"


def test_forest_block_packing_walks_reach_the_plain_walks_leaves(tmp_path):
    # the prediction layout (curfil_amd/csrc/forest_pack.cpp): every walk stepped through the 128-byte
    # blocks the way k_classify steps stays in bounds, visits the plain walk's nodes and reaches its leaf
    import subprocess
    import os
    here = os.path.dirname(__file__)
    exe = str(tmp_path / "check_forest_pack")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, os.path.join(here, "native", "check_forest_pack.cpp"),
                    os.path.join(here, "..", "curfil_amd", "csrc", "forest_pack.cpp")], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.strip()
    assert out.startswith("ok ") and int(out.split()[1]) > 30000, out
