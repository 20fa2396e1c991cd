"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bar: bit-exact for integrals, responses (double), counters, feature tables,
chosen splits, tree structure and labels; scores compared against the oracle
(FP64, glibc log2) with the maximum difference stated; posteriors within
1e-5 relative.
"""
import math
import os
import re

import numpy as np
import pytest

import curfil_amd as cf
import oracle as orc
from helpers import (CONFIG0, GOLDEN_FEATURE_HASH, boost_feature_hash, config0_dataset, golden_hash_dataset,
                     random_features, random_samples, random_u8_dataset)

pytestmark = pytest.mark.gpu


def _images_u8(ctx, rgb, depth, labels=None):
    n, H, W, _ = rgb.shape
    im = cf.Images(ctx, n, W, H)
    for i in range(n):
        im.set_rgbd_u8(i, rgb[i], depth[i])
        if labels is not None:
            im.set_labels(i, labels[i])
    return im


def _images_f32(ctx, color, depth):
    n, _, H, W = color.shape
    im = cf.Images(ctx, n, W, H)
    for i in range(n):
        im.set_rgbd_f32(i, color[i], depth[i])
    return im


def _same_float_bits(a, b):
    return np.array_equal(np.asarray(a).view(np.uint32 if a.dtype == np.float32 else np.uint64),
                          np.asarray(b).view(np.uint32 if b.dtype == np.float32 else np.uint64))


# ------------------------------------------------------------------ integrals
@pytest.mark.parametrize("W,H", [(160, 120), (640, 480), (37, 5)])
def test_integrals_u8_bit_exact(ctx, W, H):
    rgb, depth, _ = random_u8_dataset(2, W, H, seed=W)
    im = _images_u8(ctx, rgb, depth)
    ds = orc.Dataset.from_u8(rgb, depth)
    for i in range(2):
        col, dep, val = im.integrals(i)
        ocol, odep, oval = ds.integrals(i)
        assert _same_float_bits(col, ocol)
        assert np.array_equal(dep, odep) and np.array_equal(val, oval)


@pytest.mark.parametrize("W,H", [(64, 48), (640, 480), (3000, 3)])
def test_integrals_f32_kahan_bit_exact(ctx, W, H):
    """the sequential Kahan recurrence (LDS row kernel, and the global-memory kernel for rows wider
    than its LDS), with signed non-integer values as CIELab planes hold"""
    rng = np.random.default_rng(W)
    color = (rng.integers(-10000, 10000, (1, 3, H, W)).astype(np.float32) / np.float32(100.0))
    depth = rng.integers(0, 5000, (1, H, W)).astype(np.int32)
    im = _images_f32(ctx, color, depth)
    col, dep, val = im.integrals(0)
    ocol, odep, oval = orc.Dataset(color, depth).integrals(0)
    assert _same_float_bits(col, ocol)
    assert np.array_equal(dep, odep) and np.array_equal(val, oval)


def test_depth_integral_wraps_but_box_sums_stay_exact(ctx):
    # 1280x960 at 5000 mm overflows 32-bit integral entries (the reference's int
    # integral overflows here); box sums from the wrapped entries are exact.
    W, H = 1280, 960
    rgb = np.zeros((1, H, W, 3), np.uint8)
    depth = np.full((1, H, W), 5000, np.uint16)
    im = _images_u8(ctx, rgb, depth)
    _, dep, val = im.integrals(0)
    s = (np.int64(dep[H - 1, W - 1]) - dep[H - 11, W - 1] - dep[H - 1, W - 11] + dep[H - 11, W - 11]) % (1 << 32)
    assert s == 100 * 5000
    f = np.array([0, 0, 0, 5, 5, 3, 3, 3, 3, 0, 0], np.int8)[:, None]
    samples = cf.make_samples([W - 20], [H - 20], [0], [5000], [0])
    r = cf.feature_responses(ctx, im, samples, f)
    assert r[0, 0] == 0.0


# ------------------------------------------------------------------ responses / counters known answers
def test_known_answer_complex_case(ctx):
    from test_oracle_pins import COMPLEX_COUNTERS, COMPLEX_RESPONSES, complex_case
    color, depth, feats, thr, samples = complex_case()
    im = _images_f32(ctx, color, depth)
    r = cf.feature_responses(ctx, im, samples, feats)
    assert np.array_equal(np.isnan(r), np.isnan(COMPLEX_RESPONSES))
    assert np.array_equal(r[~np.isnan(r)], COMPLEX_RESPONSES[~np.isnan(COMPLEX_RESPONSES)])
    c = cf.node_counters(ctx, im, samples, feats, thr, 2)   # [F][T][C][2]
    for (f, t), per_label in COMPLEX_COUNTERS.items():
        for label, lr in enumerate(per_label):
            assert tuple(c[f, t, label]) == lr


def test_invalid_features_rejected_on_host(ctx):
    """A colour channel outside 0..2 or an unknown type is refused before any
    kernel runs (it would address outside the integral images)."""
    from test_oracle_pins import complex_case
    color, depth, feats, thr, samples = complex_case()
    im = _images_f32(ctx, color, depth)
    for p, v in ((9, 3), (10, -1), (0, 2)):
        bad = feats.copy()
        bad[0, 0] = 1
        bad[p, 0] = v
        with pytest.raises(cf.CurfilError, match="invalid"):
            cf.feature_responses(ctx, im, samples, bad)
        with pytest.raises(cf.CurfilError, match="invalid"):
            cf.node_counters(ctx, im, samples, bad, thr, 2)


def test_invalid_thresholds_depths_and_sizes_rejected_on_host(ctx):
    """A NaN threshold (the reference never draws one: sampled thresholds are finite responses), depth
    above 50000 mm (image.cpp:430) and images whose integral planes would pass 4 GiB are refused on the
    host with CF_ERR_ARG before any kernel runs."""
    from test_oracle_pins import complex_case
    color, depth, feats, thr, samples = complex_case()
    im = _images_f32(ctx, color, depth)
    bad = thr.copy()
    bad[1, 2] = np.nan
    with pytest.raises(cf.CurfilError, match="NaN threshold"):
        cf.node_counters(ctx, im, samples, feats, bad, 2)
    with pytest.raises(cf.CurfilError, match="NaN threshold"):
        cf.node_best_split(ctx, im, samples, feats, bad, 2, np.array([2, 2], np.uint32))
    W, H = 16, 12
    im8 = cf.Images(ctx, 1, W, H)
    dep = np.full((H, W), 1000, np.uint16)
    dep[3, 4] = 50001
    with pytest.raises(cf.CurfilError, match="50000"):
        im8.set_rgbd_u8(0, np.zeros((H, W, 3), np.uint8), dep)
    with pytest.raises(cf.CurfilError, match="50000"):
        im8.set_rgbd_u8_padded(0, np.zeros((H - 2, W - 2, 3), np.uint8), dep[:H - 2, :W - 2].copy())
    with pytest.raises(cf.CurfilError, match="4 GiB"):
        cf.Images(ctx, 1, 30000, 20000)


def test_known_answer_depth_simple(ctx):
    W, H = 16, 20
    depth = np.zeros((1, H, W), np.int32)
    depth[0, 3, 7], depth[0, 4, 7], depth[0, 4, 8], depth[0, 2, 3] = 1500, 1700, 4500, 3900
    im = _images_f32(ctx, np.zeros((1, 3, H, W), np.float32), depth)
    f = np.array([0, 1, 1, -3, -1, 2, 2, 1, 1, 0, 0], np.int8)[:, None]
    thr = ((np.arange(100) - 50) / np.float32(10.0)).astype(np.float32)[:, None]
    s = cf.make_samples([6, 6], [3, 3], [0, 0], [1000, 1000], [0, 1])
    r = cf.feature_responses(ctx, im, s, f)
    assert abs(r[0, 0] - ((1.5 + 1.7 + 4.5) / 3 - 3.9)) < 1e-12
    c = cf.node_counters(ctx, im, s, f, thr, 2)
    for label in range(2):
        assert np.array_equal(c[0, :, label, 0], (np.arange(100) >= 37).astype(np.uint32))
        assert np.all(c[0, :, label, 0] + c[0, :, label, 1] == 1)


# ------------------------------------------------------------------ random parity
@pytest.fixture(scope="module")
def rand_data(ctx):
    rgb, depth, labels = random_u8_dataset(3, 160, 120, seed=11)
    im = _images_u8(ctx, rgb, depth, labels)
    ds = orc.Dataset.from_u8(rgb, depth, labels)
    return rgb, depth, labels, im, ds


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_responses_bit_exact(ctx, rand_data, seed):
    rgb, depth, labels, im, ds = rand_data
    rng = np.random.default_rng(seed)
    feats = random_features(rng, 64, box_radius=100 if seed else 40, region_size=20)
    s = random_samples(rng, 500, 3, 160, 120, depth.astype(np.int32))
    s["flip"] = rng.integers(0, 2, len(s))
    r = cf.feature_responses(ctx, im, s, feats)
    o = ds.responses(s, feats)
    assert np.array_equal(np.isnan(r), np.isnan(o))
    assert _same_float_bits(np.nan_to_num(r), np.nan_to_num(o))


@pytest.mark.parametrize("seed,n,T", [(0, 1, 5), (1, 777, 10), (2, 9000, 32), (3, 300, 1)])
def test_random_counters_exact(ctx, rand_data, seed, n, T):
    rgb, depth, labels, im, ds = rand_data
    rng = np.random.default_rng(seed)
    F = 40
    feats = random_features(rng, F)
    s = random_samples(rng, n, 3, 160, 120, depth.astype(np.int32), classes=6)
    # thresholds: float-rounded responses of random samples, plus duplicates
    resp = ds.responses(random_samples(rng, T, 3, 160, 120, depth.astype(np.int32)), feats)
    thr = np.nan_to_num(resp.T, nan=0.0).astype(np.float32)
    if T > 2:
        thr[1] = thr[0]
    c = cf.node_counters(ctx, im, s, feats, thr, 6)
    o = ds.node_counters(s, feats, thr, 6)          # [C][F][T][2]
    assert np.array_equal(c, o.transpose(1, 2, 0, 3))


def test_scores_against_oracle(ctx, rand_data):
    rgb, depth, labels, im, ds = rand_data
    rng = np.random.default_rng(5)
    F, T, C = 60, 16, 6
    feats = random_features(rng, F)
    s = random_samples(rng, 2000, 3, 160, 120, depth.astype(np.int32), labels=labels)
    resp = ds.responses(random_samples(rng, T, 3, 160, 120, depth.astype(np.int32)), feats)
    thr = np.nan_to_num(resp.T, nan=0.0).astype(np.float32)
    o_cnt = ds.node_counters(s, feats, thr, C)
    hist = np.bincount(s["label"], minlength=C).astype(np.uint32)
    o = orc.scores_from_counters(o_cnt, hist)
    g = cf.scores(ctx, o_cnt.transpose(1, 2, 0, 3), hist)
    err = np.abs(g - o).max()
    print(f"max |score_gpu - score_oracle| = {err:.3e}; bit-equal fraction {np.mean(g == o):.4f}")
    assert err <= 4e-16


def test_node_best_split_matches_oracle(ctx, rand_data):
    rgb, depth, labels, im, ds = rand_data
    rng = np.random.default_rng(9)
    for trial in range(12):
        F, T, C = int(rng.integers(1, 50)), int(rng.integers(1, 21)), 6
        n = int(rng.integers(2, 3000))
        feats = random_features(rng, F)
        s = random_samples(rng, n, 3, 160, 120, depth.astype(np.int32), labels=labels)
        resp = ds.responses(random_samples(rng, T, 3, 160, 120, depth.astype(np.int32)), feats)
        thr = np.nan_to_num(resp.T, nan=0.0).astype(np.float32)
        hist = np.bincount(s["label"], minlength=C).astype(np.uint32)
        o = orc.scores_from_counters(ds.node_counters(s, feats, thr, C), hist)
        of, ot, osc = orc.argmax(o)
        gf, gt, gsc = cf.node_best_split(ctx, im, s, feats, thr, C, hist)
        assert (gf, gt) == (of, ot), trial
        assert abs(gsc - osc) <= 4e-16


# ------------------------------------------------------------------ feature generation
def _adversarial_score_matrices():
    rng = np.random.default_rng(11)
    yield "random", rng.random((20, 2000))
    yield "increasing", np.linspace(0.0, 1.0, 20 * 2000).reshape(20, 2000)   # every score is a new best
    yield "decreasing", np.linspace(1.0, 0.0, 10 * 300).reshape(10, 300)
    m = rng.random((16, 700)) * 0.5
    m[3, 5] = 0.9
    m[7, 600] = 0.9 + 5e-14     # within 1e-13 of the best: must not replace it
    m[9, 10] = 0.9 + 2e-13      # beyond the band: replaces
    m[12, 1] = 0.9 + 2e-13      # equal to the best: must not replace it
    yield "band", m
    steps = np.full((8, 129), 0.25)
    steps.flat[::7] += np.arange(steps.flat[::7].size) * 6e-14   # a staircase of sub-band increments
    yield "staircase", steps
    yield "constant", np.full((5, 64), 0.5)
    yield "single", np.array([[0.3]])


def test_argmax_scores_matches_sequential_scan(ctx):
    for name, m in _adversarial_score_matrices():
        got = cf.argmax_scores(ctx, m)
        want = orc.argmax(m)
        assert got[0] == want[0] and got[1] == want[1] and got[2] == want[2], (name, got, want)


def test_golden_feature_hash_on_device(ctx):
    color, depth, samples = golden_hash_dataset()
    im = _images_f32(ctx, color, depth)
    feats, thr = cf.generate_features(ctx, im, samples, seed=9, num_features=500, num_thresholds=50,
                                      box_radius=127, region_size=16)
    assert boost_feature_hash(feats, thr) == GOLDEN_FEATURE_HASH


# boxRadius 127 on 160x120 frames: most responses NaN, so candidates are rejected and NaN runs carry
# across 64-draw words and walk rounds; T = 64 spans several words per candidate
@pytest.mark.parametrize("seed,n,F,T,br,rs", [(17, 4000, 100, 10, 60, 10), (3, 1, 20, 4, 60, 10),
                                              (123456, 50, 300, 32, 60, 10), (99, 2500, 2000, 20, 60, 10),
                                              (5, 3000, 400, 40, 127, 20), (11, 800, 200, 64, 127, 4)])
def test_feature_generation_matches_oracle(ctx, seed, n, F, T, br, rs):
    rgb, depth, labels = config0_dataset()
    im = _images_u8(ctx, rgb, depth, labels)
    ds = orc.Dataset.from_u8(rgb, depth, labels)
    rng = np.random.default_rng(seed)
    s = random_samples(rng, n, 4, 160, 120, depth.astype(np.int32), labels=labels)
    s = s[s["depth_mm"] > 0] if n > 1 else s
    if len(s) == 0 or s["depth_mm"].min() <= 0:
        s = s[:0]
        s = random_samples(rng, 1, 4, 160, 120, np.full_like(depth, 1000, dtype=np.int32))
    gf, gt = cf.generate_features(ctx, im, s, seed, F, T, br, rs)
    of, ot = ds.generate_features(s, seed, F, T, br, rs)
    assert np.array_equal(gf, of)
    assert _same_float_bits(gt, ot)


# ------------------------------------------------------------------ training
@pytest.fixture(scope="module")
def config0(ctx):
    rgb, depth, labels = config0_dataset()
    im = _images_u8(ctx, rgb, depth, labels)
    ds = orc.Dataset.from_u8(rgb, depth, labels)
    return rgb, depth, labels, im, ds


def _assert_trees_equal(gn, gh, on, oh):
    assert len(gn) == len(on)
    for k in ("id", "level", "parent", "left", "right", "num_samples", "is_leaf"):
        assert np.array_equal(gn[k], on[k]), k
    inner = gn["is_leaf"] == 0
    assert np.array_equal(gn["feature_id"][inner], on["feature_id"][inner])
    assert np.array_equal(gn["feature"][inner], on["feature"][inner])
    assert _same_float_bits(gn["threshold"][inner], on["threshold"][inner])
    assert np.abs(gn["score"][inner] - on["score"][inner]).max(initial=0) <= 4e-16
    assert np.array_equal(gh, oh)


@pytest.mark.parametrize("mode", [cf.MODE_PLAIN, cf.MODE_BATCHED])
def test_forest_training_matches_oracle(ctx, config0, mode):
    rgb, depth, labels, im, ds = config0
    cfg = cf.TrainingConfiguration(**CONFIG0)
    forest = cf.RandomForestImage.train(ctx, im, cfg, tree_count=1, num_labels=6, mode=mode)
    ocfg = orc.config(**{k: v for k, v in CONFIG0.items()})
    of = ds.train_forest(ocfg, 1, 6)
    assert np.array_equal(forest.root_samples(0), of.roots(0))
    assert np.array_equal(forest.prior(), of.prior())
    gn, gh = forest.tree(0)
    on, oh = of.tree(0)
    print(f"mode {mode}: {len(gn)} nodes, depth {gn['level'].max()}, stats {forest.train_stats()}")
    _assert_trees_equal(gn, gh, on, oh)


def test_three_tree_forest_and_prediction_match_oracle(ctx, config0):
    rgb, depth, labels, im, ds = config0
    params = dict(CONFIG0)
    params.update(samplesPerImage=3000, maxDepth=10, featureCount=200, thresholds=20)
    cfg = cf.TrainingConfiguration(**params)
    forest = cf.RandomForestImage.train(ctx, im, cfg, tree_count=3, num_labels=6, mode=cf.MODE_BATCHED)
    of = ds.train_forest(orc.config(**params), 3, 6)
    trees, posts = [], []
    forest.normalizeHistograms(0.0, True)
    prior = of.prior()
    for t in range(3):
        gn, gh = forest.tree(t)
        on, oh = of.tree(t)
        _assert_trees_equal(gn, gh, on, oh)
        post = np.array([orc.normalize_histogram(h, prior, 0.0, True) for h in oh])
        assert np.array_equal(forest.posteriors(t), post)
        trees.append(on)
        posts.append(post)
    for img in range(4):
        gl, gp = forest.predict(im, img)
        ol, op = ds.predict(img, trees, posts, forest.num_classes)
        assert np.array_equal(gl, ol)
        assert np.array_equal(np.isnan(gp), np.isnan(op))
        assert _same_float_bits(np.nan_to_num(gp), np.nan_to_num(op))
    bl, bp = forest.predict_batch(im, [3, 1, 2])
    for k, img in enumerate([3, 1, 2]):
        sl, sp = forest.predict(im, img)
        assert np.array_equal(bl[k], sl) and _same_float_bits(np.nan_to_num(bp[k]), np.nan_to_num(sp))
    acc = np.mean(forest.predict(im, 0)[0][labels[0] > 0] == labels[0][labels[0] > 0])
    print(f"training-image pixel accuracy (no void): {acc:.3f}")


def test_random_complete_forest_prediction_matches_oracle(ctx, config0):
    """Complete trees with random splits (the dense-prediction benchmark's forest, smaller): normalised
    posteriors, per-pixel float posteriors and labels equal the oracle's CPU prediction bit for bit."""
    rgb, depth, labels, im, ds = config0
    forest = cf.synth_forest(ctx, im, depth.astype(np.int32), trees=4, depth=10, classes=5, seed=99)
    forest.normalizeHistograms(0.0, False)
    ones = np.ones(5, np.uint32)
    trees, posts = [], []
    for t in range(4):
        nodes, hist = forest.tree(t)
        assert len(nodes) == 1023 and nodes["is_leaf"].sum() == 512
        post = np.array([orc.normalize_histogram(h, ones, 0.0, False) for h in hist])
        assert np.array_equal(forest.posteriors(t), post)
        trees.append(nodes)
        posts.append(post)
    for img in (0, 3):
        gl, gp = forest.predict(im, img)
        ol, op = ds.predict(img, trees, posts, 5)
        assert np.array_equal(gl, ol)
        assert _same_float_bits(gp, op)
    bl, bp = forest.predict_batch(im, [3, 0])
    sl, sp = forest.predict(im, 3)
    assert np.array_equal(bl[0], sl) and _same_float_bits(bp[0], sp)


def test_server_results_equal_single_frame_prediction(ctx, config0):
    """Frames of three sizes submitted interleaved to the serving loop (continuous batching, and a budget
    that makes frames wait for slots) give the same posteriors and labels, bit for bit, as predicting each
    frame alone."""
    rgb0, depth0, labels0, im0, ds = config0
    forest = cf.synth_forest(ctx, im0, depth0.astype(np.int32), trees=3, depth=9, classes=5, seed=7)
    forest.normalizeHistograms(0.0, False)
    sizes = [(96, 64), (160, 120), (200, 150)]
    frames, refs = [], []
    for k, (W, H) in enumerate(sizes):
        rgb, depth, _ = cf.synth_dataset(4, W, H, sites=8, classes=5, zero_frac=0.05, seed0=100 * k + 1)
        im = _images_u8(ctx, rgb, depth)
        for i in range(4):
            frames.append((rgb[i], depth[i]))
            refs.append(forest.predict(im, i))
    order = [0, 4, 8, 1, 5, 9, 2, 6, 10, 3, 7, 11, 8, 0]
    big = (200 * 150) * (5 + 20 + 12 + 8 + 5 * 4 + 1) / (1 << 20)
    for budget in (256.0, 2.5 * big):
        server = cf.Server(ctx, forest, budget_mb=budget, max_batch=4)
        ids = [server.submit(*frames[j]) for j in order]
        for rid, j in reversed(list(zip(ids, order))):
            lab, probs, lat = server.wait(rid)
            rl, rp = refs[j]
            assert np.array_equal(lab, rl) and _same_float_bits(probs, rp), (budget, j)
            assert lat > 0
        st = server.stats()
        assert st["frames"] == len(order)
        server.close()
    with pytest.raises(cf.CurfilError):
        cf.Server(ctx, forest, budget_mb=0.01).submit(*frames[0])


def test_predict_batch_device_outputs_equal_host_outputs(ctx, config0):
    rgb, depth, labels, im, ds = config0
    forest = cf.synth_forest(ctx, im, depth.astype(np.int32), trees=3, depth=8, classes=5, seed=5)
    forest.normalizeHistograms(0.0, False)
    idx = [2, 0, 3]
    hl, hp = forest.predict_batch(im, idx)
    dp = cf.DeviceBuffer(ctx, 3 * 5 * 120 * 160 * 4)
    dl = cf.DeviceBuffer(ctx, 3 * 120 * 160)
    forest.predict_batch_device(im, idx, dp.ptr, dl.ptr)
    assert np.array_equal(dl.to_host(np.uint8, (3, 120, 160)), hl)
    assert _same_float_bits(dp.to_host(np.float32, (3, 5, 120, 160)), hp)
    dp.close()
    dl.close()



def test_exported_and_imported_forest_predicts_the_same(ctx, config0, tmp_path):
    """import_export_test.cpp on the device path: a trained forest written as tree<N>.json.gz and read back
    (without a context, then bound by its first prediction) has the same trees, and predicts the same labels
    and posteriors, bit for bit."""
    rgb, depth, labels, im, ds = config0
    params = dict(CONFIG0)
    params.update(samplesPerImage=2000, maxDepth=9, featureCount=150, thresholds=15)
    cfg = cf.TrainingConfiguration(**params)
    forest = cf.RandomForestImage.train(ctx, im, cfg, tree_count=3, num_labels=6, mode=cf.MODE_BATCHED)
    forest.normalizeHistograms(0.0, True)
    paths = forest.export(tmp_path, cfg, folder_training="synthetic")
    back = cf.RandomForestImage.from_files(paths, histogramBias=0.0, useLabelsPrior=True)
    assert back.ctx is None
    for t in range(3):
        gn, gh = forest.tree(t)
        bn, bh = back.tree(t)
        _assert_trees_equal(bn, bh, gn, gh)
        assert np.array_equal(back.posteriors(t), forest.posteriors(t))
    for img in (0, 2):
        fl, fp = forest.predict(im, img)
        bl, bp = back.predict(im, img)
        assert back.ctx is ctx
        assert np.array_equal(fl, bl) and _same_float_bits(np.nan_to_num(fp), np.nan_to_num(bp))


@pytest.mark.parametrize("variant", [dict(subsamplingType="pixelUniform"), dict(maxImages=3),
                                     dict(maxImages=2, subsamplingType="pixelUniform")])
def test_sampling_variants_train_like_oracle(ctx, config0, variant):
    """pixelUniform subsampling (missing-depth pixels included) and maxImages image subsampling: the same
    root samples, per-tree label priors and trees as the oracle, and each tree normalises with its prior."""
    rgb, depth, labels, im, ds = config0
    params = dict(CONFIG0, maxDepth=6, **variant)
    cfg = cf.TrainingConfiguration(**params)
    forest = cf.RandomForestImage.train(ctx, im, cfg, tree_count=2, num_labels=6, mode=cf.MODE_BATCHED)
    of = ds.train_forest(orc.config(**params), 2, 6)
    assert np.array_equal(forest.prior(), of.prior())
    forest.normalizeHistograms(0.0, True)
    for t in range(2):
        assert np.array_equal(forest.root_samples(t), of.roots(t))
        assert np.array_equal(forest.prior(t), of.prior(t))
        gn, gh = forest.tree(t)
        on, oh = of.tree(t)
        _assert_trees_equal(gn, gh, on, oh)
        post = np.array([orc.normalize_histogram(h, of.prior(t), 0.0, True) for h in oh])
        assert np.array_equal(forest.posteriors(t), post)


@pytest.mark.parametrize("mode", [cf.MODE_PLAIN, cf.MODE_BATCHED])
def test_horizontal_flipping_trains_like_oracle(ctx, config0, mode):
    """Horizontal flipping: Both samples evaluated in both orientations and split into NoFlip + Flip
    copies when a split separates them; trees equal the oracle's node for node."""
    rgb, depth, labels, im, ds = config0
    params = dict(CONFIG0, maxDepth=7, horizontalFlipping=True)
    forest = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=2, num_labels=6,
                                        mode=mode)
    of = ds.train_forest(orc.config(**params), 2, 6)
    for t in range(2):
        assert np.array_equal(forest.root_samples(t), of.roots(t))
        gn, gh = forest.tree(t)
        on, oh = of.tree(t)
        _assert_trees_equal(gn, gh, on, oh)
    grown = [gn["num_samples"][gn["left"][i]] + gn["num_samples"][gn["right"][i]] - gn["num_samples"][i]
             for i in np.nonzero(gn["is_leaf"] == 0)[0]]
    assert max(grown) > 0


def test_png_dataset_on_device_matches_arrays(ctx, config0, tmp_path):
    """loadImages from PNG files: equal-size images give the same integrals, labels and trained trees as the
    same frames set from arrays; a smaller image is padded as the reference pads integrated images (its own
    integral in the corner, colour NaN and depth 0 outside)."""
    rgb, depth, labels, im, ds = config0
    pal = np.array([[0, 0, 0], [200, 10, 10], [10, 200, 10], [10, 10, 200], [250, 250, 0], [0, 128, 255]],
                   np.uint8)
    for i in range(4):
        cf.write_png(tmp_path / f"f{i}_colors.png", rgb[i])
        cf.write_png(tmp_path / f"f{i}_depth.png", depth[i])
        cf.write_png(tmp_path / f"f{i}_ground_truth.png", pal[labels[i]])
    lm = cf.LabelMap()
    for c in pal:                      # colours known beforehand keep their labels
        lm.encode(c)
    data = cf.load_images(str(tmp_path), ignoredColors=["0,0,0"], label_map=lm)
    dim = data.to_images(ctx)
    for i in range(4):
        a, b = dim.integrals(i), im.integrals(i)
        assert all(_same_float_bits(x, y) for x, y in zip(a, b))
    params = dict(CONFIG0, maxDepth=6)
    f1 = cf.RandomForestImage.train(ctx, dim, cf.TrainingConfiguration(**params), tree_count=1, num_labels=6)
    f2 = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=1, num_labels=6)
    _assert_trees_equal(*f1.tree(0), *f2.tree(0))

    small = tmp_path / "small"
    small.mkdir()
    cf.write_png(small / "s_colors.png", rgb[0][:50, :70])
    cf.write_png(small / "s_depth.png", depth[0][:50, :70])
    cf.write_png(small / "s_ground_truth.png", pal[labels[0][:50, :70]])
    cf.write_png(small / "t_colors.png", rgb[1])
    cf.write_png(small / "t_depth.png", depth[1])
    cf.write_png(small / "t_ground_truth.png", pal[labels[1]])
    mixed = cf.load_images(str(small), ignoredColors=["0,0,0"], label_map=lm)
    assert (mixed.width, mixed.height) == (160, 120)
    mim = mixed.to_images(ctx)
    col, dep, val = mim.integrals(0)
    own = cf.Images(ctx, 1, 70, 50)
    own.set_rgbd_u8(0, rgb[0][:50, :70], depth[0][:50, :70])
    ocol, odep, oval = own.integrals(0)
    assert _same_float_bits(col[:, :50, :70], ocol)
    assert np.array_equal(dep[:50, :70], odep) and np.array_equal(val[:50, :70], oval)
    assert np.isnan(col[:, 50:, :]).all() and np.isnan(col[:, :, 70:]).all()
    assert (dep[50:] == 0).all() and (val[:, 70:] == 0).all()
    assert all(_same_float_bits(x, y) for x, y in zip(mim.integrals(1), im.integrals(1)))
    forest = cf.RandomForestImage.train(ctx, mim, cf.TrainingConfiguration(**params), tree_count=1, num_labels=6)
    assert len(forest.tree(0)[0]) > 1


def test_command_line_tools_end_to_end(ctx, config0, tmp_path, capfd):
    """curfil_train on a PNG folder writes tree files whose trees equal the library's training on the same
    frames; curfil_predict on a test folder writes prediction images equal to the library's prediction
    with the imported forest, and logs the pixel accuracies calculatePixelAccuracy gives."""
    rgb, depth, labels, im, ds = config0
    pal = np.array([[0, 0, 0], [200, 10, 10], [10, 200, 10], [10, 10, 200], [250, 250, 0], [0, 128, 255]],
                   np.uint8)
    train, test, out, pred = (tmp_path / d for d in ("train", "test", "out", "pred"))
    for d in (train, test):
        d.mkdir()
    for i in range(4):
        for folder, r, dp, lb in ((train, rgb, depth, labels),):
            cf.write_png(folder / f"f{i}_colors.png", r[i])
            cf.write_png(folder / f"f{i}_depth.png", dp[i])
            cf.write_png(folder / f"f{i}_ground_truth.png", pal[lb[i]])
    trgb, tdepth, tlabels = cf.synth_dataset(2, 160, 120, sites=12, classes=5, zero_frac=0.05, seed0=99)
    for i in range(2):
        cf.write_png(test / f"t{i}_colors.png", trgb[i])
        cf.write_png(test / f"t{i}_depth.png", tdepth[i])
        cf.write_png(test / f"t{i}_ground_truth.png", pal[tlabels[i]])
    # a smaller third test image: curfil_predict classifies a folder of mixed sizes image by image
    srgb, sdepth, slabels = cf.synth_dataset(1, 96, 80, sites=6, classes=5, zero_frac=0.05, seed0=123)
    cf.write_png(test / "t2_colors.png", srgb[0])
    cf.write_png(test / "t2_depth.png", sdepth[0])
    cf.write_png(test / "t2_ground_truth.png", pal[slabels[0]])
    args = [train, 2, 1000, 100, 32, 7, 60, 10, 10, out, "--useCIELab=false", "--ignoreColor", "0,0,0"]
    assert cf.run_tool("train", args) == 0
    paths = [out / "tree0.json.gz", out / "tree1.json.gz"]
    assert all(p.exists() for p in paths)
    back = cf.RandomForestImage.from_files(paths)
    data = cf.load_images(str(train), ignoredColors=["0,0,0"])
    dim = data.to_images(ctx)
    ign = [data.label_map.find((0, 0, 0))]
    cfg = cf.TrainingConfiguration(4711, 1000, 100, 32, 7, 60, 10, 10, ignored_labels=ign)
    ref = cf.RandomForestImage.train(ctx, dim, cfg, tree_count=2, num_labels=len(data.label_map))
    for t in range(2):
        _assert_trees_equal(*back.tree(t), *ref.tree(t))
    assert back.palette.tolist() == data.label_map.colors().tolist()
    capfd.readouterr()

    assert cf.run_tool("predict", [pred, test] + paths) == 0
    log = capfd.readouterr().err
    lm = cf.LabelMap()
    for c in back.palette:
        lm.encode(c)
    accs = []
    for i in range(3):
        tset = cf.Dataset.load([test / f"t{i}_colors.png"], label_map=lm)
        lab, _ = back.predict(tset.to_images(ctx), 0)
        written = cf.read_png(pred / f"t{i}_colors_prediction.png")
        assert np.array_equal(written, back.palette[lab])
        gt = tset.get(0)[2]
        accs.append(cf.calculatePixelAccuracy(lab, gt, True, ign))
    assert cf.read_png(pred / "t2_colors_prediction.png").shape == (80, 96, 3)
    m = re.search(r"pixel accuracy: ([0-9.]+)", log)
    assert m and abs(float(m.group(1)) - 100 * np.mean(accs)) < 1e-4, log[-2000:]
    assert "confusion matrix" in log


def test_cielab_dataset_integrals_match_oracle(ctx, config0, tmp_path):
    """a CIELab PNG dataset with one smaller image: its float planes integrate to the oracle's Kahan
    integrals (padded with NaN / 0 outside the smaller image)"""
    rgb, depth, labels, im, ds = config0
    pal = np.array([[0, 0, 0], [200, 10, 10], [10, 200, 10], [10, 10, 200], [250, 250, 0], [0, 128, 255]],
                   np.uint8)
    sizes = [(160, 120), (100, 70)]
    for i, (w, h) in enumerate(sizes):
        cf.write_png(tmp_path / f"f{i}_colors.png", rgb[i][:h, :w])
        cf.write_png(tmp_path / f"f{i}_depth.png", depth[i][:h, :w])
        cf.write_png(tmp_path / f"f{i}_ground_truth.png", pal[labels[i][:h, :w]])
    data = cf.load_images(str(tmp_path), useCIELab=True)
    dim = data.to_images(ctx)
    for i, (w, h) in enumerate(sizes):
        lab = data.color(i)
        ocol, odep, oval = orc.Dataset(lab[None], depth[i][:h, :w][None].astype(np.int32)).integrals(0)
        col, dep, val = dim.integrals(i)
        assert _same_float_bits(col[:, :h, :w].copy(), ocol)
        assert np.array_equal(dep[:h, :w], odep) and np.array_equal(val[:h, :w], oval)
        assert np.isnan(col[:, h:, :]).all() and np.isnan(col[:, :, w:]).all()


@pytest.mark.parametrize("mode", [cf.MODE_PLAIN, cf.MODE_BATCHED])
def test_many_thresholds_train_like_oracle(ctx, config0, mode):
    """T = 255 (the widest 8-bit bins; node histograms of 16 features need 96 KB of LDS, above the 64 KB
    default): trees equal the oracle's; histograms that exceed the device's LDS are refused with a message."""
    rgb, depth, labels, im, ds = config0
    params = dict(CONFIG0, featureCount=40, thresholds=255, maxDepth=5)
    forest = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=1, num_labels=6,
                                        mode=mode)
    of = ds.train_forest(orc.config(**params), 1, 6)
    _assert_trees_equal(*forest.tree(0), *of.tree(0))
    # 32 classes (pixel-uniform sampling: most classes have no pixels): 256 x 32 bins of 16 features need 512 KB
    wide = cf.TrainingConfiguration(**dict(params, subsamplingType="pixelUniform"))
    with pytest.raises(cf.CurfilError, match="too large for the LDS"):
        cf.RandomForestImage.train(ctx, im, wide, tree_count=1, num_labels=32, mode=mode)


def test_thirty_two_classes_train_and_predict_like_oracle(ctx):
    """32 label values (the reference's GPU limit): the class-wide score and classify paths equal the
    oracle's trees, posteriors and per-pixel predictions."""
    rgb, depth, labels = cf.synth_dataset(3, 96, 72, sites=40, classes=31, zero_frac=0.05, seed0=321)
    im = _images_u8(ctx, rgb, depth, labels)
    ds = orc.Dataset.from_u8(rgb, depth, labels)
    params = dict(CONFIG0, samplesPerImage=3200, featureCount=60, thresholds=12, maxDepth=6, minSampleCount=8,
                  subsamplingType="pixelUniform")
    C = int(labels.max()) + 1
    assert C == 32
    forest = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=2, num_labels=C,
                                        mode=cf.MODE_BATCHED)
    of = ds.train_forest(orc.config(**params), 2, C)
    forest.normalizeHistograms(0.0, True)
    trees, posts = [], []
    for t in range(2):
        gn, gh = forest.tree(t)
        on, oh = of.tree(t)
        _assert_trees_equal(gn, gh, on, oh)
        post = np.array([orc.normalize_histogram(h, of.prior(), 0.0, True) for h in oh])
        assert np.array_equal(forest.posteriors(t), post)
        trees.append(on)
        posts.append(post)
    gl, gp = forest.predict(im, 1)
    ol, op = ds.predict(1, trees, posts, C)
    assert np.array_equal(gl, ol)
    assert _same_float_bits(np.nan_to_num(gp), np.nan_to_num(op))


def test_colour_only_features_without_depth_images(ctx, config0):
    """useDepthImages = false: colour features only, every pixel at the dummy depth of 1000 mm
    (loadDummyDepthValues, image.cpp:379-394)."""
    rgb, depth, labels, im, ds = config0
    flat = np.full_like(depth, 1000)
    im2 = _images_u8(ctx, rgb, flat, labels)
    ds2 = orc.Dataset.from_u8(rgb, flat, labels)
    params = dict(CONFIG0, useDepthImages=False, maxDepth=6)
    forest = cf.RandomForestImage.train(ctx, im2, cf.TrainingConfiguration(**params), tree_count=1, num_labels=6)
    of = ds2.train_forest(orc.config(**params), 1, 6)
    gn, gh = forest.tree(0)
    _assert_trees_equal(gn, gh, *of.tree(0))
    assert np.all(gn["feature"][gn["is_leaf"] == 0][:, 0] == cf.COLOR)


REFDATA = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_testdata")


def test_reference_testdata_trains_like_oracle_and_predicts(ctx):
    """random_tree_image_test.cpp's trainTestGPU on the reference's own images (CIELab colour, seed 4711,
    500 samples per image, 500 features, 50 thresholds, depth 10, box radius 127, region 16, min samples
    100): the tree equals the oracle's on the same CIELab planes, and the prediction of testing1 is
    reported with its pixel accuracy (the reference's test expects about 73%)."""
    lm = cf.LabelMap()
    files = [os.path.join(REFDATA, f"training{i}_colors.png") for i in (1, 2, 3)]
    train = cf.Dataset.load(files, useCIELab=True, padding_color=(0, 0, 0), label_map=lm)
    im = train.to_images(ctx)
    C = len(lm)
    params = dict(randomSeed=4711, samplesPerImage=500, featureCount=500, minSampleCount=100, maxDepth=10,
                  boxRadius=127, regionSize=16, thresholds=50, maxImages=10)
    forest = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=1, num_labels=C)
    colors = np.stack([train.color(i) for i in range(3)])
    depth = np.stack([train.get(i)[1] for i in range(3)]).astype(np.int32)
    labels = np.stack([train.get(i)[2] for i in range(3)])
    ds = orc.Dataset(colors, depth, labels)
    of = ds.train_forest(orc.config(**params), 1, C)
    gn, gh = forest.tree(0)
    _assert_trees_equal(gn, gh, *of.tree(0))
    test = cf.Dataset.load([os.path.join(REFDATA, "testing1_colors.png")], useCIELab=True, label_map=lm)
    tim = test.to_images(ctx)
    forest.normalizeHistograms(0.0, True)
    pred, _ = forest.predict(tim, 0)
    acc = cf.calculatePixelAccuracy(pred, test.get(0)[2], True, [])
    print(f"testing1 pixel accuracy {100 * acc:.1f}% ({len(gn)} nodes)")
    assert acc > 0.5


def test_reference_testdata_ensemble(ctx):
    """random_tree_image_test.cpp's trainTestEnsemble (:172-237): 3 trees, 2000 samples per image, depth 14,
    box radius 120, region 10, min samples 32, seed 64684263; trees equal the oracle's, and testing1's
    pixel accuracy is reported (the reference's test expects about 76%)."""
    lm = cf.LabelMap()
    files = [os.path.join(REFDATA, f"training{i}_colors.png") for i in (1, 2, 3)]
    train = cf.Dataset.load(files, useCIELab=True, padding_color=(0, 0, 0), label_map=lm)
    im = train.to_images(ctx)
    C = len(lm)
    params = dict(randomSeed=64684263, samplesPerImage=2000, featureCount=500, minSampleCount=32, maxDepth=14,
                  boxRadius=120, regionSize=10, thresholds=50, maxImages=3)
    forest = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=3, num_labels=C)
    colors = np.stack([train.color(i) for i in range(3)])
    depth = np.stack([train.get(i)[1] for i in range(3)]).astype(np.int32)
    labels = np.stack([train.get(i)[2] for i in range(3)])
    of = orc.Dataset(colors, depth, labels).train_forest(orc.config(**params), 3, C)
    for t in range(3):
        _assert_trees_equal(*forest.tree(t), *of.tree(t))
    test = cf.Dataset.load([os.path.join(REFDATA, "testing1_colors.png")], useCIELab=True, label_map=lm)
    forest.normalizeHistograms(0.0, True)
    pred, _ = forest.predict(test.to_images(ctx), 0)
    acc = cf.calculatePixelAccuracy(pred, test.get(0)[2], True, [])
    print(f"testing1 pixel accuracy {100 * acc:.1f}% (3 trees, {[len(forest.tree(t)[0]) for t in range(3)]} nodes)")
    assert acc > 0.5


def test_reference_testdata_export_import(ctx, tmp_path):
    """import_export_test.cpp's testExportImport (:103-166) on training1 with colour 50,205,50 ignored: the
    configuration, label priors and every tree read back from tree<N>.json.gz equal what was written."""
    lm = cf.LabelMap()
    train = cf.Dataset.load([os.path.join(REFDATA, "training1_colors.png")], useCIELab=True,
                            padding_color=(50, 205, 50), label_map=lm)
    im = train.to_images(ctx)
    ign = lm.find((50, 205, 50))
    cfg = cf.TrainingConfiguration(4711, 500, 500, 32, 10, 50, 10, 10, numThreads=4, maxImages=10, useCIELab=True,
                                   ignored_labels=[ign])
    forest = cf.RandomForestImage.train(ctx, im, cfg, tree_count=3, num_labels=len(lm))
    paths = forest.export(tmp_path, cfg, palette=lm.colors(), folder_training=REFDATA)
    back = cf.RandomForestImage.from_files(paths)
    c = back.configuration
    assert (c.randomSeed, c.samplesPerImage, c.featureCount, c.minSampleCount, c.maxDepth, c.boxRadius,
            c.regionSize, c.thresholds, c.numThreads, c.maxImages) == (4711, 500, 500, 32, 10, 50, 10, 10, 4, 10)
    assert c.ignored_labels == [ign] and c.useCIELab
    assert back.palette.tolist() == lm.colors().tolist()
    for t in range(3):
        assert np.array_equal(back.prior(t), forest.prior())
        _assert_trees_equal(*back.tree(t), *forest.tree(t))
