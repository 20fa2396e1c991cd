"""GPU parity at the benchmark workloads' shapes.

The HIP path against the CPU oracle on the configurations the benchmark
runs, cut to sizes the oracle finishes in seconds:

* full-frame training (640x480 frames of the synthetic generator with 30
  sites, 4 classes + void and 8% missing depth; 3 trees, F = 2000, T = 20,
  boxRadius 120, regionSize 20, minSampleCount 50) on 10 frames, plain and
  batched, the batched levels split into several feature-group batches;
* batched training (5 trees, F = 4000, T = 32) on 4 frames;
* the small workload with device budgets that force several feature-group
  and node-slot batches per level, with and without horizontal flipping;
* dense prediction with 10 complete trees of depth 20 on 640x480 frames;
* the serving loop with frames of 320x240, 640x480 and 1280x960.

Bar: trees equal node for node (ids, structure, features, threshold bits,
histograms; scores within 4e-16 of the FP64 oracle), labels and float
posteriors bit-exact.
"""
import os

import numpy as np
import pytest

import curfil_amd as cf
import oracle as orc
from helpers import CONFIG0, config0_dataset

pytestmark = pytest.mark.gpu

FULL_FRAME = dict(randomSeed=4711, featureCount=2000, minSampleCount=50, boxRadius=120, regionSize=20,
                  thresholds=20, ignored_labels=(0,))


def _same_float_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return np.array_equal(a.view(np.uint32 if a.dtype == np.float32 else np.uint64),
                          b.view(np.uint32 if b.dtype == np.float32 else np.uint64))


def _images_u8(ctx, rgb, depth, labels=None):
    n, H, W, _ = rgb.shape
    im = cf.Images(ctx, n, W, H)
    for i in range(n):
        im.set_rgbd_u8(i, rgb[i], depth[i])
        if labels is not None:
            im.set_labels(i, labels[i])
    return im


def _assert_trees_equal(gn, gh, on, oh):
    assert len(gn) == len(on), (len(gn), len(on))
    for k in ("id", "level", "parent", "left", "right", "num_samples", "is_leaf"):
        assert np.array_equal(gn[k], on[k]), k
    inner = gn["is_leaf"] == 0
    assert np.array_equal(gn["feature_id"][inner], on["feature_id"][inner])
    assert np.array_equal(gn["feature"][inner], on["feature"][inner])
    assert _same_float_bits(gn["threshold"][inner], on["threshold"][inner])
    assert np.abs(gn["score"][inner] - on["score"][inner]).max(initial=0) <= 4e-16
    assert np.array_equal(gh, oh)


class _Env:
    """Sets environment variables for the duration of a block (the trainer reads its budgets per call)."""

    def __init__(self, **kv):
        self.kv = kv
        self.old = {}

    def __enter__(self):
        for k, v in self.kv.items():
            self.old[k] = os.environ.get(k)
            os.environ[k] = str(v)

    def __exit__(self, *exc):
        for k, v in self.old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v


def _compare_forest(forest, of, trees):
    assert np.array_equal(forest.prior(), of.prior())
    for t in range(trees):
        assert np.array_equal(forest.root_samples(t), of.roots(t))
        gn, gh = forest.tree(t)
        on, oh = of.tree(t)
        _assert_trees_equal(gn, gh, on, oh)


# ------------------------------------------------------ several batches per level
@pytest.fixture(scope="module")
def small(ctx):
    rgb, depth, labels = config0_dataset()
    return _images_u8(ctx, rgb, depth, labels), orc.Dataset.from_u8(rgb, depth, labels)


@pytest.mark.parametrize("flip", [False, True])
def test_multi_batch_levels_match_oracle(ctx, small, flip):
    """Budgets of a few KiB make every batched level run pass 1 / pass 2 in several feature-group batches
    (the flipped rows then sit at a batch-dependent offset) and the node histograms in several node-slot
    batches, as the 800-frame workloads do; the trees still equal the oracle's."""
    im, ds = small
    params = dict(CONFIG0, maxDepth=9, horizontalFlipping=flip)
    with _Env(CURFIL_BINS_BUDGET_KB=160, CURFIL_EVAL_BUDGET_KB=96):
        forest = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=2,
                                            num_labels=6, mode=cf.MODE_BATCHED)
    prof = forest.profile()
    print(f"flip {flip}: group batches {prof['group_batches_max']:.0f}, slot batches {prof['slot_batches_max']:.0f}, "
          f"nodes {[len(forest.tree(t)[0]) for t in range(2)]}")
    assert prof["group_batches_max"] >= 2 and prof["slot_batches_max"] >= 2
    _compare_forest(forest, ds.train_forest(orc.config(**params), 2, 6), 2)


# ------------------------------------------------------ full-frame training
@pytest.fixture(scope="module")
def full_frames(ctx):
    n = 10
    rgb, depth, labels = cf.synth_dataset(n, 640, 480, sites=30, classes=4, zero_frac=0.08, seed0=4711)
    return _images_u8(ctx, rgb, depth, labels), orc.Dataset.from_u8(rgb, depth, labels), (rgb, depth, labels)


@pytest.fixture(scope="module")
def full_frame_oracle(full_frames):
    _, ds, _ = full_frames
    params = dict(FULL_FRAME, samplesPerImage=12000, maxDepth=5)
    return params, ds.train_forest(orc.config(**params), 3, 5)


@pytest.mark.parametrize("mode", [cf.MODE_BATCHED, cf.MODE_PLAIN])
def test_full_frame_training_matches_oracle(ctx, full_frames, full_frame_oracle, mode):
    """The full-frame workload's shape on 10 frames of 640x480: 3 trees, 4000 samples per image and tree
    (40000 root samples each), F = 2000, T = 20, boxRadius 120, regionSize 20, minSampleCount 50, 4 levels.
    Batched levels run with a 16 MiB bin budget, so pass 1 and pass 2 take several feature-group batches."""
    im, _, _ = full_frames
    params, of = full_frame_oracle
    with _Env(CURFIL_BINS_BUDGET_MB=16):
        forest = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=3,
                                            num_labels=5, mode=mode)
    prof = forest.profile()
    print(f"mode {mode}: nodes {[len(forest.tree(t)[0]) for t in range(3)]}, "
          f"group batches {prof['group_batches_max']:.0f}, replays {prof['replays']:.0f}")
    if mode == cf.MODE_BATCHED:
        assert prof["group_batches_max"] >= 2
    _compare_forest(forest, of, 3)


def test_batched_training_shape_matches_oracle(ctx, full_frames):
    """The batched workload's shape: 5 trees trained together, F = 4000, T = 32, 4000 samples per image
    and tree, on 4 of the 640x480 frames, 4 levels."""
    _, _, (rgb, depth, labels) = full_frames
    im = _images_u8(ctx, rgb[:4], depth[:4], labels[:4])
    ds = orc.Dataset.from_u8(rgb[:4], depth[:4], labels[:4])
    params = dict(FULL_FRAME, featureCount=4000, thresholds=32, samplesPerImage=20000, maxDepth=5)
    forest = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=5, num_labels=5,
                                        mode=cf.MODE_BATCHED)
    print(f"nodes {[len(forest.tree(t)[0]) for t in range(5)]}, profile {forest.profile()}")
    _compare_forest(forest, ds.train_forest(orc.config(**params), 5, 5), 5)


# ------------------------------------------------------ dense prediction
@pytest.fixture(scope="module")
def dense_forest(ctx, full_frames):
    im, _, (rgb, depth, labels) = full_frames
    forest = cf.synth_forest(ctx, im, depth.astype(np.int32), trees=10, depth=20, classes=5, seed=4711)
    forest.normalizeHistograms(0.0, False)
    return forest


def test_dense_forest_prediction_matches_oracle(ctx, full_frames, dense_forest):
    """The dense-prediction forest (10 complete trees of depth 20, random splits over the full-frame feature
    space, Dirichlet(1) leaves over 5 classes) on two 640x480 frames: labels and float posteriors equal the
    oracle's CPU prediction bit for bit, single-frame and batched.  A single frame (0.31 Mpix) goes through
    the kernel that splits each tile's trees over a workgroup's waves; the batch of 5 frames (1.54 Mpix)
    through the one-thread-per-pixel kernel.  A batch of 10 frames runs as two launches whose results are
    copied while the other runs: into page-locked arrays directly, and into ordinary (pageable) arrays
    through the page-locked staging halves.  Every copy of a frame equals the oracle."""
    im, ds, _ = full_frames
    trees, posts = [], []
    ones = np.ones(5, np.uint32)
    for t in range(10):
        nodes, hist = dense_forest.tree(t)
        assert len(nodes) == (1 << 20) - 1
        trees.append(nodes)
        posts.append(dense_forest.posteriors(t))
    # the posteriors of a few leaves against the oracle's normalisation
    nodes, hist = dense_forest.tree(3)
    leaves = np.nonzero(nodes["is_leaf"])[0][::65536]
    post = np.array([orc.normalize_histogram(hist[i], ones, 0.0, False) for i in leaves])
    assert np.array_equal(posts[3][leaves], post)
    bl, bp = dense_forest.predict_batch(im, [7, 2])
    ql, qp = dense_forest.predict_batch(im, [2, 7, 2, 7, 2])
    ten = [7, 2] * 5
    tl, tp = dense_forest.predict_batch(im, ten)
    pl = cf.pinned_empty(ctx, tl.shape, np.uint8)
    pp = cf.pinned_empty(ctx, tp.shape, np.float32)
    dense_forest.predict_batch(im, ten, out=(pl, pp))
    for k, img in enumerate((7, 2)):
        ol, op = ds.predict(img, trees, posts, 5)
        gl, gp = dense_forest.predict(im, img)
        assert np.array_equal(gl, ol), img
        assert _same_float_bits(gp, op), img
        assert np.array_equal(bl[k], ol) and _same_float_bits(bp[k], op), img
        for q in (1 - k, 3 - k):
            assert np.array_equal(ql[q], ol) and _same_float_bits(qp[q], op), img
        for q in range(k, 10, 2):
            assert np.array_equal(tl[q], ol) and _same_float_bits(tp[q], op), (img, q)
            assert np.array_equal(pl[q], ol) and _same_float_bits(pp[q], op), (img, q)
        print(f"frame {img}: label histogram {np.bincount(ol.ravel(), minlength=5)}")


# ------------------------------------------------------ serving at the workload's frame sizes
SERVE_POLICIES = {
    "default": {},
    "bands": {"CURFIL_SERVE_BAND_MPIX": 0.1, "CURFIL_SERVE_LAUNCH_MPIX": 0.2},
    "whole": {"CURFIL_SERVE_BAND_MPIX": 0, "CURFIL_SERVE_LAUNCH_MPIX": 0},
    "gated_unsplit": {"CURFIL_SERVE_ADMIT_MPIX": 0.1, "CURFIL_CLASSIFY_SPLIT_MPIX": 0},
}


@pytest.mark.parametrize("policy", list(SERVE_POLICIES))
def test_serving_at_workload_frame_sizes(ctx, full_frames, dense_forest, policy):
    """The serving loop with the dense forest resident and frames of 320x240, 640x480 and 1280x960
    interleaved (with a budget that makes frames wait for memory): each frame's labels and posteriors
    equal its single-frame prediction at full size; the 1280x960 frame (depth sums beyond 2^31 in its
    integral) also equals the oracle.  Policies: the loop's defaults (0.35 Mpix row bands, 1.25 Mpix
    launches, every launch by the tree-split kernel); bands above 0.1 Mpix with launches of at most
    0.2 Mpix, so bands of one frame run in different launches, between other frames; whole frames in
    unbounded launches; admission gated at 0.1 Mpix (one frame admitted at a time) with every launch
    by the one-thread-per-pixel kernel."""
    sizes = [(320, 240), (640, 480), (1280, 960)]
    frames, refs = [], []
    big = None
    for k, (W, H) in enumerate(sizes):
        rgb, depth, labels = cf.synth_dataset(2, W, H, sites=30, classes=4, zero_frac=0.08, seed0=900 + 10 * k)
        im = _images_u8(ctx, rgb, depth)
        for i in range(2):
            frames.append((rgb[i], depth[i]))
            refs.append(dense_forest.predict(im, i))
        if W == 1280:
            big = (rgb, depth)
    order = [0, 2, 4, 1, 3, 5, 2, 4]
    budget = 1280 * 960 * (5 + 20 + 12 + 8 + 5 * 4 + 1) / (1 << 20) * 1.5
    with _Env(**SERVE_POLICIES[policy]):
        server = cf.Server(ctx, dense_forest, budget_mb=budget, max_batch=4)
    ids = [server.submit(*frames[j]) for j in order]
    for rid, j in zip(ids, order):
        lab, probs, lat = server.wait(rid)
        rl, rp = refs[j]
        assert np.array_equal(lab, rl) and _same_float_bits(probs, rp), j
    assert server.stats()["frames"] == len(order)
    server.close()
    rgb, depth = big
    ds = orc.Dataset.from_u8(rgb[:1], depth[:1])
    trees = [dense_forest.tree(t)[0] for t in range(10)]
    posts = [dense_forest.posteriors(t) for t in range(10)]
    ol, op = ds.predict(0, trees, posts, 5)
    assert np.array_equal(refs[4][0], ol) and _same_float_bits(refs[4][1], op)


# ------------------------------------------------------ forced outcomes (known answers)
@pytest.mark.parametrize("mode", [cf.MODE_BATCHED, cf.MODE_PLAIN])
@pytest.mark.parametrize("max_depth,min_samples", [(1, 40), (2, 40), (3, 40), (9, 40), (9, 41), (9, 39)])
def test_forced_level_loop_on_device(ctx, mode, max_depth, min_samples):
    """The tree of the forced case (its outcome does not depend on the features drawn): maxDepth, the
    minSampleCount boundary, a node without a split and the rolled-back ids, on the device path."""
    from helpers import forced_expected, forced_level_case
    from test_oracle_level_loop import FORCED_CFG
    rgb, depth, roots = forced_level_case()
    im = _images_u8(ctx, rgb, depth)
    params = dict(FORCED_CFG, minSampleCount=min_samples, maxDepth=max_depth)
    params.pop("samplesPerImage")
    cfg = cf.TrainingConfiguration(samplesPerImage=1, **params)
    forest = cf.RandomForestImage.train_tree_from_samples(ctx, im, cfg, 0, 5, 4711, roots, mode=mode)
    nodes, hist = forest.tree(0)
    want = forced_expected(max_depth, min_samples)
    assert len(nodes) == len(want)
    for i, (nid, level, parent, left, right, n, leaf, h) in enumerate(want):
        got = (nodes["id"][i], nodes["level"][i], nodes["parent"][i], nodes["left"][i], nodes["right"][i],
               nodes["num_samples"][i], nodes["is_leaf"][i])
        assert got == (nid, level, parent, left, right, n, leaf), (i, got)
        assert hist[i].tolist() == h
    ds = orc.Dataset.from_u8(rgb, depth)
    on, oh = ds.train_tree(orc.config(samplesPerImage=1, **params), 0, 5, 4711, roots)
    _assert_trees_equal(nodes, hist, on, oh)


@pytest.mark.parametrize("row", [0, 1, 2, 3])
def test_onehot_forest_known_answers_on_device(ctx, row):
    """One-split trees over the reference's known-answer colour features with one-hot leaves: the device
    prediction's probabilities (votes / trees) and labels at the four known-answer pixels follow from the
    reference's responses alone, and every pixel equals the oracle."""
    from test_oracle_level_loop import COMPLEX_RESPONSES, onehot_case, onehot_expected, onehot_forest
    color, depth, feats, thr, samples = onehot_case()
    n_img, _, H, W = color.shape
    im = cf.Images(ctx, n_img, W, H)
    for i in range(n_img):
        im.set_rgbd_f32(i, color[i], depth[i])
    table = np.concatenate([thr, COMPLEX_RESPONSES[:, [2, 1]].T.astype(np.float32)])
    features = [0, 1, 2]
    thresholds = [table[row, f] for f in features]
    trees, posts = onehot_forest(feats, thresholds)
    hists = []
    for p in posts:
        hists.append((p * 7).astype(np.uint32))
    forest = cf.RandomForestImage.from_trees(ctx, list(zip(trees, hists)), 2)
    forest.normalizeHistograms(0.0, False)
    votes = onehot_expected(thresholds, features)
    ds = orc.Dataset(color, depth)
    for k, s in enumerate(samples):
        img, x, y = int(s["image"]), int(s["x"]), int(s["y"])
        gl, gp = forest.predict(im, img)
        assert [gp[c, y, x] for c in range(2)] == [np.float32(votes[c, k] / 3) for c in range(2)], (row, k)
        assert gl[y, x] == (1 if votes[1, k] > votes[0, k] else 0)
        ol, op = ds.predict(img, trees, [forest.posteriors(t) for t in range(3)], 2)
        assert np.array_equal(gl, ol) and _same_float_bits(gp, op)


# ------------------------------------------------------ pass-1 launch shapes (the tuner's candidates)
@pytest.mark.parametrize("shape", range(len(cf.eval_shapes())))
def test_every_launch_shape_trains_like_oracle(ctx, small, shape):
    """Each pass-1 launch shape the tuner can choose, pinned, with several feature-group and node-slot
    batches per level and horizontal flipping: the trees equal the oracle's."""
    im, ds = small
    params = dict(CONFIG0, maxDepth=8, horizontalFlipping=True)
    with _Env(CURFIL_EVAL_SHAPE=shape, CURFIL_BINS_BUDGET_KB=160, CURFIL_EVAL_BUDGET_KB=96):
        forest = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=2,
                                            num_labels=6, mode=cf.MODE_BATCHED)
    assert forest.profile()["eval_shape"] == shape
    print(f"shape {shape} {cf.eval_shape_info(shape)}: nodes {[len(forest.tree(t)[0]) for t in range(2)]}")
    _compare_forest(forest, _small_oracle_flip(ds), 2)


def test_captured_level_graphs_and_eager_launches_train_like_oracle(ctx, small):
    """With CURFIL_LEVEL_GRAPH=1, levels whose pass-1 shape is settled run as one captured graph each
    (several feature-group and node-slot batches per level here); the same training with every kernel
    launched on its own (the default, CURFIL_LEVEL_GRAPH=0, and a profiling context, whose phase timers
    need events on the stream) gives the same trees, all equal to the oracle's."""
    im, ds = small
    params = dict(CONFIG0, maxDepth=8, horizontalFlipping=True)
    forests = {}
    for mode in ("graphs", "eager", "profiled"):
        ctx.set_profiling(mode == "profiled")
        try:
            with _Env(CURFIL_LEVEL_GRAPH=0 if mode == "eager" else 1, CURFIL_BINS_BUDGET_KB=160,
                      CURFIL_EVAL_BUDGET_KB=96):
                forests[mode] = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params),
                                                           tree_count=2, num_labels=6, mode=cf.MODE_BATCHED)
        finally:
            ctx.set_profiling(False)
    prof = {m: f.profile() for m, f in forests.items()}
    print({m: (p["level_graphs"], p["eval_launches"], p["group_batches_max"], p["slot_batches_max"])
           for m, p in prof.items()})
    assert prof["graphs"]["level_graphs"] >= 1
    assert prof["eager"]["level_graphs"] == 0 and prof["profiled"]["level_graphs"] == 0
    assert prof["graphs"]["group_batches_max"] > 1 and prof["graphs"]["slot_batches_max"] > 1
    assert prof["graphs"]["eval_launches"] == prof["eager"]["eval_launches"] == prof["profiled"]["eval_launches"]
    assert prof["profiled"]["eval_kernel_ms"] > 0 and prof["profiled"]["hist_ms"] > 0
    for f in forests.values():
        _compare_forest(f, _small_oracle_flip(ds), 2)


_ORACLE_CACHE = {}


def _small_oracle_flip(ds):
    if "flip8" not in _ORACLE_CACHE:
        params = dict(CONFIG0, maxDepth=8, horizontalFlipping=True)
        _ORACLE_CACHE["flip8"] = ds.train_forest(orc.config(**params), 2, 6)
    return _ORACLE_CACHE["flip8"]


def test_tuner_chooses_a_shape_and_trees_stay_exact(ctx, full_frames, full_frame_oracle):
    """With the tuner on (and allowed to time small levels), a full-frame training times every candidate
    shape on its first pass-1 launches, each after a launch of the default shape (one round here), and keeps
    one; the trees equal the oracle's."""
    im, _, _ = full_frames
    params, of = full_frame_oracle
    cf.eval_tuning_reset()
    with _Env(CURFIL_BINS_BUDGET_KB=1024, CURFIL_TUNE=1, CURFIL_TUNE_SMALL=1, CURFIL_TUNE_ROUNDS=1):
        forest = cf.RandomForestImage.train(ctx, im, cf.TrainingConfiguration(**params), tree_count=3,
                                            num_labels=5, mode=cf.MODE_BATCHED)
    table = cf.eval_tuning()
    print(f"tuning table: {table}")
    assert table and all(r["launches_timed"] >= 2 * (len(cf.eval_shapes()) - 1) for r in table)
    assert all(r["shape"] >= 0 for r in table)
    cf.eval_tuning_reset()
    _compare_forest(forest, of, 3)
