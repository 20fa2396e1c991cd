"""Pins of the oracle's level loop and prediction by cases whose outcome is
forced by construction, whatever random features and thresholds are drawn.

* Level loop (random_tree.h:1182-1191, 1249-1415): the tree of
  helpers.forced_level_case for maxDepth 1..4 and minSampleCount around the
  node sizes: the root is level 1, a level equal to maxDepth is not
  evaluated, a child grows iff impure and holding at least minSampleCount
  samples, node ids are ++idNode per evaluated node (left, right) and rolled
  back by 2 for a node without a split.
* Prediction (random_forest_image.cpp:307-345): one-split trees over the
  three colour features of random_tree_image_gpu_test.cu:653-766, whose
  responses at four pixels are the reference's known answers (-2040, 1186,
  NaN, ...).  With one-hot leaves every pixel's path, its per-class vote
  count, its probabilities (votes / trees) and its label (first maximum) are
  fixed by those answers alone.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import forced_expected, forced_level_case
from test_oracle_pins import COMPLEX_RESPONSES, complex_case

FORCED_CFG = dict(randomSeed=4711, samplesPerImage=0, featureCount=100, boxRadius=20, regionSize=10, thresholds=10)


def forced_dataset():
    rgb, depth, roots = forced_level_case()
    return orc.Dataset.from_u8(rgb, depth), roots


def _check_forced(nodes, hist, max_depth, min_samples):
    want = forced_expected(max_depth, min_samples)
    assert len(nodes) == len(want), (max_depth, min_samples, len(nodes))
    for i, (nid, level, parent, left, right, n, leaf, h) in enumerate(want):
        got = (nodes["id"][i], nodes["level"][i], nodes["parent"][i], nodes["left"][i], nodes["right"][i],
               nodes["num_samples"][i], nodes["is_leaf"][i])
        assert got == (nid, level, parent, left, right, n, leaf), (max_depth, min_samples, i, got)
        assert hist[i].tolist() == h, (max_depth, min_samples, i)


@pytest.mark.parametrize("max_depth,min_samples", [(1, 40), (2, 40), (3, 40), (4, 40), (9, 40), (9, 41), (9, 39),
                                                   (9, 1), (2, 1)])
def test_forced_level_loop(max_depth, min_samples):
    ds, roots = forced_dataset()
    cfg = orc.config(minSampleCount=min_samples, maxDepth=max_depth, **FORCED_CFG)
    nodes, hist = ds.train_tree(cfg, 0, 5, 4711, roots)
    _check_forced(nodes, hist, max_depth, min_samples)
    if len(nodes) > 1:
        # the root separates group B (image 0) from the rest
        left = nodes["left"][0]
        assert nodes["num_samples"][left] == 40


@pytest.mark.parametrize("seed", [1, 2, 3, 4711])
def test_forced_level_loop_any_seed(seed):
    """The forced outcome does not depend on the features drawn: four tree seeds, the same tree."""
    ds, roots = forced_dataset()
    cfg = orc.config(minSampleCount=40, maxDepth=9, **dict(FORCED_CFG, randomSeed=seed))
    nodes, hist = ds.train_tree(cfg, 0, 5, seed, roots)
    _check_forced(nodes, hist, 9, 40)


def onehot_case():
    """The complex known-answer frames with depth images that give the four pixels the depths the
    reference test gives its samples (everything else missing depth), and three one-split trees."""
    color, _, feats, thr, samples = complex_case()
    N, _, H, W = color.shape
    depth = np.zeros((N, H, W), np.int32)
    for s in samples:
        depth[s["image"], s["y"], s["x"]] = s["depth_mm"]
    return color, depth, feats, thr, samples


def onehot_forest(feats, thresholds, C=2):
    """One tree per feature: root split (feature f, threshold), left leaf votes class 0, right leaf class 1."""
    trees, posts = [], []
    for f, t in enumerate(thresholds):
        nodes = np.zeros(3, orc.NODE_DTYPE)
        nodes["id"] = [0, 1, 2]
        nodes["level"] = [1, 2, 2]
        nodes["parent"] = [-1, 0, 0]
        nodes["left"] = [1, -1, -1]
        nodes["right"] = [2, -1, -1]
        nodes["is_leaf"] = [0, 1, 1]
        nodes["feature"][0] = feats[:, f]
        nodes["feature_id"][0] = f
        nodes["threshold"][0] = t
        post = np.zeros((3, C))
        post[1, 0] = 1.0
        post[2, 1] = 1.0
        trees.append(nodes)
        posts.append(post)
    return trees, posts


def onehot_expected(thresholds, features):
    """Votes at the four known-answer pixels from the reference's responses alone: left iff r <= t (NaN
    never), class 0 on the left."""
    r = COMPLEX_RESPONSES[features]                       # [trees][samples]
    t = np.asarray(thresholds, np.float32)[:, None]
    left = ~np.isnan(r) & (r <= t)
    votes = np.stack([left.sum(0), (~left).sum(0)])       # [2][samples]
    return votes


@pytest.mark.parametrize("row", [0, 1, 2, 3])
@pytest.mark.parametrize("features", [(0, 1, 2), (0, 1), (2,), (1, 2, 0, 1)])
def test_onehot_forest_prediction_known_answers(row, features):
    """Rows 0 and 1: the reference test's thresholds; rows 2 and 3: thresholds equal to the responses at
    the third and the second pixel (as sampled thresholds are), which that pixel sends left."""
    color, depth, feats, thr, samples = onehot_case()
    ds = orc.Dataset(color, depth)
    table = np.concatenate([thr, COMPLEX_RESPONSES[:, [2, 1]].T.astype(np.float32)])
    thresholds = [table[row, f] for f in features]
    trees, posts = onehot_forest(feats[:, list(features)], thresholds)
    votes = onehot_expected(thresholds, list(features))
    n = len(features)
    for k, s in enumerate(samples):
        labels, probs = ds.predict(int(s["image"]), trees, posts, 2)
        x, y = int(s["x"]), int(s["y"])
        want_p = [np.float32(votes[c, k] / n) for c in range(2)]
        assert [probs[c, y, x] for c in range(2)] == want_p, (features, row, k)
        # first maximum wins (strict >), so a 1:1 vote is class 0
        assert labels[y, x] == (1 if votes[1, k] > votes[0, k] else 0), (features, row, k)
    # a pixel with missing depth answers NaN for every feature: all votes right
    labels, probs = ds.predict(0, trees, posts, 2)
    assert labels[0, 0] == 1 and probs[1, 0, 0] == np.float32(1.0) and probs[0, 0, 0] == 0.0


def test_onehot_prediction_reads_depth_at_the_pixel():
    """The pixel's own depth decides validity (the reference's GPU classifier reads (x+1, y+1),
    random_tree_image_gpu.cu:1255-1258; the CPU branch reads (x, y)): the pixels right-below the
    known-answer pixels have missing depth and vote right in every tree."""
    color, depth, feats, thr, samples = onehot_case()
    ds = orc.Dataset(color, depth)
    trees, posts = onehot_forest(feats, thr[1])
    for s in samples:
        labels, probs = ds.predict(int(s["image"]), trees, posts, 2)
        x, y = int(s["x"]) + 1, int(s["y"]) + 1
        if depth[s["image"], y, x] == 0:
            assert probs[1, y, x] == np.float32(1.0) and labels[y, x] == 1
