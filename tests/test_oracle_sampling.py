"""Pixel sampling of the oracle (CPU): pixelUniform subsampling and maxImages image subsampling.

The reference's tests hold no vectors for either (its only end-to-end training test needs PNG input,
CIELab and a CUDA device), so these functions are "parity unpinned" beyond the invariants below:
  * pixelUniform (random_tree_image.cpp:840-883): samplesPerImage/trees samples per image of the tree,
    every label read from the label image at the sample's pixel, no ignored label, invalid depth as
    Depth::INVALID (-1), the root block sort (random_tree_image.cpp:776-802), images drawn uniformly;
  * maxImages (random_forest_image.cpp:161-171): a tree samples from at most maxImages images, and its
    label prior counts exactly the labels of those images (random_tree_image.cpp:885-909);
  * trees with maxImages unset keep the forest's prior;
  * horizontal flipping (random_tree.h:122-144, 521-527, 1295-1360): root samples are Both and count twice
    in the root histogram; every split's children histograms add up to the parent's; a tree grown with
    flipping differs from one grown without.
The flipped response itself is pinned by a mathematical identity (the mirror test at the end): copies of
the oracle that keep the second box's x offset or negate a y offset fail it.
"""
import numpy as np
import pytest

import oracle as orc
from helpers import CONFIG0, config0_dataset


@pytest.fixture(scope="module")
def data():
    rgb, depth, labels = config0_dataset()
    return rgb, depth, labels, orc.Dataset.from_u8(rgb, depth, labels)


def _sorted_in_blocks(roots):
    key = (roots["image"].astype(np.int64) << 40) | (roots["label"].astype(np.int64) << 32) | \
          ((roots["y"] // 10).astype(np.int64) << 16) | (roots["x"] // 10).astype(np.int64)
    return all(np.all(np.diff(key[b:b + 200000]) >= 0) for b in range(0, len(key), 200000))


def test_pixel_uniform_invariants(data):
    rgb, depth, labels, ds = data
    params = dict(CONFIG0, maxDepth=2, subsamplingType="pixelUniform")
    of = ds.train_forest(orc.config(**params), 2, 6)
    for t in range(2):
        r = of.roots(t)
        assert len(r) == (params["samplesPerImage"] // 2) * len(labels)
        assert np.array_equal(r["label"], labels[r["image"], r["y"], r["x"]])
        assert not np.any(r["label"] == 0)            # ignored
        raw = depth[r["image"], r["y"], r["x"]].astype(np.int32)
        assert np.array_equal(r["depth_mm"], np.where(raw > 0, raw, -1))
        assert np.any(r["depth_mm"] == -1)            # pixels with missing depth are kept
        assert _sorted_in_blocks(r)
        counts = np.bincount(r["image"], minlength=len(labels))
        assert counts.min() > 0.7 * counts.mean()     # images drawn uniformly
        assert np.array_equal(of.prior(t), of.prior())
    assert not np.array_equal(of.roots(0), of.roots(1))


def test_max_images_subsets_and_priors(data):
    rgb, depth, labels, ds = data
    params = dict(CONFIG0, maxDepth=2, maxImages=2)
    of = ds.train_forest(orc.config(**params), 3, 6)
    full = np.bincount(labels.ravel(), minlength=6)
    assert np.array_equal(of.prior(), full)
    subsets = []
    for t in range(3):
        r = of.roots(t)
        imgs = np.unique(r["image"])
        assert len(imgs) == 2
        subsets.append(tuple(imgs))
        assert np.array_equal(of.prior(t), np.bincount(labels[imgs].ravel(), minlength=6))
        # class-uniform over the tree's images: samplesPerClass = ceil(2 * (1000 // 3) / 5)
        assert len(r) == int(np.ceil(2 * (1000 // 3) / 5)) * 5
    assert len(set(subsets)) > 1
    # more images allowed than there are: every image, the forest's prior
    of = ds.train_forest(orc.config(**dict(params, maxImages=10)), 1, 6)
    assert np.array_equal(of.prior(0), full)
    assert len(np.unique(of.roots(0)["image"])) == 4


def _check_histograms(nodes, hist):
    for i, n in enumerate(nodes):
        if n["is_leaf"] == 0:
            assert np.array_equal(hist[n["left"]] + hist[n["right"]], hist[i]), i


def test_horizontal_flipping_invariants(data):
    rgb, depth, labels, ds = data
    params = dict(CONFIG0, maxDepth=7, horizontalFlipping=True)
    of = ds.train_forest(orc.config(**params), 1, 6)
    r = of.roots(0)
    assert np.all(r["flip"] == 2)
    nodes, hist = of.tree(0)
    assert np.array_equal(hist[0], 2 * np.bincount(r["label"], minlength=6))
    assert nodes["num_samples"][0] == len(r)
    _check_histograms(nodes, hist)
    # a node's histogram weighs at least its samples and at most twice them
    assert np.all(hist.sum(1) >= nodes["num_samples"]) and np.all(hist.sum(1) <= 2 * nodes["num_samples"])
    # Both samples separated by a split continue as NoFlip + Flip: children hold more samples than the parent
    inner = np.nonzero(nodes["is_leaf"] == 0)[0]
    assert len(inner) > 3
    grown = [nodes["num_samples"][nodes["left"][i]] + nodes["num_samples"][nodes["right"][i]] - nodes["num_samples"][i]
             for i in inner]
    assert min(grown) >= 0 and max(grown) > 0
    plain = ds.train_forest(orc.config(**dict(params, horizontalFlipping=False)), 1, 6)
    pn, ph = plain.tree(0)
    _check_histograms(pn, ph)
    assert len(pn) != len(nodes) or not np.array_equal(pn["feature_id"], nodes["feature_id"])


def test_flipped_response_equals_mirrored_image_response():
    """The flipped response (random_tree_image.h:564-620: both boxes' x offsets negated, y offsets and
    regions kept) pinned by the mirror identity.  A box of averageRegionColor / averageRegionDepth
    (random_tree_image.h:288-316) spans the columns x - w + 1 .. x + w between its integral corners, so
    pixel x of an image corresponds to pixel W - 2 - x of the horizontally mirrored image: the flipped
    response at (x, y) equals the unflipped response of the mirrored image at (W - 2 - x, y), with the same
    sample depth, wherever both are finite (a box touching an edge can differ: column 0 never lies inside
    a box, column W - 1 can).  u8 colour on a small image keeps every colour prefix sum an exact float,
    and depth sums are integers, so the identity holds bit for bit.  Without the flip it fails."""
    from helpers import random_features, random_samples, random_u8_dataset
    rng = np.random.default_rng(77)
    W, H, n = 64, 48, 2
    rgb, depth, _ = random_u8_dataset(n, W, H, seed=5)
    ds = orc.Dataset.from_u8(rgb, depth)
    dm = orc.Dataset.from_u8(np.ascontiguousarray(rgb[:, :, ::-1]), np.ascontiguousarray(depth[:, :, ::-1]))
    feats = random_features(rng, 300, box_radius=20, region_size=6)
    s = random_samples(rng, 3000, n, W - 1, H, depth)   # x <= W - 2
    s = s[s["depth_mm"] > 0]
    s["flip"] = 1
    m = s.copy()
    m["x"] = W - 2 - s["x"]
    m["flip"] = 0
    a, b = ds.responses(s, feats), dm.responses(m, feats)
    both = np.isfinite(a) & np.isfinite(b)
    assert both.mean() > 0.3
    assert np.array_equal(a[both], b[both])
    s0 = s.copy()
    s0["flip"] = 0
    c = ds.responses(s0, feats)
    both0 = np.isfinite(c) & np.isfinite(b)
    assert np.mean(c[both0] != b[both0]) > 0.5
