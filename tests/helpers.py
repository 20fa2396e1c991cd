"""Test data builders shared by the CPU and GPU suites."""
import numpy as np

import oracle as orc

GOLDEN_FEATURE_HASH = 4763250482941632261  # feature_generation_test.cpp:386
M64 = 1 << 64


def golden_hash_dataset():
    """The data of feature_generation_test.cpp:310-361, drawn with the oracle's
    Boost-equivalent samplers: 10 images 400x400 with float colours k/100 and
    depth Depth(k/10.0f), and 500 samples per image labelled image % 10."""
    W = H = 400
    n_img, spi = 10, 500
    # RandomSource(4711): width 4711, height 4712, pixel 4713, depth 4714
    pix = orc.sampler_draws(4713, 0, 999, n_img * H * W * 3).reshape(n_img, H, W, 3)
    dep = orc.sampler_draws(4714, 1, 90, n_img * H * W).reshape(n_img, H, W)
    color = (pix.astype(np.float32) / np.float32(100.0)).transpose(0, 3, 1, 2).copy()
    depth_m = dep.astype(np.float32) / np.float32(10.0)
    depth = np.floor(depth_m.astype(np.float64) * 1000.0).astype(np.int32)
    xs = orc.sampler_draws(4711, 0, W - 1, n_img * spi)
    ys = orc.sampler_draws(4712, 0, H - 1, n_img * spi)
    samples = np.zeros(n_img * spi, orc.SAMPLE_DTYPE)
    samples["x"], samples["y"] = xs, ys
    samples["image"] = np.repeat(np.arange(n_img), spi)
    samples["label"] = samples["image"] % 10
    samples["depth_mm"] = depth[samples["image"], ys, xs]
    return color, depth, samples


def boost_feature_hash(feats, thr):
    """analyzeFeatures' hash (feature_generation_test.cpp:291-307) in Python."""
    h = 0

    def upd(v):
        nonlocal h
        h ^= (v + 0x9e3779b9 + ((h << 6) % M64) + (h >> 2)) % M64

    F = feats.shape[1]
    for row in (0, 9, 10, 1, 2, 3, 4, 5, 6, 7, 8):
        for f in range(F):
            v = int(feats[row, f])
            if row in (9, 10):
                v &= 0xFF
            upd(v % M64)
    for f in range(F):
        for t in range(thr.shape[0]):
            upd(int(np.trunc(thr[t, f])) % M64)
    return h


def random_u8_dataset(n, W, H, seed, zero_frac=0.05, classes=5):
    """Random colour and depth with smooth structure so boxes see varied content."""
    rng = np.random.default_rng(seed)
    rgb = rng.integers(0, 256, (n, H, W, 3)).astype(np.uint8)
    yy, xx = np.mgrid[0:H, 0:W]
    depth = np.zeros((n, H, W), np.uint16)
    labels = np.zeros((n, H, W), np.uint8)
    for i in range(n):
        d = 500 + (xx * rng.uniform(0, 8) + yy * rng.uniform(0, 8) + rng.uniform(0, 3000))
        d = np.clip(d, 500, 5000).astype(np.uint16)
        d[rng.random((H, W)) < zero_frac] = 0
        depth[i] = d
        labels[i] = ((xx // max(1, W // 4)) + (yy // max(1, H // 3)) * 2 + i) % (classes + 1)
    return rgb, depth, labels


def random_features(rng, F, box_radius=60, region_size=10, use_depth=True):
    feats = np.zeros((11, F), np.int8)
    feats[0] = rng.integers(0, 2, F) if use_depth else 1
    for r in (1, 2, 3, 4):
        feats[r] = rng.integers(-box_radius, box_radius + 1, F)
    for r in (5, 6, 7, 8):
        feats[r] = rng.integers(1, region_size + 1, F)
    col = feats[0] == 1
    feats[9] = np.where(col, rng.integers(0, 3, F), 0)
    feats[10] = np.where(col, rng.integers(0, 3, F), 0)
    return feats


def random_samples(rng, n, n_img, W, H, depth, labels=None, classes=5):
    s = np.zeros(n, orc.SAMPLE_DTYPE)
    s["image"] = rng.integers(0, n_img, n)
    s["x"] = rng.integers(0, W, n)
    s["y"] = rng.integers(0, H, n)
    s["depth_mm"] = depth[s["image"], s["y"], s["x"]]
    if labels is None:
        s["label"] = rng.integers(0, classes, n)
    else:
        s["label"] = labels[s["image"], s["y"], s["x"]]
    return s


def config0_dataset():
    """The plain CPU-runnable workload: 4 frames 160x120, 12 Voronoi sites,
    5 classes + void, 5% missing depth, seeds 4711 + i."""
    import curfil_amd as cf
    return cf.synth_dataset(4, 160, 120, sites=12, classes=5, zero_frac=0.05, seed0=4711)


CONFIG0 = dict(randomSeed=4711, samplesPerImage=1000, featureCount=100, minSampleCount=32, maxDepth=8,
               boxRadius=60, regionSize=10, thresholds=10, ignored_labels=(0,))


def forced_level_case():
    """Images and root samples whose tree is forced whatever features are drawn.

    Image 0 (group B): colour 50 everywhere, depth 1000 mm; image 1 (group D):
    colour 200, depth 1000 mm; image 2 (group A): depth 0 (missing) everywhere.
    Samples sit in the centre, where every box of boxRadius 20 / regionSize 10
    stays inside the 160x120 frame.  A colour feature answers v * (area1 -
    area2) on B and D (the same value for every sample of a group), a depth
    feature 0; a sample on missing depth answers NaN and always goes right
    (random_tree_image.h:485-623).  Labels: B 20 x 1 and 20 x 2, D 20 x 4,
    A 20 x 3.

    The root's best split is B | A + D (normalised information gain 2/3,
    against 0.577 for D | A + B and B + D | A).  At level 2 the B node (40
    samples, impure) cannot be split: every threshold is one of the level's
    responses, so all its samples fall on one side, the side is empty and the
    two ids are rolled back; the A + D node (40 samples) then splits
    perfectly into D (left) and A (right), whose ids follow the rolled-back
    counter: 3 and 4.  (The reference reaches its rollback branch,
    random_tree.h:1366-1394, only after throwing "no samples in left node"
    for such a node; this build keeps the node as a leaf and rolls the ids
    back as that branch intends.  With maxDepth 2 or minSampleCount above 40
    no node has an empty side and the reference's rule applies unchanged.)
    """
    W, H = 160, 120
    rgb = np.zeros((3, H, W, 3), np.uint8)
    rgb[0] = 50
    rgb[1] = 200
    rgb[2] = 120
    depth = np.zeros((3, H, W), np.uint16)
    depth[0] = 1000
    depth[1] = 1000
    rng = np.random.default_rng(5)

    def group(img, labels):
        s = np.zeros(len(labels), orc.SAMPLE_DTYPE)
        s["image"] = img
        s["x"] = rng.integers(40, 120, len(labels))
        s["y"] = rng.integers(35, 85, len(labels))
        s["depth_mm"] = depth[img, s["y"], s["x"]]
        s["label"] = labels
        return s

    roots = np.concatenate([group(0, [1] * 20 + [2] * 20), group(1, [4] * 20), group(2, [3] * 20)])
    return rgb, depth, roots


FORCED_TREE = {
    # index: (id, level, parent, left, right, num_samples, is_leaf, histogram over labels 0..4)
    0: (0, 1, -1, 1, 2, 80, 0, [0, 20, 20, 20, 20]),
    1: (1, 2, 0, -1, -1, 40, 1, [0, 20, 20, 0, 0]),
    2: (2, 2, 0, 3, 4, 40, 0, [0, 0, 0, 20, 20]),
    3: (3, 3, 2, -1, -1, 20, 1, [0, 0, 0, 0, 20]),
    4: (4, 3, 2, -1, -1, 20, 1, [0, 0, 0, 20, 0]),
}


def forced_expected(max_depth, min_samples):
    """The forced tree's nodes for a maxDepth / minSampleCount: the root is level 1 and a level equal
    to maxDepth is not evaluated (random_tree.h:1256-1259); a child grows iff it is impure and holds at
    least minSampleCount samples (random_tree.h:1182-1191); the root is always evaluated."""
    if max_depth <= 1:
        keep = [0]
    elif max_depth == 2 or min_samples > 40:
        keep = [0, 1, 2]
    else:
        keep = [0, 1, 2, 3, 4]
    out = []
    for i in keep:
        node = list(FORCED_TREE[i])
        if node[3] not in keep:
            node[3] = node[4] = -1
            node[6] = 1
        out.append(node)
    return out
