"""Tree files (host side, no GPU): the per-tree JSON layout RandomTreeExport writes and RandomTreeImport reads
(export.cpp:180-313, import.cpp:42-265), plain or gzip.

Checks, after the reference's import_export_test.cpp (export, import, then compare configuration, prior and
every node's id, level, histogram, featureId, threshold and score):
  * a hand-written file in the reference's layout (every value a string, `\\/` escapes, the verbose writer's
    extra `timers`/`countsPerImage` keys, an empty array written as "") reads into the expected node table,
    configuration, label colours and ignored labels;
  * export -> import of random complete trees gives back identical node tables, histograms and priors, bit
    for bit (thresholds and scores included), and exporting the imported forest writes the same document;
  * each imported tree normalises with its own label prior (RandomForestImage(treeFiles) calls every
    tree's normalizeHistograms, random_forest_image.cpp:366-376), checked against the oracle's
    detail::normalizeHistogram restatement;
  * the reference's rejections: histogram sum below the sample count, unknown feature type, offsets that
    are not pairs, trees of different configurations or class counts, missing keys, missing files.
"""
import copy
import gzip
import json
import os
import shutil

import numpy as np
import pytest

import curfil_amd as cf
import oracle as orc

HERE = os.path.dirname(os.path.abspath(__file__))
FIXTURE = os.path.join(HERE, "golden", "tree_reference_layout.json")

# feature column order of cf_tree_node.feature: type, offset1 x/y, offset2 x/y, region1 x/y, region2 x/y, c1, c2
O1X, O1Y, O2X, O2Y, R1X, R1Y, R2X, R2Y, C1, C2 = range(1, 11)


def _same_bits(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))


def random_complete_tree(rng, depth, C, tree_id=0):
    """A complete tree of the given depth in the library's node order (root, then each level's children
    pairs), random splits and leaf counts; inner histograms are the sums of their leaves."""
    n = (1 << (depth + 1)) - 1
    nodes = np.zeros(n, cf.NODE_DTYPE)
    hist = np.zeros((n, C), np.uint32)
    for i in range(n):
        lvl = int(np.floor(np.log2(i + 1)))
        nd = nodes[i]
        nd["id"], nd["level"] = i, lvl
        nd["parent"] = -1 if i == 0 else (i - 1) // 2
        if lvl < depth:
            nd["left"], nd["right"] = 2 * i + 1, 2 * i + 2
            nd["is_leaf"] = 0
            nd["feature_id"] = int(rng.integers(0, 2000))
            nd["threshold"] = np.float32(rng.normal() * 100)
            nd["score"] = float(rng.random())
            f = np.zeros(11, np.int8)
            f[0] = rng.integers(0, 2)
            f[O1X:R2Y + 1] = rng.integers(-120, 121, 8)
            f[R1X:R2Y + 1] = rng.integers(1, 21, 4)
            if f[0] == cf.COLOR:
                f[C1], f[C2] = rng.integers(0, 3, 2)
            nd["feature"] = f
        else:
            nd["left"] = nd["right"] = -1
            nd["is_leaf"] = 1
            hist[i] = rng.integers(0, 1 << 20, C)
    for i in range(n - 1, -1, -1):
        if nodes[i]["is_leaf"] == 0:
            hist[i] = hist[2 * i + 1] + hist[2 * i + 2]
        nodes[i]["num_samples"] = min(int(hist[i].sum()), 2**31 - 1)
    return nodes, hist


def _config(**kw):
    p = dict(randomSeed=4711, samplesPerImage=500, featureCount=500, minSampleCount=32, maxDepth=10,
             boxRadius=50, regionSize=10, thresholds=10, numThreads=4, maxImages=10)
    p.update(kw)
    return cf.TrainingConfiguration(**p)


def _read(path):
    op = gzip.open if str(path).endswith(".gz") else open
    with op(path, "rt") as f:
        return json.load(f)


def _write(path, doc):
    op = gzip.open if str(path).endswith(".gz") else open
    with op(path, "wt") as f:
        json.dump(doc, f, indent=4)


def _tables(forest):
    return [forest.tree(t) for t in range(forest.num_trees)]


def test_reference_layout_file_reads():
    f = cf.RandomForestImage.from_files([FIXTURE])
    assert f.num_trees == 1 and f.num_classes == 3
    assert f.palette.tolist() == [[50, 205, 50], [128, 0, 0], [0, 0, 128]]
    c = f.configuration
    assert (c.randomSeed, c.samplesPerImage, c.featureCount, c.minSampleCount, c.maxDepth) == (4711, 500, 500, 32, 10)
    assert (c.boxRadius, c.regionSize, c.thresholds, c.numThreads, c.maxImages) == (50, 10, 10, 4, 10)
    assert c.subsamplingType == "classUniform" and c.useDepthImages
    assert c.ignored_labels == [0]      # "50,205,50" is the colour of label 0
    nodes, hist = f.tree(0)
    # breadth-first: root, its children (ids 1, 2), then node 2's children (ids 3, 4)
    assert nodes["id"].tolist() == [0, 1, 2, 3, 4]
    assert nodes["level"].tolist() == [0, 1, 1, 2, 2]
    assert nodes["parent"].tolist() == [-1, 0, 0, 2, 2]
    assert nodes["is_leaf"].tolist() == [0, 1, 0, 1, 1]
    assert nodes["left"][[0, 2]].tolist() == [1, 3] and nodes["right"][[0, 2]].tolist() == [2, 4]
    assert nodes["num_samples"].tolist() == [500, 200, 300, 250, 50]
    assert nodes["feature_id"][[0, 2]].tolist() == [217, 9]
    assert _same_bits(nodes["threshold"][[0, 2]], np.array([-12.75, 0.1], np.float32))
    assert nodes["score"][0] == 0.31415926535897931 and nodes["score"][2] == 0.125
    depth_f = nodes["feature"][0]
    assert depth_f[0] == cf.DEPTH
    assert (depth_f[O1X], depth_f[O1Y], depth_f[R1X], depth_f[R1Y]) == (-3, 14, 5, 2)
    assert (depth_f[O2X], depth_f[O2Y], depth_f[R2X], depth_f[R2Y]) == (7, -20, 1, 9)
    col_f = nodes["feature"][2]
    assert col_f[0] == cf.COLOR and (col_f[C1], col_f[C2]) == (2, 0)
    assert (col_f[O1X], col_f[O1Y], col_f[R1X], col_f[R1Y]) == (0, -1, 4, 4)
    assert (col_f[O2X], col_f[O2Y], col_f[R2X], col_f[R2Y]) == (-60, 33, 10, 3)
    assert hist.tolist() == [[120, 300, 80], [100, 90, 10], [20, 210, 70], [0, 200, 50], [20, 10, 20]]
    prior = np.array([120, 300, 80], np.uint32)
    assert f.prior().tolist() == prior.tolist()
    post = np.array([orc.normalize_histogram(h, prior, 0.0, True) for h in hist])
    assert np.array_equal(f.posteriors(0), post)
    f.close()


def test_gzip_and_plain_files_read_the_same(tmp_path):
    gz = tmp_path / "tree0.json.gz"
    with open(FIXTURE, "rb") as src, gzip.open(gz, "wb") as dst:
        shutil.copyfileobj(src, dst)
    a = cf.RandomForestImage.from_files([FIXTURE])
    b = cf.RandomForestImage.from_files([gz])
    (an, ah), (bn, bh) = a.tree(0), b.tree(0)
    assert _same_bits(an, bn) and np.array_equal(ah, bh)


@pytest.mark.parametrize("depth,C,trees", [(0, 2, 1), (6, 5, 3), (9, 32, 2)])
def test_export_import_round_trip_is_exact(tmp_path, depth, C, trees):
    rng = np.random.default_rng(depth * 100 + C)
    tabs = [random_complete_tree(rng, depth, C) for _ in range(trees)]
    prior = rng.integers(1, 1 << 24, C).astype(np.uint32)
    forest = cf.RandomForestImage.from_trees(None, tabs, C, prior=prior)
    cfg = _config(ignored_labels=[C - 1], subsamplingType="pixelUniform", useDepthImages=False)
    palette = rng.integers(0, 256, (C, 3)).astype(np.uint8)
    palette[:, 0] = np.arange(C)      # distinct colours
    paths = forest.export(tmp_path / "out", cfg, palette=palette, folder_training="/data/train")
    assert [os.path.basename(p) for p in paths] == [f"tree{t}.json.gz" for t in range(trees)]
    doc = _read(paths[0])
    assert doc["folderTraining"] == "/data/train" and doc["subsamplingType"] == "pixelUniform"

    def leaves(v):
        if isinstance(v, dict):
            for x in v.values():
                yield from leaves(x)
        elif isinstance(v, list):
            for x in v:
                yield from leaves(x)
        else:
            yield v
    assert all(isinstance(v, str) for v in leaves(doc))     # Boost.PropertyTree writes strings only
    back = cf.RandomForestImage.from_files(paths)
    assert back.num_trees == trees and back.num_classes == C
    assert back.palette.tolist() == palette.tolist()
    assert back.prior().tolist() == prior.tolist()
    c = back.configuration
    assert (c.randomSeed, c.samplesPerImage, c.featureCount, c.minSampleCount, c.maxDepth, c.boxRadius,
            c.regionSize, c.thresholds, c.numThreads, c.maxImages) == (4711, 500, 500, 32, 10, 50, 10, 10, 4, 10)
    assert c.ignored_labels == [C - 1] and c.subsamplingType == "pixelUniform" and not c.useDepthImages
    for (n0, h0), (n1, h1) in zip(tabs, _tables(back)):
        assert _same_bits(n0, n1)
        assert np.array_equal(h0, h1)
    # the imported forest writes the same document (the date and host aside)
    again = back.export(tmp_path / "again", back.configuration, palette=back.palette, folder_training="/data/train")
    for p, q in zip(paths, again):
        d0, d1 = _read(p), _read(q)
        for k in ("date", "hostname"):
            d0.pop(k), d1.pop(k)
        assert d0 == d1


def test_plain_json_export_without_palette(tmp_path):
    rng = np.random.default_rng(5)
    forest = cf.RandomForestImage.from_trees(None, [random_complete_tree(rng, 3, 4)], 4)
    p = tmp_path / "tree.json"
    forest.export_tree(0, p, _config())
    with open(p, "rb") as fh:
        assert fh.read(1) == b"{"          # not compressed
    doc = _read(p)
    assert list(doc["tree"]["histogram"]) == ["0,0,0 (0)", "1,1,1 (1)", "2,2,2 (2)", "3,3,3 (3)"]
    back = cf.RandomForestImage.from_files([p])
    assert _same_bits(back.tree(0)[0], forest.tree(0)[0])


def test_each_tree_normalises_with_its_own_prior(tmp_path):
    doc = _read(FIXTURE)
    other = copy.deepcopy(doc)
    other["randomSeed"] = "17"                     # trees of different seeds make one forest
    other["classLabelPriorDistribution"] = {"0": "10", "1": "10", "2": "480"}
    _write(tmp_path / "tree0.json", doc)
    _write(tmp_path / "tree1.json.gz", other)
    f = cf.RandomForestImage.from_files([tmp_path / "tree0.json", tmp_path / "tree1.json.gz"],
                                        histogramBias=0.05)
    for t, pr in enumerate(([120, 300, 80], [10, 10, 480])):
        _, hist = f.tree(t)
        prior = np.array(pr, np.uint32)
        post = np.array([orc.normalize_histogram(h, prior, 0.05, True) for h in hist])
        assert np.array_equal(f.posteriors(t), post), t
    assert not np.array_equal(f.posteriors(0), f.posteriors(1))
    f.normalizeHistograms(0.0, False)
    _, hist = f.tree(1)
    ones = np.ones(3, np.uint32)
    assert np.array_equal(f.posteriors(1), np.array([orc.normalize_histogram(h, ones, 0.0, False) for h in hist]))


def _broken(tmp_path, name, edit):
    doc = _read(FIXTURE)
    edit(doc)
    p = tmp_path / name
    _write(p, doc)
    return p


@pytest.mark.parametrize("edit,message", [
    (lambda d: d["tree"]["left"].update(samples="201"), "incorrect histogram sum"),
    (lambda d: d["tree"]["split"]["feature"].update(type="gradient"), "illegal feature type"),
    (lambda d: d["tree"]["split"]["feature"].update(offset1=["1", "2", "3"]), "offset1"),
    (lambda d: d["tree"]["right"]["split"]["feature"].update(channel1="3"), "channel"),
    (lambda d: d.pop("hostname"), "hostname"),
    (lambda d: d["tree"]["right"].pop("split"), "split"),
    (lambda d: d["tree"]["left"]["histogram"].pop("0,0,128 (2)"), "different numbers of classes"),
    (lambda d: d.update(maxDepth="ten"), "not an integer"),
])
def test_malformed_files_are_rejected(tmp_path, edit, message):
    p = _broken(tmp_path, "bad.json", edit)
    with pytest.raises(cf.CurfilError, match=message) as e:
        cf.RandomForestImage.from_files([p])
    assert e.value.code == cf.CF_ERR_ARG


def test_forests_of_mixed_trees_are_rejected(tmp_path):
    deeper = _broken(tmp_path, "deeper.json", lambda d: d.update(maxDepth="12"))
    with pytest.raises(cf.CurfilError, match="different configurations"):
        cf.RandomForestImage.from_files([FIXTURE, deeper])

    def two_classes(d):
        d["classLabelPriorDistribution"] = {"0": "1", "1": "1"}
        for node in (d["tree"], d["tree"]["left"], d["tree"]["right"], d["tree"]["right"]["left"],
                     d["tree"]["right"]["right"]):
            node["histogram"].pop("0,0,128 (2)")
            node["samples"] = "0"
    fewer = _broken(tmp_path, "fewer.json", two_classes)
    with pytest.raises(cf.CurfilError, match="different number of classes"):
        cf.RandomForestImage.from_files([FIXTURE, fewer])
    with pytest.raises(cf.CurfilError, match="cannot open"):
        cf.RandomForestImage.from_files([tmp_path / "missing.json.gz"])
    with pytest.raises(cf.CurfilError, match="JSON"):
        p = tmp_path / "cut.json"
        p.write_text(open(FIXTURE).read()[:500])
        cf.RandomForestImage.from_files([p])


def test_deeply_nested_file_is_rejected_not_crashing(tmp_path):
    p = tmp_path / "deep.json"
    p.write_text('{"tree": ' + "[" * 100000 + "]" * 100000 + "}")
    with pytest.raises(cf.CurfilError, match="nested too deeply"):
        cf.RandomForestImage.from_files([p])
