"""ctypes bindings of the CPU oracle (oracle/_build/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker or the timed CPU baseline,
never as the thing measured or shipped.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("ORC_LIB") or os.path.join(HERE, "_build", "liborc.so")

SAMPLE_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("image", "<u4"), ("depth_mm", "<i4"),
                         ("label", "u1"), ("flip", "u1"), ("pad0", "u1"), ("pad1", "u1")])
NODE_DTYPE = np.dtype([("id", "<i4"), ("level", "<i4"), ("parent", "<i4"), ("left", "<i4"), ("right", "<i4"),
                       ("num_samples", "<i4"), ("feature_id", "<i4"), ("threshold", "<f4"), ("score", "<f8"),
                       ("feature", "i1", (11,)), ("is_leaf", "u1"), ("pad", "<i4")])
ARG_ORDER = 7  # g++ evaluates both getNext() of one constructor right to left (pinned by the golden hash)


class _Config(ctypes.Structure):
    _fields_ = [("randomSeed", ctypes.c_int), ("samplesPerImage", ctypes.c_int), ("featureCount", ctypes.c_int),
                ("minSampleCount", ctypes.c_int), ("maxDepth", ctypes.c_int), ("boxRadius", ctypes.c_int),
                ("regionSize", ctypes.c_int), ("thresholds", ctypes.c_int), ("numThreads", ctypes.c_int),
                ("maxImages", ctypes.c_int), ("useDepthImages", ctypes.c_int), ("horizontalFlipping", ctypes.c_int),
                ("subsamplingPixelUniform", ctypes.c_int), ("numIgnored", ctypes.c_int),
                ("ignored", ctypes.c_int * 8)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def _load():
    if not os.path.exists(LIB_PATH):
        build()
    lib = ctypes.CDLL(LIB_PATH)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.orc_mt19937_nth.argtypes = [ctypes.c_uint32, I]
    lib.orc_mt19937_nth.restype = ctypes.c_uint32
    lib.orc_sampler_draws.argtypes = [I, I, I, I, P]
    lib.orc_integral_color.argtypes = [P, I, I, P]
    lib.orc_integral_int.argtypes = [P, I, I, P]
    lib.orc_dataset_create.argtypes = [I, I, I, P, P, P]
    lib.orc_dataset_create.restype = P
    lib.orc_dataset_free.argtypes = [P]
    lib.orc_dataset_integrals.argtypes = [P, I, P, P, P]
    lib.orc_pixel_depth.argtypes = [P, I, I, I]
    lib.orc_responses.argtypes = [P, P, I, P, I, P]
    lib.orc_generate_features.argtypes = [P, P, I, I, I, I, I, I, I, I, I, P, P]
    lib.orc_node_counters.argtypes = [P, P, I, P, P, I, I, I, P]
    lib.orc_scores.argtypes = [P, P, I, I, I, P]
    lib.orc_score.argtypes = [I, P, P, P]
    lib.orc_score.restype = D
    lib.orc_information_gain.argtypes = [I, P, P, P]
    lib.orc_information_gain.restype = D
    lib.orc_argmax.argtypes = [P, I, I, P, P, P]
    lib.orc_normalize_histogram.argtypes = [P, P, I, D, I, P]
    lib.orc_label_prior.argtypes = [P, I, P, I]
    lib.orc_train_forest.argtypes = [P, P, I, I, I, P]
    lib.orc_train_forest.restype = P
    lib.orc_forest_free.argtypes = [P]
    lib.orc_forest_num_classes.argtypes = [P]
    lib.orc_forest_tree_size.argtypes = [P, I]
    lib.orc_forest_tree.argtypes = [P, I, P, P]
    lib.orc_forest_root_count.argtypes = [P, I]
    lib.orc_forest_roots.argtypes = [P, I, P]
    lib.orc_forest_prior.argtypes = [P, P]
    lib.orc_reservoir.argtypes = [I, I, I, I, I, P]
    lib.orc_reservoir.restype = I
    lib.orc_forest_tree_prior.argtypes = [P, I, P]
    lib.orc_train_tree.argtypes = [P, P, I, I, I, P, I, I, P, P, P, I]
    lib.orc_predict.argtypes = [P, I, I, P, P, P, I, P, P]
    lib.orc_feature_generation_hash.argtypes = [I, P, P]
    lib.orc_feature_generation_hash.restype = ctypes.c_uint64
    return lib


lib = _load()


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def config(randomSeed, samplesPerImage, featureCount, minSampleCount, maxDepth, boxRadius, regionSize, thresholds,
           numThreads=1, maxImages=0, useDepthImages=True, horizontalFlipping=False, ignored_labels=(),
           subsamplingType="classUniform"):
    c = _Config()
    c.subsamplingPixelUniform = int(subsamplingType == "pixelUniform")
    c.randomSeed, c.samplesPerImage, c.featureCount = randomSeed, samplesPerImage, featureCount
    c.minSampleCount, c.maxDepth, c.boxRadius = minSampleCount, maxDepth, boxRadius
    c.regionSize, c.thresholds, c.numThreads, c.maxImages = regionSize, thresholds, numThreads, maxImages
    c.useDepthImages, c.horizontalFlipping = int(useDepthImages), int(horizontalFlipping)
    c.numIgnored = len(ignored_labels)
    for i, l in enumerate(ignored_labels):
        c.ignored[i] = l
    return c


class Dataset:
    """Images with the reference's host integrals (double Kahan, image.cpp:236-277)."""

    def __init__(self, color, depth, labels=None):
        color = np.ascontiguousarray(color, np.float32)     # [N][3][H][W]
        depth = np.ascontiguousarray(depth, np.int32)       # [N][H][W]
        self.N, _, self.H, self.W = color.shape
        lab = None if labels is None else np.ascontiguousarray(labels, np.uint8)
        self.h = lib.orc_dataset_create(self.N, self.W, self.H, _p(color), _p(depth), _p(lab))

    @classmethod
    def from_u8(cls, rgb, depth, labels=None):
        color = np.ascontiguousarray(np.transpose(rgb, (0, 3, 1, 2)).astype(np.float32))
        return cls(color, depth.astype(np.int32), labels)

    def integrals(self, img):
        col = np.zeros((3, self.H, self.W), np.float32)
        dep = np.zeros((self.H, self.W), np.int32)
        val = np.zeros((self.H, self.W), np.int32)
        lib.orc_dataset_integrals(self.h, img, _p(col), _p(dep), _p(val))
        return col, dep, val

    def pixel_depth(self, img, x, y):
        return lib.orc_pixel_depth(self.h, img, x, y)

    def responses(self, samples, features):
        samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
        features = np.ascontiguousarray(features, np.int8)
        F = features.shape[1]
        out = np.zeros((F, len(samples)), np.float64)
        lib.orc_responses(self.h, _p(samples), len(samples), _p(features), F, _p(out))
        return out

    def generate_features(self, samples, seed, F, T, boxRadius, regionSize, useDepthImages=True, sort=True,
                          arg_order=ARG_ORDER):
        samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
        feats = np.zeros((11, F), np.int8)
        thr = np.zeros((T, F), np.float32)
        st = lib.orc_generate_features(self.h, _p(samples), len(samples), seed, F, T, boxRadius, regionSize,
                                       int(useDepthImages), arg_order, int(sort), _p(feats), _p(thr))
        if st != 0:
            raise RuntimeError("failed to generate random features. max loops exceeded")
        return feats, thr

    def node_counters(self, samples, features, thresholds, C):
        """counters in the CPU layout [C][F][T][2] (random_tree_image.cpp:70-71)."""
        samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
        features = np.ascontiguousarray(features, np.int8)
        thresholds = np.ascontiguousarray(thresholds, np.float32)
        T, F = thresholds.shape
        out = np.zeros((C, F, T, 2), np.uint32)
        lib.orc_node_counters(self.h, _p(samples), len(samples), _p(features), _p(thresholds), F, T, C, _p(out))
        return out

    def label_prior(self, num_labels):
        out = np.zeros(256, np.uint32)
        n = lib.orc_label_prior(self.h, num_labels, _p(out), 256)
        return out[:n]

    def train_forest(self, cfg, num_trees, num_labels):
        st = ctypes.c_int()
        f = lib.orc_train_forest(self.h, ctypes.byref(cfg), num_trees, num_labels, ARG_ORDER, ctypes.byref(st))
        forest = Forest(f)
        if st.value != 0:
            raise RuntimeError(f"oracle training failed with status {st.value}")
        return forest

    def train_tree(self, cfg, tree_id, C, rs_seed, roots, cap=1 << 20):
        roots = np.ascontiguousarray(roots, SAMPLE_DTYPE)
        nodes = np.zeros(cap, NODE_DTYPE)
        hist = np.zeros((cap, C), np.uint32)
        n = ctypes.c_int()
        st = lib.orc_train_tree(self.h, ctypes.byref(cfg), tree_id, C, rs_seed, _p(roots), len(roots), ARG_ORDER,
                                ctypes.byref(n), _p(nodes), _p(hist), cap)
        if st != 0:
            raise RuntimeError(f"oracle tree training failed with status {st}")
        return nodes[:n.value].copy(), hist[:n.value].copy()

    def predict(self, img, trees, posteriors, C):
        """trees: list of node arrays; posteriors: list of [nodes][C] float64."""
        counts = np.array([len(t) for t in trees], np.int32)
        nodes = np.ascontiguousarray(np.concatenate(trees), NODE_DTYPE)
        post = np.ascontiguousarray(np.concatenate(posteriors), np.float64)
        probs = np.zeros((C, self.H, self.W), np.float32)
        labels = np.zeros((self.H, self.W), np.uint8)
        lib.orc_predict(self.h, img, len(trees), _p(counts), _p(nodes), _p(post), C, _p(probs), _p(labels))
        return labels, probs

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_dataset_free(self.h)
            self.h = None


class Forest:
    def __init__(self, h):
        self.h = h

    @property
    def num_classes(self):
        return lib.orc_forest_num_classes(self.h)

    def tree(self, t):
        n = lib.orc_forest_tree_size(self.h, t)
        nodes = np.zeros(n, NODE_DTYPE)
        hist = np.zeros((n, self.num_classes), np.uint32)
        lib.orc_forest_tree(self.h, t, _p(nodes), _p(hist))
        return nodes, hist

    def roots(self, t):
        n = lib.orc_forest_root_count(self.h, t)
        out = np.zeros(n, SAMPLE_DTYPE)
        lib.orc_forest_roots(self.h, t, _p(out))
        return out

    def prior(self, t=None):
        """the label prior over all images, or over tree t's images (maxImages)"""
        out = np.zeros(self.num_classes, np.uint32)
        if t is None:
            lib.orc_forest_prior(self.h, _p(out))
        else:
            lib.orc_forest_tree_prior(self.h, t, _p(out))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib.orc_forest_free(self.h)
            self.h = None


def integral_color(channel):
    channel = np.ascontiguousarray(channel, np.float32)
    H, W = channel.shape
    out = np.zeros_like(channel)
    lib.orc_integral_color(_p(channel), W, H, _p(out))
    return out


def integral_int(channel):
    channel = np.ascontiguousarray(channel, np.int32)
    H, W = channel.shape
    out = np.zeros_like(channel)
    lib.orc_integral_int(_p(channel), W, H, _p(out))
    return out


def score(left, right, all_):
    left, right, all_ = (np.ascontiguousarray(a, np.uint32) for a in (left, right, all_))
    return lib.orc_score(len(left), _p(left), _p(right), _p(all_))


def information_gain(left, right, all_):
    left, right, all_ = (np.ascontiguousarray(a, np.uint32) for a in (left, right, all_))
    return lib.orc_information_gain(len(left), _p(left), _p(right), _p(all_))


def scores_from_counters(counters_cft2, hist):
    """counters [C][F][T][2] -> scores [T][F]"""
    counters_cft2 = np.ascontiguousarray(counters_cft2, np.uint32)
    hist = np.ascontiguousarray(hist, np.uint32)
    C, F, T, _ = counters_cft2.shape
    out = np.zeros((T, F), np.float64)
    lib.orc_scores(_p(counters_cft2), _p(hist), C, F, T, _p(out))
    return out


def argmax(scores):
    scores = np.ascontiguousarray(scores, np.float64)
    T, F = scores.shape
    bt, bf, bs = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    lib.orc_argmax(_p(scores), T, F, ctypes.byref(bt), ctypes.byref(bf), ctypes.byref(bs))
    return bf.value, bt.value, bs.value


def normalize_histogram(hist, prior, bias, use_prior):
    hist = np.ascontiguousarray(hist, np.uint32)
    prior = np.ascontiguousarray(prior, np.uint32)
    out = np.zeros(len(hist), np.float64)
    lib.orc_normalize_histogram(_p(hist), _p(prior), len(hist), bias, int(use_prior), _p(out))
    return out


def reservoir(seed, lo, hi, n, k):
    """ReservoirSampler of 0..n-1 keeping k, drawing from RandomSource(seed).uniformSampler(lo, hi)"""
    out = np.zeros(k, np.int32)
    got = lib.orc_reservoir(seed, lo, hi, n, k, _p(out))
    return out[:got]


def sampler_draws(seed, lo, hi, n):
    out = np.zeros(n, np.int32)
    lib.orc_sampler_draws(seed, lo, hi, n, _p(out))
    return out


def mt19937_nth(seed, n):
    return lib.orc_mt19937_nth(seed, n)


def feature_generation_hash(arg_order=ARG_ORDER, want_features=False):
    feats = np.zeros((11, 500), np.int8)
    thr = np.zeros((50, 500), np.float32)
    h = lib.orc_feature_generation_hash(arg_order, _p(feats), _p(thr))
    return (h, feats, thr) if want_features else h
