"""curfil_amd -- MI355X-native hot path of CURFIL (random-forest training and
dense prediction on RGB-D images), exposed from libcurfil_hip.so through its
C ABI (include/curfil_hip.h).

The Python layer mirrors the reference's library surface for this path:
TrainingConfiguration (random_tree.h:187-493), RandomForestImage train /
predict / normalizeHistograms (random_forest_image.h:44-168), and the plugin
seams ImageFeatureEvaluation::generateRandomFeatures / calculateScores and the
counters of calculateFeatureResponsesAndHistograms.  All compute runs in the
HIP library; importing this package fails if the library is not built.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("CURFIL_LIB") or os.path.join(_HERE, "libcurfil_hip.so")

CF_OK = 0
CF_ERR_ARG = -1
CF_ERR_HIP = -2
CF_ERR_RUNTIME = -3
CF_ERR_UNSUPPORTED = -4
CF_ERR_NOMEM = -5

MODE_PLAIN = 0
MODE_BATCHED = 1

SAMPLE_DTYPE = np.dtype([("x", "<u2"), ("y", "<u2"), ("image", "<u4"), ("depth_mm", "<i4"),
                         ("label", "u1"), ("flip", "u1"), ("pad0", "u1"), ("pad1", "u1")])
NODE_DTYPE = np.dtype([("id", "<i4"), ("level", "<i4"), ("parent", "<i4"), ("left", "<i4"), ("right", "<i4"),
                       ("num_samples", "<i4"), ("feature_id", "<i4"), ("threshold", "<f4"), ("score", "<f8"),
                       ("feature", "i1", (11,)), ("is_leaf", "u1"), ("pad", "<i4")])
assert SAMPLE_DTYPE.itemsize == 16 and NODE_DTYPE.itemsize == 56

# feature parameter rows of the [11][F] SoA (random_tree_image.h:702-773)
FEATURE_ROWS = ("type", "offset1X", "offset1Y", "offset2X", "offset2Y", "region1X", "region1Y",
                "region2X", "region2Y", "channel1", "channel2")
DEPTH, COLOR = 0, 1


class CurfilError(RuntimeError):
    def __init__(self, code, message):
        super().__init__(f"[{code}] {message}")
        self.code = code


class _TrainConfig(ctypes.Structure):
    _fields_ = [("random_seed", ctypes.c_int32), ("samples_per_image", ctypes.c_uint32),
                ("feature_count", ctypes.c_uint32), ("min_sample_count", ctypes.c_uint32),
                ("max_depth", ctypes.c_int32), ("box_radius", ctypes.c_uint16), ("region_size", ctypes.c_uint16),
                ("thresholds", ctypes.c_uint16), ("pad0", ctypes.c_uint16), ("num_threads", ctypes.c_int32),
                ("max_images", ctypes.c_int32), ("use_depth_images", ctypes.c_int32),
                ("horizontal_flipping", ctypes.c_int32), ("subsampling_pixel_uniform", ctypes.c_int32),
                ("num_ignored", ctypes.c_int32), ("ignored_labels", ctypes.c_int32 * 8),
                ("use_depth_filling", ctypes.c_int32), ("use_cielab", ctypes.c_int32)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m curfil_amd.build` (or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    P, I, D, U32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_uint32
    sig = {
        "cf_ctx_create": [I, ctypes.POINTER(P)], "cf_ctx_destroy": [P], "cf_ctx_synchronize": [P],
        "cf_images_create": [P, I, I, I, ctypes.POINTER(P)], "cf_images_destroy": [P],
        "cf_images_set_rgbd_u8": [P, I, P, P], "cf_images_set_rgbd_f32": [P, I, P, P],
        "cf_images_set_labels": [P, I, P], "cf_images_get_integrals": [P, I, P, P, P],
        "cf_generate_features": [P, P, P, I, I, I, I, I, I, I, P, P],
        "cf_feature_responses": [P, P, P, I, P, I, P],
        "cf_feature_responses_paired": [P, P, P, I, P, P],
        "cf_node_counters": [P, P, P, I, P, P, I, I, I, P],
        "cf_scores": [P, P, P, I, I, I, P],
        "cf_argmax_scores": [P, P, I, I, P, P, P],
        "cf_node_best_split": [P, P, P, I, P, P, I, I, I, P, P, P, P],
        "cf_forest_train": [P, P, P, I, I, I, ctypes.POINTER(P)],
        "cf_forest_sample": [P, P, P, I, I, ctypes.POINTER(P)], "cf_forest_grow": [P, P, P, P, I],
        "cf_ctx_set_profiling": [P, I], "cf_forest_profile": [P, P, I],
        "cf_tree_train_from_samples": [P, P, P, I, I, I, P, I, I, ctypes.POINTER(P)],
        "cf_forest_create": [P, I, P, P, P, I, P, ctypes.POINTER(P)],
        "cf_forest_destroy": [P], "cf_forest_num_trees": [P], "cf_forest_num_classes": [P],
        "cf_forest_tree_size": [P, I], "cf_forest_get_tree": [P, I, P, P],
        "cf_forest_get_root_samples": [P, I, P, I], "cf_forest_get_prior": [P, P], "cf_forest_get_tree_prior": [P, I, P],
        "cf_forest_train_stats": [P, P, P], "cf_forest_normalize": [P, D, I],
        "cf_forest_get_posteriors": [P, I, P],
        "cf_forest_export_tree": [P, I, ctypes.c_char_p, P, P, ctypes.c_char_p],
        "cf_forest_import": [P, ctypes.POINTER(ctypes.c_char_p), I, P, P, ctypes.POINTER(P)],
        "cf_predict": [P, P, P, I, P, P], "cf_predict_batch": [P, P, P, P, I, P, P],
        "cf_predict_batch_device": [P, P, P, P, I, P, P],
        "cf_predict_batch_device_noalloc": [P, P, P, P, I, P, P],
        "cf_device_alloc": [P, ctypes.c_size_t, ctypes.POINTER(P)], "cf_device_free": [P, P],
        "cf_host_alloc": [P, ctypes.c_size_t, ctypes.POINTER(P)], "cf_host_free": [P, P],
        "cf_device_to_host": [P, P, P, ctypes.c_size_t],
        "cf_synth_rgbd": [U32, U32, I, I, I, I, D, P, P, P],
        "cf_images_set_rgbd_u8_padded": [P, I, P, P, I, I],
        "cf_label_map_create": [ctypes.POINTER(P)], "cf_label_map_destroy": [P],
        "cf_label_map_encode": [P, P, ctypes.POINTER(I)], "cf_label_map_find": [P, P],
        "cf_label_map_size": [P], "cf_label_map_colors": [P, P],
        "cf_dataset_load": [P, I, I, I, I, P, P, I, ctypes.POINTER(P)], "cf_dataset_destroy": [P],
        "cf_dataset_get_color": [P, I, P], "cf_images_set_rgbd_f32_padded": [P, I, P, P, I, I],
        "cf_pixel_accuracy": [P, P, ctypes.c_size_t, I, P, I, P, I, P], "cf_confusion_normalize": [P, I],
        "cf_confusion_class_accuracy": [P, I, I, I, P, I, P],
        "cf_dataset_info": [P, P, P, P, P], "cf_dataset_image_size": [P, I, P, P],
        "cf_dataset_get": [P, I, P, P, P], "cf_dataset_to_images": [P, P, ctypes.POINTER(P)],
        "cf_png_read_info": [ctypes.c_char_p, P, P, P, P], "cf_png_read": [ctypes.c_char_p, P, ctypes.c_size_t],
        "cf_png_write": [ctypes.c_char_p, I, I, I, I, P],
        "cf_tool_train": [I, ctypes.POINTER(ctypes.c_char_p)], "cf_tool_predict": [I, ctypes.POINTER(ctypes.c_char_p)],
        "cf_server_create": [P, P, D, I, ctypes.POINTER(P)],
        "cf_server_submit": [P, P, P, I, I, ctypes.POINTER(ctypes.c_uint64)],
        "cf_server_wait": [P, ctypes.c_uint64, P, P, P],
        "cf_server_stats": [P, P, I], "cf_server_destroy": [P],
    }
    for name, args in sig.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = ctypes.c_int
    lib.cf_eval_shape_info.argtypes = [I, P, P, P, P]
    lib.cf_eval_shape_info.restype = I
    lib.cf_eval_tuning.argtypes = [P, I]
    lib.cf_eval_tuning.restype = I
    lib.cf_eval_tuning_reset.argtypes = []
    lib.cf_eval_tuning_reset.restype = I
    lib.cf_last_error.argtypes = [P]
    lib.cf_last_error.restype = ctypes.c_char_p
    lib.cf_server_error.argtypes = [P]
    lib.cf_server_error.restype = ctypes.c_char_p
    return lib


lib = _load()


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(ctx, st):
    """Raise on a failed call; calls made without a context report through cf_last_error(NULL)."""
    if ctx is not None:
        ctx.check(st)
    elif st != CF_OK:
        raise CurfilError(st, lib.cf_last_error(None).decode())


class Context:
    """One HIP device + stream (cf_ctx).  Handles created on it (images,
    forests) are released before the context itself."""

    def __init__(self, device=0):
        import weakref
        h = ctypes.c_void_p()
        st = lib.cf_ctx_create(device, ctypes.byref(h))
        if st != CF_OK:
            raise CurfilError(st, f"cf_ctx_create(device={device}) failed")
        self.h = h
        self._children = weakref.WeakSet()

    def _adopt(self, obj):
        self._children.add(obj)

    def check(self, st):
        if st != CF_OK:
            raise CurfilError(st, lib.cf_last_error(self.h).decode())

    def synchronize(self):
        self.check(lib.cf_ctx_synchronize(self.h))

    def set_profiling(self, enable=True):
        self.check(lib.cf_ctx_set_profiling(self.h, int(enable)))

    def close(self):
        if self.h:
            # servers first: they use forests and images of this context
            children = sorted(self._children, key=lambda c: 0 if isinstance(c, Server) else 1)
            for child in children:
                child.close()
            lib.cf_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Images:
    """A set of equally sized RGB-D images with device-resident integrals."""

    def __init__(self, ctx, count, width, height, _handle=None):
        self.ctx, self.count, self.width, self.height = ctx, count, width, height
        h = _handle
        if h is None:
            h = ctypes.c_void_p()
            ctx.check(lib.cf_images_create(ctx.h, count, width, height, ctypes.byref(h)))
        self.h = h
        ctx._adopt(self)

    def set_rgbd_u8_padded(self, index, rgb, depth):
        """an image smaller than the set, padded as loadImages pads integrated images"""
        rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
        depth = np.ascontiguousarray(depth, dtype=np.uint16)
        h, w = depth.shape
        assert rgb.shape == (h, w, 3)
        self.ctx.check(lib.cf_images_set_rgbd_u8_padded(self.h, index, _ptr(rgb), _ptr(depth), w, h))

    def set_rgbd_u8(self, index, rgb, depth):
        rgb = np.ascontiguousarray(rgb, dtype=np.uint8)
        depth = np.ascontiguousarray(depth, dtype=np.uint16)
        assert rgb.shape == (self.height, self.width, 3) and depth.shape == (self.height, self.width)
        self.ctx.check(lib.cf_images_set_rgbd_u8(self.h, index, _ptr(rgb), _ptr(depth)))

    def set_rgbd_f32(self, index, color, depth):
        color = np.ascontiguousarray(color, dtype=np.float32)
        depth = np.ascontiguousarray(depth, dtype=np.int32)
        assert color.shape == (3, self.height, self.width) and depth.shape == (self.height, self.width)
        self.ctx.check(lib.cf_images_set_rgbd_f32(self.h, index, _ptr(color), _ptr(depth)))

    def set_labels(self, index, labels):
        labels = np.ascontiguousarray(labels, dtype=np.uint8)
        assert labels.shape == (self.height, self.width)
        self.ctx.check(lib.cf_images_set_labels(self.h, index, _ptr(labels)))

    def integrals(self, index):
        col = np.empty((3, self.height, self.width), np.float32)
        dep = np.empty((self.height, self.width), np.int32)
        val = np.empty((self.height, self.width), np.int32)
        self.ctx.check(lib.cf_images_get_integrals(self.h, index, _ptr(col), _ptr(dep), _ptr(val)))
        return col, dep, val

    def close(self):
        if self.h:
            lib.cf_images_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def make_samples(x, y, image, depth_mm, label, flip=0):
    n = len(x)
    s = np.zeros(n, SAMPLE_DTYPE)
    s["x"], s["y"], s["image"], s["depth_mm"], s["label"], s["flip"] = x, y, image, depth_mm, label, flip
    return s


def generate_features(ctx, images, samples, seed, num_features, num_thresholds, box_radius, region_size,
                      use_depth_images=True):
    """ImageFeatureEvaluation::generateRandomFeatures (CPU generator semantics), on the device.
    Returns (features int8 [11][F], thresholds float32 [T][F]), sorted by feature key."""
    samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
    feats = np.zeros((11, num_features), np.int8)
    thr = np.zeros((num_thresholds, num_features), np.float32)
    ctx.check(lib.cf_generate_features(ctx.h, images.h, _ptr(samples), len(samples), seed, num_features,
                                       num_thresholds, box_radius, region_size, int(use_depth_images),
                                       _ptr(feats), _ptr(thr)))
    return feats, thr


def feature_responses(ctx, images, samples, features):
    samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
    features = np.ascontiguousarray(features, np.int8)
    F = features.shape[1]
    out = np.zeros((F, len(samples)), np.float64)
    ctx.check(lib.cf_feature_responses(ctx.h, images.h, _ptr(samples), len(samples), _ptr(features), F, _ptr(out)))
    return out


def node_counters(ctx, images, samples, features, thresholds, num_classes):
    """counters [F][T][C][2] (left, right) of one node."""
    samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
    features = np.ascontiguousarray(features, np.int8)
    thresholds = np.ascontiguousarray(thresholds, np.float32)
    T, F = thresholds.shape
    out = np.zeros((F, T, num_classes, 2), np.uint32)
    ctx.check(lib.cf_node_counters(ctx.h, images.h, _ptr(samples), len(samples), _ptr(features), _ptr(thresholds),
                                   F, T, num_classes, _ptr(out)))
    return out


def scores(ctx, counters, histogram):
    """NormalizedInformationGainScore of every (threshold, feature): [T][F] float64."""
    counters = np.ascontiguousarray(counters, np.uint32)
    histogram = np.ascontiguousarray(histogram, np.uint32)
    F, T, C, _ = counters.shape
    out = np.zeros((T, F), np.float64)
    ctx.check(lib.cf_scores(ctx.h, _ptr(counters), _ptr(histogram), F, T, C, _ptr(out)))
    return out


def argmax_scores(ctx, scores):
    """The reference's threshold-major scan with the 1e-13 tie rule over scores [T][F]:
    (best feature, best threshold, best score)."""
    scores = np.ascontiguousarray(scores, np.float64)
    T, F = scores.shape
    bf, bt, bs = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    ctx.check(lib.cf_argmax_scores(ctx.h, _ptr(scores), F, T, ctypes.byref(bf), ctypes.byref(bt), ctypes.byref(bs)))
    return bf.value, bt.value, bs.value


def node_best_split(ctx, images, samples, features, thresholds, num_classes, histogram):
    samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
    features = np.ascontiguousarray(features, np.int8)
    thresholds = np.ascontiguousarray(thresholds, np.float32)
    histogram = np.ascontiguousarray(histogram, np.uint32)
    T, F = thresholds.shape
    bf, bt, bs = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
    ctx.check(lib.cf_node_best_split(ctx.h, images.h, _ptr(samples), len(samples), _ptr(features), _ptr(thresholds),
                                     F, T, num_classes, _ptr(histogram), ctypes.byref(bf), ctypes.byref(bt),
                                     ctypes.byref(bs)))
    return bf.value, bt.value, bs.value


class TrainingConfiguration:
    """TrainingConfiguration(randomSeed, samplesPerImage, featureCount, minSampleCount, maxDepth, boxRadius,
    regionSize, thresholds, numThreads, maxImages, imageCacheSize, maxSamplesPerBatch, accelerationMode,
    useCIELab=True, useDepthFilling=False, deviceIds=[0], subsamplingType="classUniform", ignoredColors=[],
    useDepthImages=True, horizontalFlipping=False) -- random_tree.h:222-273.
    `ignored_labels` takes label ids (the reference maps ignored colours to labels)."""

    def __init__(self, randomSeed, samplesPerImage, featureCount, minSampleCount, maxDepth, boxRadius, regionSize,
                 thresholds, numThreads=1, maxImages=0, imageCacheSize=0, maxSamplesPerBatch=0,
                 accelerationMode="gpu", useCIELab=False, useDepthFilling=False, deviceIds=(0,),
                 subsamplingType="classUniform", ignored_labels=(), useDepthImages=True, horizontalFlipping=False):
        if maxImages > 0 and maxImages < imageCacheSize:
            raise ValueError(f"illegal configuration: maxImages ({maxImages}) must not be lower than "
                             f"imageCacheSize ({imageCacheSize})")
        if accelerationMode not in ("cpu", "gpu", "compare"):
            raise ValueError(f"illegal acceleration mode: {accelerationMode}")
        if subsamplingType not in ("classUniform", "pixelUniform"):
            raise ValueError(f"unknown subsamplingType: {subsamplingType}")
        if len(ignored_labels) > 8:
            raise ValueError("at most 8 ignored labels")
        self.randomSeed, self.samplesPerImage, self.featureCount = randomSeed, samplesPerImage, featureCount
        self.minSampleCount, self.maxDepth, self.boxRadius = minSampleCount, maxDepth, boxRadius
        self.regionSize, self.thresholds, self.numThreads = regionSize, thresholds, numThreads
        self.maxImages, self.imageCacheSize, self.maxSamplesPerBatch = maxImages, imageCacheSize, maxSamplesPerBatch
        self.accelerationMode, self.useCIELab, self.useDepthFilling = accelerationMode, useCIELab, useDepthFilling
        self.deviceIds, self.subsamplingType = list(deviceIds), subsamplingType
        self.ignored_labels, self.useDepthImages = list(ignored_labels), useDepthImages
        self.horizontalFlipping = horizontalFlipping

    @classmethod
    def from_c(cls, c):
        return cls(c.random_seed, c.samples_per_image, c.feature_count, c.min_sample_count, c.max_depth,
                   c.box_radius, c.region_size, c.thresholds, numThreads=c.num_threads, maxImages=c.max_images,
                   subsamplingType="pixelUniform" if c.subsampling_pixel_uniform else "classUniform",
                   ignored_labels=[c.ignored_labels[i] for i in range(c.num_ignored)],
                   useDepthImages=bool(c.use_depth_images), horizontalFlipping=bool(c.horizontal_flipping),
                   useDepthFilling=bool(c.use_depth_filling), useCIELab=bool(c.use_cielab))

    def to_c(self):
        c = _TrainConfig()
        c.random_seed, c.samples_per_image, c.feature_count = self.randomSeed, self.samplesPerImage, self.featureCount
        c.min_sample_count, c.max_depth = self.minSampleCount, self.maxDepth
        c.box_radius, c.region_size, c.thresholds = self.boxRadius, self.regionSize, self.thresholds
        c.num_threads, c.max_images = self.numThreads, self.maxImages
        c.use_depth_images, c.horizontal_flipping = int(self.useDepthImages), int(self.horizontalFlipping)
        c.subsampling_pixel_uniform = int(self.subsamplingType == "pixelUniform")
        c.use_depth_filling, c.use_cielab = int(self.useDepthFilling), int(self.useCIELab)
        c.num_ignored = len(self.ignored_labels)
        for i, l in enumerate(self.ignored_labels):
            c.ignored_labels[i] = l
        return c


class RandomForestImage:
    """RandomForestImage (random_forest_image.h:44-168) on the HIP path."""

    def __init__(self, ctx, h):
        self.ctx, self.h = ctx, h
        if ctx is not None:
            ctx._adopt(self)

    def _check(self, st):
        _check(self.ctx, st)

    def bind(self, ctx):
        """Attach a forest made without a context (from_files / from_trees with ctx=None) to the context
        that predicts with it; the library binds it at its first device call as well."""
        if self.ctx is None:
            self.ctx = ctx
            ctx._adopt(self)
        elif self.ctx is not ctx:
            raise ValueError("forest belongs to another context")
        return self

    @classmethod
    def train(cls, ctx, images, configuration, tree_count, num_labels, mode=MODE_BATCHED):
        h = ctypes.c_void_p()
        c = configuration.to_c()
        ctx.check(lib.cf_forest_train(ctx.h, images.h, ctypes.byref(c), tree_count, num_labels, mode,
                                      ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def sample(cls, ctx, images, configuration, tree_count, num_labels):
        """Pixel sampling of every tree; grow() then trains from these roots."""
        h = ctypes.c_void_p()
        c = configuration.to_c()
        ctx.check(lib.cf_forest_sample(ctx.h, images.h, ctypes.byref(c), tree_count, num_labels, ctypes.byref(h)))
        f = cls(ctx, h)
        f._images = images
        return f

    def grow(self, configuration, mode=MODE_BATCHED, images=None):
        c = configuration.to_c()
        im = images if images is not None else self._images
        self.ctx.check(lib.cf_forest_grow(self.ctx.h, im.h, self.h, ctypes.byref(c), mode))

    PROFILE_KEYS = ("evaluations", "seconds", "eval_kernel_ms", "eval_launches", "responses", "featgen_ms",
                    "score_ms", "partition_ms", "levels", "replays", "nodes", "hist_ms", "group_batches_max",
                    "slot_batches_max", "eval_shape", "level_graphs")

    def profile(self):
        out = np.zeros(len(self.PROFILE_KEYS), np.float64)
        lib.cf_forest_profile(self.h, _ptr(out), len(out))
        return dict(zip(self.PROFILE_KEYS, out.tolist()))

    @classmethod
    def train_tree_from_samples(cls, ctx, images, configuration, tree_id, num_classes, rs_seed, roots,
                                mode=MODE_BATCHED):
        roots = np.ascontiguousarray(roots, SAMPLE_DTYPE)
        h = ctypes.c_void_p()
        c = configuration.to_c()
        ctx.check(lib.cf_tree_train_from_samples(ctx.h, images.h, ctypes.byref(c), tree_id, num_classes, rs_seed,
                                                 _ptr(roots), len(roots), mode, ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_trees(cls, ctx, trees, num_classes, prior=None):
        """trees: list of (nodes NODE_DTYPE array, histograms uint32 [nodes][C])"""
        counts = np.array([len(n) for n, _ in trees], np.int32)
        nodes = np.ascontiguousarray(np.concatenate([n for n, _ in trees]), NODE_DTYPE)
        hist = np.ascontiguousarray(np.concatenate([h for _, h in trees]), np.uint32)
        pr = None if prior is None else np.ascontiguousarray(prior, np.uint32)
        h = ctypes.c_void_p()
        _check(ctx, lib.cf_forest_create(ctx.h if ctx is not None else None, len(trees), _ptr(counts),
                                         _ptr(nodes), _ptr(hist), num_classes, _ptr(pr), ctypes.byref(h)))
        return cls(ctx, h)

    @classmethod
    def from_files(cls, tree_files, ctx=None, histogramBias=0.0, useLabelsPrior=True):
        """RandomForestImage(treeFiles, deviceIds, accelerationMode, histogramBias, useLabelsPrior)
        (random_forest_image.cpp:47-107): reads tree<N>.json[.gz] files (RandomTreeImport::readJSON,
        import.cpp:96-265) and normalizes every tree with its own label prior.  `configuration` holds the
        first tree's training configuration, `palette` [C][3] the colour of each label."""
        paths = [os.fsencode(str(p)) for p in tree_files]
        arr = (ctypes.c_char_p * len(paths))(*paths)
        c = _TrainConfig()
        pal = np.zeros((32, 3), np.uint8)
        h = ctypes.c_void_p()
        _check(ctx, lib.cf_forest_import(ctx.h if ctx is not None else None, arr, len(paths), ctypes.byref(c),
                                         _ptr(pal), ctypes.byref(h)))
        f = cls(ctx, h)
        f.configuration = TrainingConfiguration.from_c(c)
        f.palette = pal[:f.num_classes].copy()
        f.normalizeHistograms(histogramBias, useLabelsPrior)
        return f

    def export_tree(self, tree, path, configuration, palette=None, folder_training=""):
        """One tree as JSON, gzip-compressed when `path` ends in .gz (RandomTreeExport::writeJSON,
        export.cpp:180-313).  palette [C][3] gives the label colours of the histogram keys."""
        pal = None if palette is None else np.ascontiguousarray(palette, np.uint8).reshape(-1, 3)
        if pal is not None and len(pal) < self.num_classes:
            raise ValueError("palette needs a colour per class")
        c = configuration.to_c()
        self._check(lib.cf_forest_export_tree(self.h, tree, os.fsencode(str(path)), ctypes.byref(c), _ptr(pal),
                                              os.fsencode(str(folder_training))))

    def export(self, output_folder, configuration, palette=None, folder_training=""):
        """Every tree as <output_folder>/tree<N>.json.gz (RandomTreeExport::writeJSON's file names,
        export.cpp:292-313); returns the paths."""
        os.makedirs(output_folder, exist_ok=True)
        paths = []
        for t in range(self.num_trees):
            p = os.path.join(str(output_folder), f"tree{t}.json.gz")
            self.export_tree(t, p, configuration, palette, folder_training)
            paths.append(p)
        return paths

    @property
    def num_trees(self):
        return lib.cf_forest_num_trees(self.h)

    @property
    def num_classes(self):
        return lib.cf_forest_num_classes(self.h)

    def tree(self, t):
        n = lib.cf_forest_tree_size(self.h, t)
        nodes = np.zeros(n, NODE_DTYPE)
        hist = np.zeros((n, self.num_classes), np.uint32)
        self._check(lib.cf_forest_get_tree(self.h, t, _ptr(nodes), _ptr(hist)))
        return nodes, hist

    def root_samples(self, t):
        n = lib.cf_forest_get_root_samples(self.h, t, None, 0)
        out = np.zeros(n, SAMPLE_DTYPE)
        lib.cf_forest_get_root_samples(self.h, t, _ptr(out), n)
        return out

    def prior(self, t=None):
        """the forest's label prior, or tree t's (its own when it trained on a subset of the images)"""
        out = np.zeros(self.num_classes, np.uint32)
        if t is None:
            lib.cf_forest_get_prior(self.h, _ptr(out))
        elif lib.cf_forest_get_tree_prior(self.h, t, _ptr(out)) < 0:
            raise IndexError(t)
        return out

    def train_stats(self):
        ev, sec = ctypes.c_double(), ctypes.c_double()
        replays = lib.cf_forest_train_stats(self.h, ctypes.byref(ev), ctypes.byref(sec))
        return {"evaluations": ev.value, "seconds": sec.value, "replays": replays}

    def normalizeHistograms(self, histogramBias=0.0, useLabelsPrior=True):
        self._check(lib.cf_forest_normalize(self.h, histogramBias, int(useLabelsPrior)))

    def posteriors(self, t):
        n = lib.cf_forest_tree_size(self.h, t)
        out = np.zeros((n, self.num_classes), np.float64)
        self._check(lib.cf_forest_get_posteriors(self.h, t, _ptr(out)))
        return out

    def predict(self, images, index, probabilities=True):
        C, H, W = self.num_classes, images.height, images.width
        probs = np.zeros((C, H, W), np.float32) if probabilities else None
        labels = np.zeros((H, W), np.uint8)
        self.bind(images.ctx)
        self.ctx.check(lib.cf_predict(self.ctx.h, self.h, images.h, index, _ptr(probs), _ptr(labels)))
        return labels, probs

    def predict_batch(self, images, indices, probabilities=True, out=None):
        """Labels [n][H][W] and (with `probabilities`) posteriors [n][C][H][W] of frames `indices`.  `out`
        = (labels, probs) arrays to fill, e.g. from pinned_empty (page-locked: copied at the link's
        rate); probs may be None."""
        idx = np.ascontiguousarray(indices, np.int32)
        C, H, W = self.num_classes, images.height, images.width
        if out is not None:
            labels, probs = out
            assert labels.dtype == np.uint8 and labels.shape == (len(idx), H, W) and labels.flags.c_contiguous
            if probs is not None:
                assert probs.dtype == np.float32 and probs.shape == (len(idx), C, H, W) and probs.flags.c_contiguous
        else:
            probs = np.empty((len(idx), C, H, W), np.float32) if probabilities else None
            labels = np.empty((len(idx), H, W), np.uint8)
        self.bind(images.ctx)
        self.ctx.check(lib.cf_predict_batch(self.ctx.h, self.h, images.h, _ptr(idx), len(idx), _ptr(probs),
                                            _ptr(labels)))
        return labels, probs

    def predict_batch_device(self, images, indices, probs_ptr, labels_ptr):
        """Outputs into caller-owned device memory (integer addresses, e.g. torch tensors' data_ptr();
        0 skips an output): probs [n][C][H][W] float32, labels [n][H][W] uint8."""
        idx = np.ascontiguousarray(indices, np.int32)
        self.bind(images.ctx)
        self.ctx.check(lib.cf_predict_batch_device(self.ctx.h, self.h, images.h, _ptr(idx), len(idx),
                                                   ctypes.c_void_p(probs_ptr or None),
                                                   ctypes.c_void_p(labels_ptr or None)))

    def predict_batch_device_noalloc(self, images, indices, probs_ptr, labels_ptr):
        """predict_batch_device without any allocation (the forest must already be on the device: predict
        with it once first); at most 65536 frames per call."""
        idx = np.ascontiguousarray(indices, np.int32)
        ctx = images.ctx
        ctx.check(lib.cf_predict_batch_device_noalloc(ctx.h, self.h, images.h, _ptr(idx), len(idx),
                                                      ctypes.c_void_p(probs_ptr or None),
                                                      ctypes.c_void_p(labels_ptr or None)))

    def close(self):
        if self.h:
            lib.cf_forest_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- datasets (PNG files)

def parse_color(text):
    """RGBColor(colorString) (image.cpp:66-91): "r,g,b" with every value in 0..255."""
    parts = str(text).split(",")
    try:
        if len(parts) != 3:
            raise ValueError("expected three values")
        rgb = tuple(int(p) for p in parts)
    except ValueError as e:
        raise ValueError(f"illegal color: '{text}': {e}") from None
    if any(v < 0 or v > 255 for v in rgb):
        raise ValueError(f"illegal color: '{text}': color value out of range [0-255]: {text}")
    return rgb


class LabelMap:
    """The colour -> label map of ground-truth images (LabelImage::encodeColor / decodeLabel,
    image.cpp:656-734): labels in the order colours are met; shared by training and test sets."""

    def __init__(self):
        h = ctypes.c_void_p()
        _check(None, lib.cf_label_map_create(ctypes.byref(h)))
        self.h = h

    def encode(self, rgb):
        c = np.ascontiguousarray(rgb, np.uint8)
        out = ctypes.c_int()
        _check(None, lib.cf_label_map_encode(self.h, _ptr(c), ctypes.byref(out)))
        return out.value

    def find(self, rgb):
        r = lib.cf_label_map_find(self.h, _ptr(np.ascontiguousarray(rgb, np.uint8)))
        return None if r < 0 else r

    def __len__(self):
        return lib.cf_label_map_size(self.h)

    def colors(self):
        out = np.zeros((max(len(self), 1), 3), np.uint8)
        lib.cf_label_map_colors(self.h, _ptr(out))
        return out[:len(self)]

    def close(self):
        if self.h:
            lib.cf_label_map_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def list_image_filenames(folder):
    """listImageFilenames (image.cpp:844-866): every file under `folder` whose name contains
    '_colors.png', recursively, sorted by path."""
    out = []
    for root, _, files in os.walk(folder):
        out.extend(os.path.join(root, f) for f in files if "_colors.png" in f)
    return sorted(out)


class Dataset:
    """Images loaded from `*_colors.png` / `*_depth.png` / `*_ground_truth.png` triples."""

    def __init__(self, h, label_map, files):
        self.h, self.label_map, self.files = h, label_map, list(files)
        c, w, hh, pad = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        lib.cf_dataset_info(h, ctypes.byref(c), ctypes.byref(w), ctypes.byref(hh), ctypes.byref(pad))
        self.count, self.width, self.height, self.padding_label = c.value, w.value, hh.value, pad.value

    @classmethod
    def load(cls, color_files, useDepthImages=True, useDepthFilling=False, padding_color=None, label_map=None,
             threads=0, useCIELab=False):
        label_map = label_map if label_map is not None else LabelMap()
        files = [os.fsencode(str(f)) for f in color_files]
        arr = (ctypes.c_char_p * len(files))(*files)
        pc = None if padding_color is None else np.ascontiguousarray(padding_color, np.uint8)
        h = ctypes.c_void_p()
        _check(None, lib.cf_dataset_load(arr, len(files), int(useDepthImages), int(useDepthFilling), int(useCIELab),
                                         _ptr(pc), label_map.h, threads, ctypes.byref(h)))
        return cls(h, label_map, color_files)

    def color(self, index):
        """colour planes [3][h][w] float as the forest sees them (RGB values or CIELab)"""
        w, h = self.image_size(index)
        out = np.zeros((3, h, w), np.float32)
        _check(None, lib.cf_dataset_get_color(self.h, index, _ptr(out)))
        return out

    def image_size(self, index):
        w, h = ctypes.c_int(), ctypes.c_int()
        _check(None, lib.cf_dataset_image_size(self.h, index, ctypes.byref(w), ctypes.byref(h)))
        return w.value, h.value

    def get(self, index):
        """rgb [h][w][3] u8, depth [h][w] u16 mm (after filling), labels [h][w] u8, at the image's own size"""
        w, h = self.image_size(index)
        rgb = np.zeros((h, w, 3), np.uint8)
        depth = np.zeros((h, w), np.uint16)
        labels = np.zeros((h, w), np.uint8)
        _check(None, lib.cf_dataset_get(self.h, index, _ptr(rgb), _ptr(depth), _ptr(labels)))
        return rgb, depth, labels

    def to_images(self, ctx):
        """all images on the device at the largest size, labels set"""
        h = ctypes.c_void_p()
        ctx.check(lib.cf_dataset_to_images(self.h, ctx.h, ctypes.byref(h)))
        return Images(ctx, self.count, self.width, self.height, _handle=h)

    def close(self):
        if self.h:
            lib.cf_dataset_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def load_images(folder, useCIELab=False, useDepthImages=True, useDepthFilling=False, ignoredColors=(),
                label_map=None, threads=0):
    """loadImages(folder, useCIELab, useDepthImages, useDepthFilling, ignoredColors, numLabels)
    (image.cpp:869-965).  Returns the Dataset; its label_map holds the colours met so far, and
    len(label_map) is the reference's numLabels."""
    files = list_image_filenames(folder)
    if not files:
        raise CurfilError(CF_ERR_ARG, f"found no files in {folder}")
    padding = parse_color(ignoredColors[0]) if ignoredColors else (0, 0, 0)
    return Dataset.load(files, useDepthImages, useDepthFilling, padding, label_map, threads, useCIELab)


class ConfusionMatrix:
    """ConfusionMatrix (predict.h / predict.cpp:44-140): counts [C][C], rows = labels, columns = predictions."""

    def __init__(self, numClasses, ignoredLabels=()):
        self.data = np.zeros((numClasses, numClasses), np.float64)
        self.ignoredLabels = np.ascontiguousarray(list(ignoredLabels), np.int32)
        self.normalized = False

    @property
    def numClasses(self):
        return self.data.shape[0]

    def __iadd__(self, other):
        if self.normalized:
            raise RuntimeError("confusion matrix is already normalized")
        if other.numClasses != self.numClasses:
            raise RuntimeError(f"different number of classes in confusion matrix: {self.numClasses} and "
                               f"{other.numClasses}")
        self.data += other.data
        return self

    def normalize(self):
        if self.normalized:
            raise RuntimeError("confusion matrix is already normalized")
        _check(None, lib.cf_confusion_normalize(_ptr(self.data), self.numClasses))
        self.normalized = True

    def averageClassAccuracy(self, includeVoid=True):
        out = ctypes.c_double()
        _check(None, lib.cf_confusion_class_accuracy(_ptr(self.data), self.numClasses, int(self.normalized),
                                                     int(includeVoid), _ptr(self.ignoredLabels),
                                                     len(self.ignoredLabels), ctypes.byref(out)))
        return out.value


def calculatePixelAccuracy(prediction, groundTruth, includeVoid=True, ignoredLabels=(), confusionMatrix=None):
    """calculatePixelAccuracy (predict.cpp:142-199)"""
    p = np.ascontiguousarray(prediction, np.uint8)
    g = np.ascontiguousarray(groundTruth, np.uint8)
    if p.shape != g.shape:
        raise ValueError("prediction and ground truth differ in size")
    ign = np.ascontiguousarray(list(ignoredLabels), np.int32)
    conf = None
    if confusionMatrix is not None:
        need = int(g.max()) + 1 if g.size else 0
        if confusionMatrix.numClasses < need:       # the reference grows the matrix to the labels seen
            confusionMatrix.data = np.zeros((need, need), np.float64)
        conf = confusionMatrix.data
    out = ctypes.c_double()
    _check(None, lib.cf_pixel_accuracy(_ptr(p), _ptr(g), p.size, int(includeVoid), _ptr(ign), len(ign), _ptr(conf),
                                       0 if conf is None else conf.shape[0], ctypes.byref(out)))
    return out.value


def run_tool(name, args):
    """Runs curfil_train ('train') or curfil_predict ('predict') in this process with the given command-line
    arguments (the same as the executables in curfil_amd/bin take); returns the exit status."""
    fn = {"train": lib.cf_tool_train, "predict": lib.cf_tool_predict}[name]
    argv = [f"curfil_{name}"] + [str(a) for a in args]
    arr = (ctypes.c_char_p * len(argv))(*[os.fsencode(a) for a in argv])
    return fn(len(argv), arr)


def read_png(path):
    """PNG samples as stored: [h][w] for one channel, else [h][w][channels]; uint8 for 8-bit, uint16 for 16-bit"""
    w, h, c, d = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    p = os.fsencode(str(path))
    _check(None, lib.cf_png_read_info(p, ctypes.byref(w), ctypes.byref(h), ctypes.byref(c), ctypes.byref(d)))
    out = np.zeros((h.value, w.value, c.value), np.uint16)
    _check(None, lib.cf_png_read(p, _ptr(out), out.size))
    out = out.astype(np.uint8) if d.value == 8 else out
    return out[:, :, 0] if c.value == 1 else out


def write_png(path, pixels):
    """[h][w][3] uint8 RGB, or [h][w] uint8 / uint16 grayscale"""
    a = np.ascontiguousarray(pixels)
    if a.ndim == 3 and a.shape[2] == 3 and a.dtype == np.uint8:
        ch, bits = 3, 8
    elif a.ndim == 2 and a.dtype in (np.uint8, np.uint16):
        ch, bits = 1, 8 * a.dtype.itemsize
    else:
        raise ValueError("write_png takes [h][w][3] uint8 or [h][w] uint8/uint16")
    _check(None, lib.cf_png_write(os.fsencode(str(path)), a.shape[1], a.shape[0], ch, bits, _ptr(a)))


def synth_rgbd(seed, width, height, sites, classes, zero_frac, class_seed=4711):
    """One synthetic labelled RGB-D frame (rgb [H][W][3] u8, depth [H][W] u16 mm, labels [H][W] u8)."""
    rgb = np.zeros((height, width, 3), np.uint8)
    depth = np.zeros((height, width), np.uint16)
    labels = np.zeros((height, width), np.uint8)
    st = lib.cf_synth_rgbd(seed, class_seed, width, height, sites, classes, zero_frac, _ptr(rgb), _ptr(depth),
                           _ptr(labels))
    if st != CF_OK:
        raise CurfilError(st, "cf_synth_rgbd failed")
    return rgb, depth, labels


class DeviceBuffer:
    """Device memory owned by the caller (cf_device_alloc); `ptr` is its address."""

    def __init__(self, ctx, nbytes):
        self.ctx, self.nbytes = ctx, nbytes
        p = ctypes.c_void_p()
        ctx.check(lib.cf_device_alloc(ctx.h, nbytes, ctypes.byref(p)))
        self.ptr = p.value
        ctx._adopt(self)

    def to_host(self, dtype, shape):
        out = np.empty(shape, dtype)
        assert out.nbytes <= self.nbytes
        self.ctx.check(lib.cf_device_to_host(self.ctx.h, _ptr(out), ctypes.c_void_p(self.ptr), out.nbytes))
        return out

    def close(self):
        if getattr(self, "ptr", None):
            lib.cf_device_free(self.ctx.h, ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        self.close()


class _PinnedBlock:
    """Page-locked host memory (cf_host_alloc), freed when the last array over it goes away."""

    def __init__(self, ctx, nbytes):
        self.ctx = ctx
        p = ctypes.c_void_p()
        ctx.check(lib.cf_host_alloc(ctx.h, max(1, nbytes), ctypes.byref(p)))
        self.ptr = p.value
        self.nbytes = nbytes

    def __del__(self):
        if getattr(self, "ptr", None) and self.ctx.h:
            lib.cf_host_free(self.ctx.h, ctypes.c_void_p(self.ptr))
            self.ptr = None


def pinned_empty(ctx, shape, dtype):
    """An uninitialised numpy array in page-locked host memory (device copies into it run at DMA rate)."""
    dtype = np.dtype(dtype)
    n = int(np.prod(shape)) * dtype.itemsize
    block = _PinnedBlock(ctx, n)
    buf = (ctypes.c_char * max(1, n)).from_address(block.ptr)
    buf._block = block     # the ctypes buffer keeps the memory alive; numpy keeps the buffer
    return np.frombuffer(buf, dtype=dtype, count=int(np.prod(shape))).reshape(shape)


class Server:
    """Serving loop over a resident forest (cf_server_*): submit frames as they arrive, wait for results."""

    def __init__(self, ctx, forest, budget_mb=8192.0, max_batch=64):
        self.ctx, self.forest = ctx, forest.bind(ctx)
        self.C = forest.num_classes
        h = ctypes.c_void_p()
        ctx.check(lib.cf_server_create(ctx.h, forest.h, float(budget_mb), max_batch, ctypes.byref(h)))
        self.h = h
        self.sizes = {}
        ctx._adopt(self)

    def submit(self, rgb, depth):
        rgb = np.ascontiguousarray(rgb, np.uint8)
        depth = np.ascontiguousarray(depth, np.uint16)
        H, W = depth.shape
        rid = ctypes.c_uint64()
        st = lib.cf_server_submit(self.h, _ptr(rgb), _ptr(depth), W, H, ctypes.byref(rid))
        if st != CF_OK:
            raise CurfilError(st, "cf_server_submit failed" + (": " + lib.cf_server_error(self.h).decode()))
        self.sizes[rid.value] = (H, W)
        return rid.value

    def wait(self, rid, probabilities=True, labels=True):
        H, W = self.sizes.pop(rid)
        probs = np.zeros((self.C, H, W), np.float32) if probabilities else None
        lab = np.zeros((H, W), np.uint8) if labels else None
        lat = np.zeros(4, np.float64)
        st = lib.cf_server_wait(self.h, rid, _ptr(probs), _ptr(lab), _ptr(lat))
        if st != CF_OK:
            raise CurfilError(st, "cf_server_wait failed: " + lib.cf_server_error(self.h).decode())
        self.last_breakdown = lat[1:].copy()
        return lab, probs, float(lat[0])

    STATS_KEYS = ("frames", "pixels", "launches", "device_mb", "queued", "passes", "longest_pass_ms",
                  "passes_over_1ms", "slot_allocs", "pinned_allocs", "event_allocs")

    def stats(self):
        out = np.zeros(len(self.STATS_KEYS), np.float64)
        lib.cf_server_stats(self.h, _ptr(out), len(out))
        return dict(zip(self.STATS_KEYS, out.tolist()))

    def close(self):
        if getattr(self, "h", None):
            lib.cf_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


def feature_responses_paired(ctx, images, samples, features):
    """Response of sample i to feature i (features [11][n] int8) -> float64 [n]."""
    samples = np.ascontiguousarray(samples, SAMPLE_DTYPE)
    features = np.ascontiguousarray(features, np.int8)
    out = np.zeros(len(samples), np.float64)
    ctx.check(lib.cf_feature_responses_paired(ctx.h, images.h, _ptr(samples), len(samples), _ptr(features),
                                              _ptr(out)))
    return out


def synth_forest(ctx, images, depth_mm, trees=10, depth=20, classes=5, seed=4711, box_radius=120, region_size=20,
                 retries=10):
    """A forest of complete trees with random splits (benchmark workload; no reference counterpart).

    Per split: type U{0,1}, offsets U[-box_radius, box_radius]^2, regions U[1, region_size]^2, colour
    channels U{0,1,2}; the threshold is the (float) response at a random pixel of a random frame of
    `images` (redrawn up to `retries` times while NaN, else 0).  Leaves: Dirichlet(1) posteriors over
    `classes` (normalised Exp(1) draws) stored as counts scaled to 2^20, so normalizeHistograms(0,
    False) reproduces them.  Trees are complete, in breadth-first order (children of i at 2i+1, 2i+2).
    `depth_mm`: the raw depth frames [n][H][W] the images were built from (pixel validity)."""
    rng = np.random.default_rng(seed)
    n_frames, H, W = depth_mm.shape
    n_nodes = (1 << depth) - 1
    n_split = (1 << (depth - 1)) - 1
    idx = np.arange(n_nodes, dtype=np.int64)
    out = []
    for _ in range(trees):
        nodes = np.zeros(n_nodes, NODE_DTYPE)
        nodes["id"] = idx
        nodes["level"] = np.floor(np.log2(idx + 1)).astype(np.int32) + 1
        nodes["parent"] = np.where(idx > 0, (idx - 1) // 2, -1)
        split = idx < n_split
        nodes["left"] = np.where(split, 2 * idx + 1, -1)
        nodes["right"] = np.where(split, 2 * idx + 2, -1)
        nodes["is_leaf"] = ~split
        f = np.zeros((11, n_split), np.int8)
        f[0] = rng.integers(0, 2, n_split)
        for r in (1, 2, 3, 4):
            f[r] = rng.integers(-box_radius, box_radius + 1, n_split)
        for r in (5, 6, 7, 8):
            f[r] = rng.integers(1, region_size + 1, n_split)
        col = f[0] == COLOR
        f[9] = np.where(col, rng.integers(0, 3, n_split), 0)
        f[10] = np.where(col, rng.integers(0, 3, n_split), 0)
        thr = np.zeros(n_split, np.float32)
        todo = np.arange(n_split)
        for _ in range(retries):
            if len(todo) == 0:
                break
            s = np.zeros(len(todo), SAMPLE_DTYPE)
            s["image"] = rng.integers(0, n_frames, len(todo))
            s["x"] = rng.integers(0, W, len(todo))
            s["y"] = rng.integers(0, H, len(todo))
            s["depth_mm"] = depth_mm[s["image"], s["y"], s["x"]]
            r = feature_responses_paired(ctx, images, s, np.ascontiguousarray(f[:, todo]))
            ok = ~np.isnan(r)
            thr[todo[ok]] = r[ok].astype(np.float32)
            todo = todo[~ok]
        nodes["feature"][:n_split] = f.T
        nodes["threshold"][:n_split] = thr
        hist = np.ones((n_nodes, classes), np.uint32)
        e = rng.exponential(1.0, (n_nodes - n_split, classes))
        hist[n_split:] = np.maximum(np.rint(e / e.sum(axis=1, keepdims=True) * (1 << 20)), 0).astype(np.uint32)
        out.append((nodes, hist))
    return RandomForestImage.from_trees(ctx, out, classes)


def eval_shape_info(shape):
    """(lanes, samples per lane, features per step, workgroups per CU) of a pass-1 launch shape, or None."""
    v = [ctypes.c_int() for _ in range(4)]
    if lib.cf_eval_shape_info(shape, *[ctypes.byref(x) for x in v]) != CF_OK:
        return None
    return tuple(x.value for x in v)


def eval_shapes():
    out = []
    while eval_shape_info(len(out)) is not None:
        out.append(eval_shape_info(len(out)))
    return out


def eval_tuning():
    """The pass-1 tuner's table: one dict per (threshold width, level size class)."""
    n = lib.cf_eval_tuning(None, 0)
    buf = np.zeros(6 * max(n, 1), np.float64)
    lib.cf_eval_tuning(_ptr(buf), len(buf))
    rows = []
    for k in range(n):
        tm, cls, best, ns_best, ns_default, timed = buf[6 * k:6 * k + 6]
        rows.append({"threshold_width": int(tm), "size_class": int(cls), "shape": int(best),
                     "ns_per_response": ns_best, "default_ns_per_response": ns_default, "launches_timed": int(timed)})
    return rows


def eval_tuning_reset():
    lib.cf_eval_tuning_reset()


def synth_dataset(n, width, height, sites, classes, zero_frac, seed0=4711, threads=8):
    """n frames with seeds seed0 + i (ctypes releases the GIL, so frames build in parallel)."""
    import concurrent.futures
    rgb = np.zeros((n, height, width, 3), np.uint8)
    depth = np.zeros((n, height, width), np.uint16)
    labels = np.zeros((n, height, width), np.uint8)

    def one(i):
        st = lib.cf_synth_rgbd(seed0 + i, seed0, width, height, sites, classes, zero_frac,
                               _ptr(rgb[i]), _ptr(depth[i]), _ptr(labels[i]))
        if st != CF_OK:
            raise CurfilError(st, "cf_synth_rgbd failed")

    with concurrent.futures.ThreadPoolExecutor(max_workers=threads) as ex:
        list(ex.map(one, range(n)))
    return rgb, depth, labels
