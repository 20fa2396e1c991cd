/*
 * curfil_hip.h -- C ABI of the MI355X-native CURFIL hot path.
 *
 * Plain C, plain pointers and sizes, no exceptions across the boundary: every
 * call returns a status (CF_OK == 0) and cf_last_error() gives the message of
 * the last failure on that context.  Host buffers are owned by the caller;
 * device buffers are owned by the handles.  One context per host thread (a
 * context serialises its own calls with an internal lock).
 *
 * Each entry point names the reference interface it replaces
 * (deeplearningais/curfil, src/curfil/...).
 */
#ifndef CURFIL_HIP_H
#define CURFIL_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CF_OK = 0,
    CF_ERR_ARG = -1,         /* invalid argument (reference: assert / runtime_error on bad input) */
    CF_ERR_HIP = -2,         /* HIP runtime failure (reference: cudaSafeCall, utils.h:48) */
    CF_ERR_RUNTIME = -3,     /* algorithmic runtime_error of the reference (message says which) */
    CF_ERR_UNSUPPORTED = -4, /* option not implemented on this path */
    CF_ERR_NOMEM = -5        /* device memory exhausted / HBM budget exceeded */
};

typedef struct cf_ctx cf_ctx;
typedef struct cf_images cf_images;
typedef struct cf_forest cf_forest;

/* One training pixel; the fields of Samples (random_tree_image.h:922-978) and
 * PixelInstance (random_tree_image.h:129-392).  depth_mm <= 0 means invalid.
 * flip: 0 NoFlip, 1 Flip, 2 Both (random_tree.h:58-60). */
typedef struct cf_sample {
    uint16_t x, y;
    uint32_t image;
    int32_t depth_mm;
    uint8_t label, flip, pad0, pad1;
} cf_sample;

/* One node of a trained tree, in creation order (root first, then per level
 * the children pairs of the level's nodes in frontier order).  Mirrors
 * RandomTree (random_tree.h:499-1003) + SplitFunction (random_tree.h:89-181).
 * feature[] is the 11-byte SoA column of ImageFeaturesAndThresholds
 * (random_tree_image.h:702-773): type, offset1 x/y, offset2 x/y,
 * region1 x/y, region2 x/y, channel1, channel2. */
typedef struct cf_tree_node {
    int32_t id, level, parent, left, right;
    int32_t num_samples;
    int32_t feature_id;
    float threshold;
    double score;
    int8_t feature[11];
    uint8_t is_leaf;
    int32_t pad;
} cf_tree_node;

/* TrainingConfiguration (random_tree.h:222-273), the fields the hot path reads. */
typedef struct cf_train_config {
    int32_t random_seed;
    uint32_t samples_per_image;
    uint32_t feature_count;
    uint32_t min_sample_count;
    int32_t max_depth;
    uint16_t box_radius;
    uint16_t region_size;
    uint16_t thresholds;
    uint16_t pad0;
    int32_t num_threads;
    int32_t max_images;
    int32_t use_depth_images;
    int32_t horizontal_flipping;
    int32_t subsampling_pixel_uniform; /* 0: classUniform (default) */
    int32_t num_ignored;
    int32_t ignored_labels[8];
    int32_t use_depth_filling;         /* recorded in tree files; the loader fills depth */
    int32_t use_cielab;                /* recorded in tree files; the loader converts colour */
} cf_train_config;

/* Training execution mode. */
enum {
    CF_MODE_PLAIN = 0,   /* one node per launch, host sync per node (the reference GPU path's shape) */
    CF_MODE_BATCHED = 1  /* every open node of every tree at a level in one launch sequence */
};

/* ---- context -------------------------------------------------------------- */
/* replaces ImageFeatureEvaluation::initDevice / selectDevice (random_tree_image_gpu.cu:1722-1760) */
int cf_ctx_create(int device, cf_ctx** out);
int cf_ctx_destroy(cf_ctx* ctx);
/* the message of the last failure on ctx; with NULL, of the last failure of a
 * call made without a context on this host thread (tree files, datasets, PNG) */
const char* cf_last_error(const cf_ctx* ctx);
int cf_ctx_synchronize(cf_ctx* ctx);

/* ---- images (device-resident integral images) ----------------------------- */
/* replaces RGBDImage + calculateIntegral (image.cpp:531-556) and the image
 * cache upload (random_tree_image_gpu.cu:580-1059).  All images of one set
 * have the same size (loadImages pads to the maximum size, image.cpp:910-955). */
int cf_images_create(cf_ctx* ctx, int count, int width, int height, cf_images** out);
int cf_images_destroy(cf_images* images);
/* 8-bit RGB (interleaved [H][W][3]) + uint16 depth in mm (0 = missing). The
 * integrals are built on the device (exact integer sums, colour rounded once
 * to float: identical to the reference's double Kahan integration for
 * integer input). */
int cf_images_set_rgbd_u8(cf_images* images, int index, const uint8_t* rgb, const uint16_t* depth_mm);
/* float colour planes [3][H][W] + int32 depth in mm (<= 0 invalid).  The
 * colour integral follows the reference's sequential double Kahan recurrence
 * exactly (image.cpp:236-277). */
int cf_images_set_rgbd_f32(cf_images* images, int index, const float* color, const int32_t* depth_mm);
/* an image smaller than the set (width <= W, height <= H) as loadImages pads
 * it: integrated at its own size, then colour NaN and depth 0 outside it
 * (RGBDImage::resizeImage on integrated images, image.cpp:449-475 via
 * image.cpp:945-960).  Equal sizes: cf_images_set_rgbd_u8. */
int cf_images_set_rgbd_u8_padded(cf_images* images, int index, const uint8_t* rgb, const uint16_t* depth_mm,
        int width, int height);
/* the same for float colour planes [3][height][width] (e.g. CIELab) and int32 depth */
int cf_images_set_rgbd_f32_padded(cf_images* images, int index, const float* color, const int32_t* depth_mm,
        int width, int height);
/* label image [H][W] (host copy, read by pixel sampling) */
int cf_images_set_labels(cf_images* images, int index, const uint8_t* labels);
/* read back the integrals of one image: color [3][H][W], depth sum and valid count [H][W] */
int cf_images_get_integrals(cf_images* images, int index, float* color, int32_t* depth, int32_t* valid);

/* ---- training seams -------------------------------------------------------- */
/* replaces ImageFeatureEvaluation::generateRandomFeatures (CPU generator,
 * random_tree_image.cpp:577-700, sorted as in :504-566).  features_out
 * [11][F] int8, thresholds_out [T][F] float.  Runs on the device. */
int cf_generate_features(cf_ctx* ctx, cf_images* images, const cf_sample* samples, int n, int seed,
        int num_features, int num_thresholds, int box_radius, int region_size, int use_depth_images,
        int8_t* features_out, float* thresholds_out);

/* feature responses (double) of n samples for F features -> out[F][n];
 * replaces featureResponseKernel (random_tree_image_gpu.cu:1455-1545) as
 * observed by random_tree_image_gpu_test.cu. */
int cf_feature_responses(cf_ctx* ctx, cf_images* images, const cf_sample* samples, int n,
        const int8_t* features, int num_features, double* out);

/* response of sample i to feature i (features [11][n] int8) -> out[n]:
 * ImageFeatureFunction::calculateFeatureResponse for one instance
 * (random_tree_image.h:485-623). */
int cf_feature_responses_paired(cf_ctx* ctx, cf_images* images, const cf_sample* samples, int n,
        const int8_t* features, double* out);

/* per-node left/right class counters, layout [F][T][C][2] (counters(feat,
 * thresh, label, lr), random_tree_image_gpu.cu:1069-1085); replaces
 * calculateFeatureResponsesAndHistograms (random_tree_image_gpu.cu:1958-2135). */
int cf_node_counters(cf_ctx* ctx, cf_images* images, const cf_sample* samples, int n,
        const int8_t* features, const float* thresholds, int num_features, int num_thresholds,
        int num_classes, uint32_t* counters);

/* normalized information-gain scores [T][F] of one node from its counters
 * [F][T][C][2] and node histogram [C]; replaces calculateScores
 * (random_tree_image.cpp:188-237 / scoreKernel random_tree_image_gpu.cu:1555-1601). */
int cf_scores(cf_ctx* ctx, const uint32_t* counters, const uint32_t* histogram, int num_features,
        int num_thresholds, int num_classes, double* scores);

/* the reference's host argmax over one node's scores [T][F]: threshold-major
 * sequential scan with detail::isScoreBetter's 1e-13 rule
 * (random_tree_image.cpp:445-457, random_tree.cpp:36-52), on the device. */
int cf_argmax_scores(cf_ctx* ctx, const double* scores, int num_features, int num_thresholds, int* best_feature,
        int* best_threshold, double* best_score);

/* best split of one node (plain path): counters, scores and the reference's
 * sequential argmax with its 1e-13 tie rule (random_tree_image.cpp:445-457,
 * random_tree.cpp:36-52), all on the device. */
int cf_node_best_split(cf_ctx* ctx, cf_images* images, const cf_sample* samples, int n,
        const int8_t* features, const float* thresholds, int num_features, int num_thresholds,
        int num_classes, const uint32_t* histogram, int* best_feature, int* best_threshold, double* best_score);

/* ---- forest training ------------------------------------------------------- */
/* RandomForestImage(treeCount, configuration).train(images, numLabels)
 * (random_forest_image.cpp:132-182): per-tree class-uniform pixel sampling on
 * the host, then level-wise training on the device in the given mode. */
int cf_forest_train(cf_ctx* ctx, cf_images* images, const cf_train_config* config, int num_trees,
        int num_labels, int mode, cf_forest** out);
/* the two halves of cf_forest_train: pixel sampling of every tree (host,
 * one thread per tree; RandomTreeImage::train up to doTrain,
 * random_tree_image.cpp:746-812), then growing all trees from the stored
 * roots on the device.  grow may be called repeatedly; every call restarts
 * each tree's RandomSource where sampling left it, so results repeat. */
int cf_forest_sample(cf_ctx* ctx, cf_images* images, const cf_train_config* config, int num_trees,
        int num_labels, cf_forest** out);
int cf_forest_grow(cf_ctx* ctx, cf_images* images, cf_forest* forest, const cf_train_config* config, int mode);
/* per-phase device timing of training (hipEvents; the reference's utils::Profile,
 * utils.h:159-203, and per-node timers, random_tree.h:634-679) */
int cf_ctx_set_profiling(cf_ctx* ctx, int enable);
/* out[0..15]: evaluations, wall seconds, response-kernel ms, response-kernel
 * launches, responses, feature-generation ms, score+argmax ms, partition ms,
 * levels, sequential-replay nodes, evaluated nodes, node-histogram ms (the
 * second pass of batched levels), the most feature-group batches and the most
 * node-slot batches of one batched level, the pass-1 launch shape used most
 * (cf_eval_shape_info), the batched levels evaluated as one captured graph
 * (CURFIL_LEVEL_GRAPH=0 launches them one by one).  The response kernel is the one that
 * gathers the integrals: k_eval_hist (direct path) or k_eval_bins (batched
 * levels).  Returns the number of entries. */
int cf_forest_profile(const cf_forest* forest, double* out, int n);

/* Launch shapes of the batched levels' response kernel and their tuner (no
 * counterpart in the reference, whose kernels have fixed shapes,
 * random_tree_image_gpu.cu:2034-2065, 2095-2108).  A shape: lanes per
 * workgroup, samples per lane, features whose gathers are issued together,
 * workgroups per CU.  Every shape gives the same histograms; the tuner times
 * them on the first launches of each (threshold width, level size class),
 * each shape's launch right after a launch of shape 0, and keeps the shape
 * with the lowest mean ratio to it (CURFIL_EVAL_SHAPE=i pins shape i,
 * CURFIL_TUNE=0 keeps shape 0, CURFIL_TUNE_ROUNDS=r sets the ratios per
 * shape, CURFIL_TUNE_FILE=path keeps the choices between runs). */
int cf_eval_shape_info(int shape, int* lanes, int* per_lane, int* features_per_step, int* workgroups_per_cu);
/* per tuned key, 6 values: threshold width, size class, chosen shape (-1 while
 * open), ns per response of shape 0 on its timed launches and of the chosen
 * shape estimated as that times the chosen shape's mean ratio to shape 0
 * (in that order: chosen, shape 0), launches timed; returns the number of
 * keys */
int cf_eval_tuning(double* out, int n);
int cf_eval_tuning_reset(void);
/* one tree from given root samples (RandomTreeImage::doTrain,
 * random_tree_image.cpp:714-730); rs_seed seeds the tree's RandomSource. */
int cf_tree_train_from_samples(cf_ctx* ctx, cf_images* images, const cf_train_config* config, int tree_id,
        int num_classes, int rs_seed, const cf_sample* roots, int n, int mode, cf_forest** out);
/* build a forest from node tables (e.g. imported trees); histograms [nodes][C] */
int cf_forest_create(cf_ctx* ctx, int num_trees, const int* num_nodes, const cf_tree_node* nodes,
        const uint32_t* histograms, int num_classes, const uint32_t* class_prior, cf_forest** out);
int cf_forest_destroy(cf_forest* forest);
int cf_forest_num_trees(const cf_forest* forest);
int cf_forest_num_classes(const cf_forest* forest);
int cf_forest_tree_size(const cf_forest* forest, int tree);
int cf_forest_get_tree(const cf_forest* forest, int tree, cf_tree_node* nodes, uint32_t* histograms);
int cf_forest_get_root_samples(const cf_forest* forest, int tree, cf_sample* out, int capacity);
int cf_forest_get_prior(const cf_forest* forest, uint32_t* prior);
/* the label prior of one tree (RandomTreeImage::getClassLabelPriorDistribution,
 * random_tree_image.h:1149-1151): differs from the forest's when the tree
 * trained on a subset of the images (maxImages) or was read from its own file.
 * Returns the number of entries. */
int cf_forest_get_tree_prior(const cf_forest* forest, int tree, uint32_t* prior);
/* training statistics of the last train call: evaluated feature-threshold
 * pairs, device seconds, per-level timing count */
int cf_forest_train_stats(const cf_forest* forest, double* evaluations, double* seconds);

/* RandomForestImage::normalizeHistograms (random_forest_image.cpp:366-376)
 * and upload of the packed forest for prediction. */
int cf_forest_normalize(cf_forest* forest, double histogram_bias, int use_labels_prior);
/* leaf posteriors [nodes][C] (double) of one tree, after normalize */
int cf_forest_get_posteriors(const cf_forest* forest, int tree, double* out);

/* ---- tree files ------------------------------------------------------------- */
/* one tree as JSON (every value a string, as Boost.PropertyTree writes it),
 * gzip-compressed when `path` ends in ".gz"; replaces RandomTreeExport::writeJSON
 * (export.cpp:180-313).  palette [C][3] gives the colour of each label in the
 * histogram keys "R,G,B (label)"; NULL writes "l,l,l (l)".  Floats are
 * written with enough digits to read back unchanged. */
int cf_forest_export_tree(cf_forest* forest, int tree, const char* path, const cf_train_config* config,
        const uint8_t* palette, const char* folder_training);
/* a forest from tree files (plain or gzip JSON); replaces
 * RandomForestImage(treeFiles, ...) (random_forest_image.cpp:47-107) with
 * RandomTreeImport::readJSON (import.cpp:96-265).  Every tree keeps its own
 * label prior for cf_forest_normalize.  config_out (may be NULL) receives the
 * first tree's configuration, palette (may be NULL, else [32][3]) the colour of
 * each label.  Fails with CF_ERR_ARG on a malformed file, a histogram whose
 * sum is below the node's sample count, or trees of different configurations
 * or class counts. */
int cf_forest_import(cf_ctx* ctx, const char* const* paths, int count, cf_train_config* config_out, uint8_t* palette,
        cf_forest** out);

/* ---- prediction ------------------------------------------------------------- */
/* RandomForestImage::predict (random_forest_image.cpp:258-352) with the CPU
 * branch's arithmetic (:307-345): probs [C][H][W] float, labels [H][W] uint8.
 * Either output may be NULL. */
int cf_predict(cf_ctx* ctx, cf_forest* forest, cf_images* images, int index, float* probs, uint8_t* labels);
/* several frames of one image set in one launch sequence; outputs stacked per frame */
int cf_predict_batch(cf_ctx* ctx, cf_forest* forest, cf_images* images, const int* indices, int count,
        float* probs, uint8_t* labels);

/* the same with outputs in device memory owned by the caller (d_probs
 * [count][C][H][W] float, d_labels [count][H][W] uint8; either may be NULL):
 * for pipelines whose next stage runs on the GPU.  Returns when the
 * outputs are written. */
int cf_predict_batch_device(cf_ctx* ctx, cf_forest* forest, cf_images* images, const int* indices, int count,
        float* d_probs, uint8_t* d_labels);

/* cf_predict_batch_device with no allocation at all, for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops); it has no
 * counterpart in the reference.  The batch's frame indices are staged in a
 * fixed device array of the library (65536 frames at most, 256 KiB per
 * device, held until the process ends) instead of a buffer that grows.  The
 * forest must already be resident on the device (any earlier prediction with
 * it uploads it), or CF_ERR_ARG.  Calls that overlap on one device (other
 * contexts, unsynchronised threads) race on that array. */
int cf_predict_batch_device_noalloc(cf_ctx* ctx, cf_forest* forest, cf_images* images, const int* indices,
        int count, float* d_probs, uint8_t* d_labels);

/* device memory for callers without an allocator of their own (for example
 * for cf_predict_batch_device's outputs); no counterpart in the reference */
int cf_device_alloc(cf_ctx* ctx, size_t bytes, void** ptr);
int cf_device_free(cf_ctx* ctx, void* ptr);
/* page-locked host memory (no counterpart in the reference): outputs of
 * cf_predict / cf_predict_batch written there are copied by DMA at the link's
 * rate instead of being staged through pageable memory */
int cf_host_alloc(cf_ctx* ctx, size_t bytes, void** ptr);
int cf_host_free(cf_ctx* ctx, void* ptr);
int cf_device_to_host(cf_ctx* ctx, void* host, const void* device, size_t bytes);

/* ---- serving loop ------------------------------------------------------------ */
/* RandomForestImage::predict (random_forest_image.cpp:258-352) as a service for
 * frames that arrive one at a time, at any resolution: a worker thread admits
 * submitted frames in arrival order while device memory of `budget_mb` is
 * free (integral images, staging and outputs of each frame count), builds
 * their integral images, classifies frames of equal size together (at most
 * `max_batch` per launch) and copies results back, with the three stages on
 * separate streams.  Frames submitted while others run join the next batch.
 * A frame's results equal cf_predict's for the same frame.  The forest must
 * stay unchanged while a server uses it.  (The reference classifies one frame
 * per call; this loop has no single counterpart there.) */
typedef struct cf_server cf_server;
int cf_server_create(cf_ctx* ctx, cf_forest* forest, double budget_mb, int max_batch, cf_server** out);
/* copies the frame (rgb [H][W][3] u8, depth [H][W] u16 mm); CF_ERR_NOMEM if one
 * frame of this size exceeds the budget */
int cf_server_submit(cf_server* server, const uint8_t* rgb, const uint16_t* depth_mm, int width, int height,
        uint64_t* id);
/* blocks until frame `id` is done; probs [C][H][W] and labels [H][W] may be
 * NULL; latency_ms[0..3] (may be NULL): completion (results in host memory)
 * - the cf_server_submit call (its input copy included), and its parts:
 * the call and waiting for admission, admission to the batch's launch,
 * launch to completion */
int cf_server_wait(cf_server* server, uint64_t id, float* probs, uint8_t* labels, double* latency_ms);
/* out[0..10]: frames done, pixels done, classify launches, device MB in use,
 * frames queued, worker passes (rounds of admission, launches and completion
 * scan, without the waits for work), the longest pass in ms, passes over
 * 1 ms, device slots allocated, page-locked host buffers allocated, events
 * created; returns the number of values the server has (11) */
int cf_server_stats(cf_server* server, double* out, int n);
const char* cf_server_error(cf_server* server);
int cf_server_destroy(cf_server* server);

/* ---- datasets (PNG files) --------------------------------------------------- */
/* The colour -> label map of the ground-truth images (the reference's
 * process-wide `colors` map with LabelImage::encodeColor / decodeLabel,
 * image.cpp:656-734): labels are given in the order colours are met.  Training
 * and test sets, and imported trees' colours, share one map. */
typedef struct cf_label_map cf_label_map;
int cf_label_map_create(cf_label_map** out);
int cf_label_map_destroy(cf_label_map* map);
/* the label of a colour, added as the next label if new */
int cf_label_map_encode(cf_label_map* map, const uint8_t* rgb, int* label);
/* the label of a colour, or -1 */
int cf_label_map_find(const cf_label_map* map, const uint8_t* rgb);
int cf_label_map_size(const cf_label_map* map);
/* colours by label [size][3]; returns the size */
int cf_label_map_colors(const cf_label_map* map, uint8_t* rgb);

/* loadImages (image.cpp:869-965) over the given `*_colors.png` files (the
 * caller lists them as listImageFilenames does, image.cpp:844-866): each with
 * its `*_depth.png` (16-bit grayscale mm; skipped, depth 1000 everywhere, when
 * use_depth_images is 0; the reference's checks of values and average depth)
 * and `*_ground_truth.png` (RGB colours -> labels through `map`, column by
 * column), depth filling when asked (image.cpp:279-377).  The padding colour
 * (the first ignored colour, else black) is encoded first, as getPaddingLabel
 * does; NULL encodes none (test images loaded one at a time, which are not
 * padded; the padding label is then -1).  Files are decoded on `threads` host threads (0 = 8); labels are
 * encoded in file order.  Errors (missing files, non-RGB colour images, a bad
 * depth image) fail the call; cf_last_error(NULL) has the message. */
typedef struct cf_dataset cf_dataset;
int cf_dataset_load(const char* const* color_files, int count, int use_depth_images, int use_depth_filling,
        int use_cielab, const uint8_t* padding_color, cf_label_map* map, int threads, cf_dataset** out);
int cf_dataset_destroy(cf_dataset* dataset);
/* number of images, the largest width and height, the padding label */
int cf_dataset_info(const cf_dataset* dataset, int* count, int* width, int* height, int* padding_label);
/* an image's own size */
int cf_dataset_image_size(const cf_dataset* dataset, int index, int* width, int* height);
/* an image at its own size: rgb [h][w][3], depth [h][w] (after filling), labels [h][w]; any may be NULL */
int cf_dataset_get(const cf_dataset* dataset, int index, uint8_t* rgb, uint16_t* depth_mm, uint8_t* labels);
/* the colour planes [3][h][w] as the forest sees them: RGB values, or CIELab
 * (VIGRA's RGB2LabFunctor<double> restated, image.cpp:89-96; parity unpinned)
 * when the set was loaded with use_cielab */
int cf_dataset_get_color(const cf_dataset* dataset, int index, float* color);
/* every image on the device at the largest size (smaller images padded as
 * loadImages pads them, labels padded with the padding label) */
int cf_dataset_to_images(const cf_dataset* dataset, cf_ctx* ctx, cf_images** out);

/* PNG files (VIGRA importImage / exportImage in the reference).  read_info:
 * size, channels (1 gray, 2 gray+alpha, 3 RGB incl. palette images, 4 RGBA)
 * and bit depth (8 or 16); read: samples [h][w][channels] as uint16. */
int cf_png_read_info(const char* path, int* width, int* height, int* channels, int* bit_depth);
int cf_png_read(const char* path, uint16_t* samples, size_t capacity);
/* 8-bit RGB (pixels [h][w][3] u8) or 8/16-bit grayscale ([h][w] u8 / u16) */
int cf_png_write(const char* path, int width, int height, int channels, int bit_depth, const void* pixels);

/* ---- prediction quality ------------------------------------------------------ */
/* calculatePixelAccuracy (predict.cpp:142-199) over n pixels: with include_void
 * every pixel counts, else pixels whose ground truth is an ignored label are
 * skipped; confusion [C][C] (may be NULL; rows: labels, columns: predictions)
 * is incremented for the pixels counted. */
int cf_pixel_accuracy(const uint8_t* prediction, const uint8_t* ground_truth, size_t n, int include_void,
        const int* ignored_labels, int num_ignored, double* confusion, int num_classes, double* accuracy);
/* ConfusionMatrix::normalize (predict.cpp:102-140): rows scaled to sum 1 (zero rows stay zero) */
int cf_confusion_normalize(double* confusion, int num_classes);
/* ConfusionMatrix::averageClassAccuracy (predict.cpp:71-100): mean diagonal of the
 * row-normalised matrix, without the ignored labels unless include_void */
int cf_confusion_class_accuracy(const double* confusion, int num_classes, int normalized, int include_void,
        const int* ignored_labels, int num_ignored, double* out);

/* ---- command-line tools --------------------------------------------------------- */
/* main() of curfil_train (train_main.cpp:40-187) and curfil_predict
 * (predict_main.cpp:59-166 with test(), predict.cpp:202-400): the same options,
 * positional order and defaults; log lines on stderr; returns the exit status.
 * The executables curfil_amd/bin/curfil_train and curfil_predict call these. */
int cf_tool_train(int argc, char** argv);
int cf_tool_predict(int argc, char** argv);

/* ---- synthetic workload (benchmark / tests; no counterpart in the reference) -- */
/* One labelled RGB-D frame: Voronoi labels (site k -> label k mod (classes+1),
 * 0 = void), per-class colour means from class_seed plus N(0, 20) noise,
 * per-cell tilted depth planes in [500, 5000] mm with a fraction zero_frac of
 * missing (0) pixels.  rgb [H][W][3], depth [H][W], labels [H][W]. */
int cf_synth_rgbd(uint32_t seed, uint32_t class_seed, int width, int height, int sites, int classes,
        double zero_frac, uint8_t* rgb, uint16_t* depth, uint8_t* labels);

#ifdef __cplusplus
}
#endif

#endif /* CURFIL_HIP_H */
