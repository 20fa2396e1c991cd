// Host-side training driver and forest handle.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "cf_internal.h"
#include "cf_rng.h"

namespace cf {

struct TreeBuild {
    std::vector<cf_tree_node> nodes;  // creation order
    std::vector<uint32_t> hist;       // [nodes][C]
};

struct TrainStats {
    double evaluations = 0;  // sample x feature x threshold evaluations
    double seconds = 0;      // wall time of the level loop (device work included)
    int levels = 0;
    int replays = 0;         // nodes resolved by the sequential argmax replay
    double nodes = 0;        // evaluated nodes
    double responses = 0;    // (sample, feature) responses binned by the histogram kernel
    // device time per phase (ms), filled when the context profiles
    double evalKernelMs = 0, evalLaunches = 0;
    double featGenMs = 0, scoreMs = 0, partitionMs = 0;
    double histMs = 0;   // two-pass evaluation: node histograms from the bin rows
    // two-pass evaluation: most feature-group batches and node-slot batches of one level
    int groupBatchesMax = 0, slotBatchesMax = 0;
    // two-pass evaluation: pass-1 launches per launch shape (see evalShape)
    double shapeLaunches[16] = {};
    // two-pass evaluation: levels evaluated as one captured graph
    int levelGraphs = 0;
};

// One tree to grow: its id, random source state and root samples.
struct TreeJob {
    int treeId;
    HostRandomSource rs;
    const std::vector<cf_sample>* roots;   // the tree's root samples, held by the forest (not copied)
};

// Level-wise training of several trees on one device (RandomTreeTrain::train,
// random_tree.h:1249-1415, for every tree at once).
void trainTrees(cf_ctx* ctx, cf_images* images, const cf_train_config& cfg, int numClasses,
        std::vector<TreeJob>& jobs, int mode, std::vector<TreeBuild>& out, TrainStats& stats);

// RandomTreeImage::calculateLabelPriorDistribution (random_tree_image.cpp:885-909)
// over the images `ids`, or over every image
std::vector<uint32_t> labelPrior(const cf_images* images, const std::vector<int>& ids, int numLabels);
std::vector<uint32_t> labelPrior(const cf_images* images, int numLabels);

// the images one tree trains on: a reservoir of maxImages images drawn with the
// tree's RandomSource, or all (random_forest_image.cpp:161-171)
std::vector<int> sampleImages(int count, int maxImages, HostRandomSource& rs);

// RandomTreeImage::subsampleTrainingDataClassUniform with one sampling chunk
// (random_tree_image.cpp:911-1059) followed by the root sort
// (random_tree_image.cpp:776-802, image index in place of the image address).
std::vector<cf_sample> sampleClassUniform(const cf_images* images, const std::vector<int>& ids,
        const std::vector<uint32_t>& prior, const cf_train_config& cfg, HostRandomSource& rs, size_t subsampleCount);

// RandomTreeImage::subsampleTrainingDataPixelUniform (random_tree_image.cpp:840-883)
// followed by the same root sort
std::vector<cf_sample> samplePixelUniform(const cf_images* images, const std::vector<int>& ids,
        const cf_train_config& cfg, HostRandomSource& rs, size_t subsampleCount);

}  // namespace cf

struct cf_forest {
    cf_ctx* ctx = nullptr;
    int C = 0;
    std::vector<cf::TreeBuild> trees;
    std::vector<std::vector<cf_sample>> roots;
    std::vector<int> treeIds;
    std::vector<int> rsSeeds;  // RandomSource seed counter of each tree after pixel sampling
    std::vector<uint32_t> prior;
    // label prior of each tree when the trees were trained apart (imported
    // tree files carry their own); empty: every tree uses `prior`
    std::vector<std::vector<uint32_t>> treePriors;
    std::vector<std::vector<double>> post;  // normalized histograms [nodes][C] per tree
    bool uploaded = false;
    cf::DevBuf<cf::NodeBlock> dBlocks;
    cf::DevBuf<double> dPost;
    cf::DevBuf<int32_t> dRoots;
    // prediction staging kept between calls (grown on demand)
    cf::DevBuf<int> dFrames;
    cf::DevBuf<float> dProbs;
    cf::DevBuf<uint8_t> dLabels;
    cf::TrainStats stats;
    // the uploaded forest as the prediction kernels read it
    cf::DevForest devForest() const {
        return cf::DevForest{dBlocks.get(), dPost.get(), dRoots.get(), static_cast<int>(trees.size()), C};
    }
};

namespace cf {
// the message of the last failed call made without a context, per host thread
std::string& hostErrorSlot();

// a decoded PNG: samples as stored (palette images expanded to RGB), row-major,
// interleaved; 8-bit images hold values 0..255
struct PngImage {
    int width = 0, height = 0, channels = 0, bitDepth = 0, colorType = 0;
    std::vector<uint16_t> samples;
};
PngImage readPng(const std::string& path);
// 8-bit RGB (pixels [H][W][3] u8) or 8/16-bit grayscale ([H][W] u8 / u16)
void writePng(const std::string& path, int width, int height, int channels, int bitDepth, const void* pixels);

// a forest made without a context (imported, or built from node tables) takes
// the context of its first device call; later calls must use the same one
void bindForest(cf_forest* fo, cf_ctx* ctx);
// packs a normalized forest into device node records (once; later calls return at once)
void uploadForest(cf_forest* fo);
// resolves the allocation-free prediction's device array on `device` (current), once per device
void resolvePredictScratch(int device);

// one tree as a JSON file, gzip-compressed when `path` ends in ".gz"
// (RandomTreeExport::writeJSON, export.cpp:180-313); palette [C][3] or null
void writeTreeFile(const cf_forest* fo, int tree, const char* path, const cf_train_config& cfg,
        const uint8_t* palette, const char* folderTraining);
// a forest from tree files, plain or gzip (RandomForestImage(treeFiles),
// random_forest_image.cpp:47-107 with RandomTreeImport::readJSON,
// import.cpp:96-265); palette receives [C][3] when not null
std::unique_ptr<cf_forest> readTreeFiles(cf_ctx* ctx, const char* const* paths, int count, cf_train_config* cfgOut,
        uint8_t* palette);
}  // namespace cf
