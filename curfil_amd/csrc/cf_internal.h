// Internal host-side declarations: handles, device buffers, kernel launchers.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "cf_common.h"

namespace cf {

// XCDs of an MI355X: the dispatcher sends workgroup i of a launch to XCD i % kNumXcds, each XCD with
// its own L2.
constexpr int kNumXcds = 8;

// Thrown inside the library, turned into a status at the C ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CF_HIP(call)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (call);                                                                        \
        if (e_ != hipSuccess)                                                                          \
            throw ::cf::Error(CF_ERR_HIP, std::string(#call) + ": " + hipGetErrorString(e_));          \
    } while (0)

// Owning device buffer.
template<class T>
class DevBuf {
public:
    DevBuf() = default;
    explicit DevBuf(size_t n) { resize(n); }
    ~DevBuf() { release(); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : p_(o.p_), n_(o.n_), cap_(o.cap_) { o.p_ = nullptr; o.n_ = o.cap_ = 0; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        if (this != &o) {
            release();
            p_ = o.p_; n_ = o.n_; cap_ = o.cap_;
            o.p_ = nullptr; o.n_ = o.cap_ = 0;
        }
        return *this;
    }
    // grows (never shrinks) the allocation; contents are not preserved
    void resize(size_t n) {
        if (n > cap_) {
            release();
            if (n > 0) {
                hipError_t e = hipMalloc(reinterpret_cast<void**>(&p_), n * sizeof(T));
                if (e != hipSuccess) {
                    p_ = nullptr;
                    throw Error(e == hipErrorOutOfMemory ? CF_ERR_NOMEM : CF_ERR_HIP,
                            std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) + " bytes: " +
                            hipGetErrorString(e));
                }
            }
            cap_ = n;
        }
        n_ = n;
    }
    void release() {
        if (p_) (void)hipFree(p_);
        p_ = nullptr;
        n_ = cap_ = 0;
    }
    T* get() const { return p_; }
    size_t size() const { return n_; }
    size_t bytes() const { return n_ * sizeof(T); }

private:
    T* p_ = nullptr;
    size_t n_ = 0, cap_ = 0;
};

// Uniform-size image set: image i's planes start at col + i*3*W*H and dep + i*W*H.
struct ImageSet {
    const float* col;
    const int2* dep;
    int W, H;
};

}  // namespace cf

struct cf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    // batched prediction to host memory: one event per launch, whose results are copied while the next runs,
    // and one per copy into the page-locked staging (two halves) for outputs in pageable memory
    std::vector<hipEvent_t> launchDone, copyDone;
    void* hostStage = nullptr;
    size_t hostStageBytes = 0;
    std::string error;
    std::recursive_mutex mu;
    bool profiling = false;
    // scratch buffers reused across calls
    cf::DevBuf<uint8_t> scratch0, scratch1, scratch2;
};

struct cf_images {
    cf_ctx* ctx = nullptr;
    int count = 0, W = 0, H = 0;
    cf::DevBuf<float> col;         // [count][3][H][W]
    cf::DevBuf<int2> dep;          // [count][H][W]
    cf::DevBuf<cf::DevImage> desc; // [count]
    std::vector<uint8_t> labels;   // host [count][H][W]
    std::vector<uint8_t> labelsSet;
    std::vector<int32_t> depthHost; // host raw depth mm [count][H][W] (pixel validity for sampling)
    std::vector<uint8_t> depthSet;
    cf::ImageSet set() const;
};

namespace cf {

// ---- launchers (implemented in the .hip files) ----
void launchIntegralU8(hipStream_t s, const uint8_t* rgb, const uint16_t* depth, int W, int H, uint32_t* tmp,
        float* col, int2* dep);
void launchIntegralF32(hipStream_t s, const float* color, const int32_t* depth, int W, int H, uint32_t* tmp,
        double* work, float* col, int2* dep);
// integral of a w x h image (srcCol [3][h][w], srcDep [h][w]) into a W x H slot, NaN / 0 outside
void launchIntegralPad(hipStream_t s, const float* srcCol, const int2* srcDep, int w, int h, int W, int H, float* col,
        int2* dep);

// Feature generation job of one (tree, level).
struct FeatGenJob {
    const cf_sample* samples; // the level's samples of this tree, frontier order
    int n;                    // number of samples
    int seed;                 // seed of the level's RandomSource
    int outOffset;            // feature slot offset (tree * F)
    int pad;
};

// Per-(tree, level) feature tables on the device, slot = tree * F + f.
struct FeatTables {
    DevBuf<FeatRec> feat;      // sorted features [trees*F]
    DevBuf<float> thrFT;       // thresholds feature-major [trees*F][T]
    DevBuf<float> thrTF;       // thresholds threshold-major [trees][T][F] (reference layout)
    DevBuf<uint16_t> rank;     // rank[tf*T + t] = #{k : thr[k] < thr[t]} within the feature
    DevBuf<uint16_t> order;    // thresholds of each feature in non-decreasing rank order
    DevBuf<int8_t> soa;        // [trees][11][F] (reference SoA layout)
    // candidate / unsorted staging
    DevBuf<FeatRec> cand;
    DevBuf<FeatRec> unsortedFeat;
    DevBuf<float> unsortedThr;
    DevBuf<int> unsortedKey;
    DevBuf<int> err;
    int trees = 0, F = 0, T = 0;
};

void launchFeatureGeneration(hipStream_t s, const ImageSet& imgs, const FeatGenJob* jobsDev, int numJobs,
        int F, int T, int boxRadius, int regionSize, int useDepth, const int* seedsHost, FeatTables& tabs);
// thresholds given in reference layout [trees][T][F] + features SoA [trees][11][F] -> tables
void launchPrepareTables(hipStream_t s, int trees, int F, int T, FeatTables& tabs);

// Evaluation work item: a chunk of one node's samples.
struct Chunk {
    int slot;   // node slot within the launch batch
    int tree;   // tree index (feature table)
    int begin;  // first sample (index into the level sample array)
    int count;  // samples in this chunk
};

struct NodeBest {
    double score;
    int feat;     // sorted feature index
    int thresh;   // threshold index
    int flag;     // 1: near-tie band not clean -> sequential replay needed
    int pad;
};

struct FeatSummary {
    double best;   // max score over thresholds
    double second; // largest score strictly below best (-inf if none)
    int thresh;    // smallest threshold index reaching best
    int pad;
};

// histogram bins G[slot][F][T+1][C] (interval histograms)
void launchEvalHistograms(hipStream_t s, const ImageSet& imgs, const cf_sample* samples, const Chunk* chunks,
        int numChunks, const FeatTables& tabs, int C, uint32_t* G);

// ---- two-pass level evaluation (k_bins.hip)
// LDS a workgroup may use on the current device (MI355X: 160 KB); kernels whose
// histograms need more than the default 64 KB opt in before launch
size_t deviceLdsLimit();
void allowLds(const void* kernel, size_t bytes);

constexpr int kBinGroup = 16;            // features per bin row (16 x 8-bit bins = one uint4 per sample)
constexpr int kMaxBinThresholds = 255;   // bins 0..T must fit in 8 bits
// a block of consecutive entries of the image-major permutation (at most a pass-1 workgroup's samples),
// all of one (tree, image)
struct BinBlock {
    int tree;
    int begin;   // first position in the permutation
    int count;
    int image;   // the image of every sample in the block
};
void launchFeatureEvalOrder(hipStream_t s, const FeatTables& tabs, int trees, uint16_t* evalOrder);
// keys ((image * keyMul + keyBase) << yBits | y) and values (level index) of one tree's level segment;
// with flipN > 0 also the flipped evaluation of every Both sample at flipN + index (value |
// kFlipEntry), and segment absentKey for the other samples there
constexpr uint32_t kFlipEntry = 0x80000000u;
void launchLevelKeys(hipStream_t s, const cf_sample* samples, int begin, int n, uint32_t keyBase, uint32_t keyMul,
        uint32_t* keys, uint32_t* vals, int flipN, uint32_t absentKey, int yBits);
// label per level index, bit 7 set for samples still set to Both (counted in both orientations)
void launchLevelLabels(hipStream_t s, const cf_sample* samples, int n, uint8_t* labels);
// Counting sort of n level entries by key (< levelSortBuckets(numSegs, yBits)): out = the values in key
// order (input order inside every segment bucket of up to 64 entries), starts[seg] = first position of segment seg (key >> yBits) or -1 if it is empty, and
// starts[numSegs] = the number of entries with a key below numSegs << yBits.  count and cursor hold
// levelSortBuckets entries, blockSum levelSortScanBlocks.
size_t levelSortBuckets(int numSegs, int yBits);
// p[0 .. n) = v by a kernel (usable inside a captured graph, where memset nodes are avoided)
void launchFillU32(hipStream_t s, uint32_t* p, size_t n, uint32_t v);
size_t levelSortScanBlocks(int numSegs, int yBits);
void levelCountingSort(hipStream_t s, const uint32_t* keys, const uint32_t* vals, int n, int numSegs, int yBits,
        uint32_t* count, uint32_t* cursor, uint32_t* blockSum, uint32_t* out, int* starts);
// Work list of pass 1 for ngb feature groups, built on the device: segment i (its blocks
// segFirst[i] .. + segCount[i]) belongs to XCD i % 8; XCD x takes its segments in order, all groups
// of a segment (group major, block minor) before the next, and workgroup w of the launch runs on
// XCD w % 8, so work[pos * 8 + x] is XCD x's item pos, (-1, 0) past its end.  xcdPrefix[i] counts
// the blocks of XCD x's segments before segment i; xcdBlocks[x] all of them; maxLen is the longest
// XCD list in items.
void launchBuildWork(hipStream_t s, const int* segFirst, const int* segCount, const int* xcdPrefix,
        const int* xcdBlocks, int numSegs, int ngb, int maxLen, int2* work);
// bins[g][level index] for the work items (block, g) of feature groups g0 + g of the evaluation order;
// permutation entries with kFlipEntry are evaluated flipped into bins[flipOffset + g * N + index]
// with launch shape `shape` (evalShape): blocks of shape.lanes * shape.perLane samples
void launchEvalBins(hipStream_t s, const ImageSet& imgs, const cf_sample* samples, const uint32_t* perm,
        const BinBlock* blocks, const int2* work, int numWork, int g0, const FeatRec* gFeat, const float* gThr,
        const FeatTables& tabs, int N, uint4* bins, size_t flipOffset, int shape);
// A launch shape of pass 1: lanes per workgroup, samples per lane (a workgroup's block holds
// lanes * perLane samples), features whose corner gathers are issued together, workgroups per CU (held
// by an LDS reservation).
struct EvalShape {
    int lanes, perLane, featPerStep, wgPerCu;
};
int numEvalShapes();
const EvalShape& evalShape(int i);
// Pass-1 shape tuner: the shape for a launch of threshold width TM at a level of `entries`
// permutation entries.  *measure asks the caller to time the launch and report it with
// evalTunerRecord.  CURFIL_EVAL_SHAPE=i pins shape i, CURFIL_TUNE=0 keeps shape 0,
// CURFIL_TUNE_FILE keeps the choices between runs.
int evalTunerChoose(int TM, size_t entries, bool* measure);
void evalTunerRecord(int TM, size_t entries, int shape, double ms, double responses);
// padded threshold width of the bin kernel for T thresholds
int binThresholdWidth(int T);
// features and +inf-padded thresholds in evaluation order: gFeat[trees * F], gThr[trees * F * binThresholdWidth(T)]
void launchBinTables(hipStream_t s, const FeatTables& tabs, int trees, const uint16_t* evalOrder, FeatRec* gFeat,
        float* gThr);
// G[slot][ng * 16][T + 1][C] from the bin rows of the chunks' samples
// (samples still set to Both also count their flipped row at flipOffset)
void launchHistBins(hipStream_t s, const uint4* bins, const uint8_t* labels, const Chunk* chunks, int numChunks,
        int g0, int ng, const FeatTables& tabs, int C, int N, uint32_t* G, size_t flipOffset = 0);
// Scores and per-feature summaries of the slots' histograms.  With allScores, also every score:
// allScores[T][F] of slot onlySlot alone, or allScores[slot][T][F] of all slots (onlySlot < 0).
void launchScores(hipStream_t s, const uint32_t* G, const uint32_t* nodeHist, const int* slotTree, int numSlots,
        const FeatTables& tabs, int C, FeatSummary* summ, double* allScores, int onlySlot, int fOffset = 0,
        int Fb = -1, const uint16_t* evalOrder = nullptr);
void launchScoresFromCounters(hipStream_t s, const uint32_t* counters, const uint32_t* hist, int F, int T, int C,
        double* scores);
void launchReduce(hipStream_t s, const FeatSummary* summ, int numSlots, int F, NodeBest* best);
// The reference's sequential argmax of scores[node][T][F] into best[node], for numNodes nodes.
void launchSequentialArgmax(hipStream_t s, const double* scores, int T, int F, NodeBest* best, int numNodes = 1);
void launchCountersFromBins(hipStream_t s, const uint32_t* G, const FeatTables& tabs, int C, uint32_t* counters);
void launchResponses(hipStream_t s, const DevImage* imgs, const cf_sample* samples, int n, const FeatRec* feats,
        int F, double* out);
void launchResponsesPaired(hipStream_t s, const DevImage* imgs, const cf_sample* samples, int n, const FeatRec* feats,
        double* out);

// Partition
// split record of node i from best[i] and the feature table of tree nodeTree[i]
void launchSplitRecords(hipStream_t s, const NodeBest* best, const int* nodeTree, int numNodes, const FeatTables& tabs,
        SplitRec* out);
void launchPartitionDirections(hipStream_t s, const DevImage* imgs, const cf_sample* samples, const Chunk* chunks,
        int numChunks, const SplitRec* splits, int C, uint8_t* dir, int2* chunkCounts, uint32_t* leftHist);
void launchPartitionScatter(hipStream_t s, const cf_sample* samples, const Chunk* chunks, int numChunks,
        const uint8_t* dir, const int2* dst, cf_sample* out);

// Prediction
//
// The split nodes of a forest, packed in 128-byte blocks (one cache line): a block holds a subtree of up to
// three split levels, its nodes at heap positions 0..6 (position p's children at 2p+1 and 2p+2).  Slots 0..6
// are node records {threshold bits, offsets o1x o1y o2x o2y, regions r1x r1y r2x r2y, type c1 c2 -} (signed
// bytes, as in FeatRec); slot 7 is the block's tail {first child block, first leaf, leaf mask, child-block
// mask | first exit position << 16}.  A child position q below the first exit position (7, or 1 / 3 for a
// tree's top block of one or two levels) holds a split of the block; at or past it, a child that is a
// split starts another block.  Bit q of the leaf mask marks a leaf child, numbered from the first leaf in
// position order; bit q of the child-block mask a child block, numbered from the first child block in
// position order.
struct NodeBlock {
    uint4 slot[8];
};
// doubles per posterior row of a packed forest: C rounded up to even, so rows are 16-byte aligned
__host__ __device__ constexpr int postStride(int C) { return (C + 1) & ~1; }
struct DevForest {
    const NodeBlock* blocks;    // all trees' blocks, each tree's breadth-first from its root block
    const double* post;         // [leaves][postStride(C)], a block's leaves together
    const int32_t* treeRoot;    // root block of each tree, or -(leaf index)-1 for a one-node tree
    int numTrees;
    int C;
};
// probs [frames][C][H][W] and labels [frames][H][W] may be null
void launchClassify(hipStream_t s, const DevImage* imgs, const int* frameIdx, int numFrames, int W, int H,
        const DevForest& forest, float* probs, uint8_t* labels);
// one frame of a serving batch: its integrals and its outputs ([C][H][W] posteriors, [H][W] labels)
// Frames of any sizes share a launch: job i classifies rows rowBegin .. rowEnd - 1 of its frame (its
// whole integral image is there) with the tiles firstTile .. (job i + 1's firstTile) - 1 of the launch,
// firstTile ascending from 0 in job order (set by planClassifyJobs).
struct FrameJob {
    DevImage im;
    float* probs;
    uint8_t* labels;
    int firstTile;
    int rowBegin, rowEnd;
    int pad;
};
// the shape of a launch over jobs: the tree-split kernel for few pixels, and the launch's tiles
struct ClassifyPlan {
    bool split = false;
    int totalTiles = 0;
};
// chooses the shape by the jobs' pixels (the tree-split kernel up to splitMaxPixels) and sets their firstTile
// (host copies of the jobs)
ClassifyPlan planClassifyJobs(FrameJob* jobs, int numJobs, const DevForest& forest, double splitMaxPixels);
// the bound below which launches split the trees (CURFIL_CLASSIFY_SPLIT_MPIX or the default)
double classifySplitPixels();
void launchClassifyJobs(hipStream_t s, const FrameJob* jobs, int numJobs, const ClassifyPlan& plan,
        const DevForest& forest);

}  // namespace cf
