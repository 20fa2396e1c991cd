// Dense per-pixel forest prediction with the CPU branch's arithmetic of
// RandomForestImage::predict (random_forest_image.cpp:307-345):
//   * pixel depth by inclusion-exclusion on the integrals, invalid -> NaN
//     responses (PixelInstance, random_tree_image.h:140-165); every split
//     sends NaN right (SplitFunction::split, random_tree.h:122-144);
//   * per tree, the leaf's double posterior is added into a float
//     accumulator; the double sum normalises; the label is the first class
//     with the largest probability (strict >, starting from -1).
// One thread per pixel, a wave per 16 x 4 pixel tile; one launch classifies a
// whole batch of frames against all trees, each XCD taking a contiguous run
// of the batch's frames.  The
// forest holds split nodes only, in 128-byte blocks of up to three levels
// (NodeBlock): a walk reads one line per three levels, and ends at the last
// split with the leaf's number from its block, without a leaf record.
#include <cstdlib>
#include <type_traits>

#include "cf_internal.h"

namespace cf {

// Loads and stores through address-space-qualified pointers.  The image planes, the forest and a serving
// job's outputs are reached through pointers read from memory (frame descriptors, the forest's fields,
// the jobs): through generic pointers their accesses are flat instructions, which count on both wait
// counters and take a 64-bit address per lane; through global ones a plane's gathers take its base in
// scalar registers and a 32-bit offset per lane (measured 2.5% faster on 64-frame batches of 640x480).
// Vector values are loaded as the compiler's own vector types: the HIP vector classes' copy constructors
// take generic references, through which the address space is lost.
#define CF_GLOBAL __attribute__((address_space(1)))
typedef int i32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef double f64x2 __attribute__((ext_vector_type(2)));

// element i of a depth integral plane / a colour integral plane (a plane is far below 4 GiB)
__device__ __forceinline__ int2 depthAt(const int2* plane, int i) {
    const uint32_t off = static_cast<uint32_t>(i) * 8u;
    const i32x2 v = *(const CF_GLOBAL i32x2*)((const CF_GLOBAL char*)plane + off);
    return make_int2(v.x, v.y);
}
__device__ __forceinline__ float colorAt(const float* plane, int i) {
    const uint32_t off = static_cast<uint32_t>(i) * 4u;
    return *(const CF_GLOBAL float*)((const CF_GLOBAL char*)plane + off);
}
// record s of block b of the forest
__device__ __forceinline__ uint4 blockSlot(const NodeBlock* blocks, int b, int s) {
    const u32x4 v = ((const CF_GLOBAL u32x4*)blocks)[static_cast<size_t>(b) * 8 + s];
    return make_uint4(v.x, v.y, v.z, v.w);
}
// the k-th pair of doubles from row (16-byte aligned)
__device__ __forceinline__ double2 pairAt(const double* row, int k) {
    const f64x2 v = ((const CF_GLOBAL f64x2*)row)[k];
    return make_double2(v.x, v.y);
}

// A block's node record {threshold bits, o1x o1y o2x o2y, r1x r1y r2x r2y, type c1 c2 -} as a feature.
__device__ __forceinline__ Feat unpackNode(const uint4& w) {
    Feat f;
    f.o1x = sbyte(w.y, 0);
    f.o1y = sbyte(w.y, 1);
    f.o2x = sbyte(w.y, 2);
    f.o2y = sbyte(w.y, 3);
    f.r1x = sbyte(w.z, 0);
    f.r1y = sbyte(w.z, 1);
    f.r2x = sbyte(w.z, 2);
    f.r2y = sbyte(w.z, 3);
    f.type = sbyte(w.w, 0);
    f.c1 = static_cast<int>((w.w >> 8) & 0xffu);
    f.c2 = static_cast<int>((w.w >> 16) & 0xffu);
    return f;
}

// PixelInstance depth (random_tree_image.h:140-165) by inclusion-exclusion on the integral: true and
// the depth in mm when the pixel's depth is valid.
__device__ __forceinline__ bool pixelDepthAt(const DevImage& im, int x, int y, int32_t& depth) {
    const int W = im.W;
    auto at = [&](int xx, int yy) -> int2 { return depthAt(im.dep, yy * W + xx); };
    const int2 c00 = at(x, y);
    const int2 cl = x > 0 ? at(x - 1, y) : make_int2(0, 0);
    const int2 ca = y > 0 ? at(x, y - 1) : make_int2(0, 0);
    const int2 cal = (x > 0 && y > 0) ? at(x - 1, y - 1) : make_int2(0, 0);
    const int valid = static_cast<int>(static_cast<uint32_t>(c00.y) -
            (static_cast<uint32_t>(cl.y) + static_cast<uint32_t>(ca.y) - static_cast<uint32_t>(cal.y)));
    depth = -1;
    if (valid == 1) {
        depth = static_cast<int32_t>(static_cast<uint32_t>(c00.x) -
                (static_cast<uint32_t>(cl.x) + static_cast<uint32_t>(ca.x) - static_cast<uint32_t>(cal.x)));
    }
    return depth > 0;
}

// The leaf index one tree sends the pixel to, from the tree's root reference.  Per split the 8 corner
// gathers are issued before the response is formed; SplitFunction::split: left iff r <= threshold, NaN
// (invalid depth, a box outside the image, a depth box without valid pixels) goes right.  A block's tail
// is loaded with its first record (one line) and kept while the walk stays in the block.
__device__ __forceinline__ int walkTree(const DevImage& im, int x, int y, bool depthOk, double rinv,
        const DevForest& forest, int root) {
    if (root < 0) return -root - 1;
    const int W = im.W, H = im.H;
    const size_t plane = static_cast<size_t>(W) * H;
    int b = root;
    uint4 tail = blockSlot(forest.blocks, b, 7);
    int p = 0;
    for (;;) {
        const uint4 rec = blockSlot(forest.blocks, b, p);
        // the record as one 16-byte load: the compiler would split off the offsets and regions into the
        // branch that uses them, and every load instruction costs a tag lookup per line its lanes touch
        __asm__ volatile("" : : "v"(rec.x), "v"(rec.y), "v"(rec.z), "v"(rec.w));
        const Feat f = unpackNode(rec);
        int idx[8];
        const bool isDepth = f.type == T_DEPTH;
        const bool ok = depthOk && responseCorners(f, x, y, W, H, rinv, false, idx);
        // one 8-byte load per depth corner and one 4-byte load per colour corner, in separate branches
        // (a shared corner array makes the compiler split the depth loads in two)
        double r = cf_nan();
        if (ok && isDepth) {
            int2 dv[8];
#pragma unroll
            for (int c = 0; c < 8; c++) dv[c] = depthAt(im.dep, idx[c]);
            // all 8 corners are needed here: the compiler may not split a corner's two halves and sink the
            // sums behind the count check (that doubles the loads and serialises them)
#pragma unroll
            for (int c = 0; c < 8; c++) __asm__ volatile("" : : "v"(dv[c].x), "v"(dv[c].y));
            r = depthResponseFromCorners(dv);
        } else if (ok) {
            const float* colA = im.col + static_cast<size_t>(f.c1) * plane;
            const float* colB = im.col + static_cast<size_t>(f.c2) * plane;
            float cv[8];
#pragma unroll
            for (int c = 0; c < 4; c++) cv[c] = colorAt(colA, idx[c]);
#pragma unroll
            for (int c = 4; c < 8; c++) cv[c] = colorAt(colB, idx[c]);
            r = colorResponseFromCorners(cv);
        }
        const float thr = __uint_as_float(rec.x);
        const int q = 2 * p + ((r <= static_cast<double>(thr)) ? 1 : 2);
        const uint32_t below = (1u << q) - 1u;
        if ((tail.z >> q) & 1u) return static_cast<int>(tail.y + __popc(tail.z & below));
        if (q < static_cast<int>(tail.w >> 16)) {
            p = q;
            continue;
        }
        b = static_cast<int>(tail.x + __popc(tail.w & 0xffffu & below));
        tail = blockSlot(forest.blocks, b, 7);
        p = 0;
    }
}

// The double sum of the float accumulator normalises it; the label is the first class with the
// largest probability (strict >, from -1).
template<int CM>
__device__ __forceinline__ void finishPixel(const float (&acc)[CM], int C, size_t plane, int px,
        float* __restrict__ pout, uint8_t* __restrict__ lout) {
    double sum = 0.0;
#pragma unroll
    for (int c = 0; c < CM; c++) {
        if (c < C) sum += static_cast<double>(acc[c]);
    }
    float bestProb = -1.0f;
    int label = 0;
#pragma unroll
    for (int c = 0; c < CM; c++) {
        if (c < C) {
            const float p = static_cast<float>(static_cast<double>(acc[c]) / sum);
            if (pout) ((CF_GLOBAL float*)pout)[c * plane + px] = p;
            if (p > bestProb) {
                label = c;
                bestProb = p;
            }
        }
    }
    if (lout) ((CF_GLOBAL uint8_t*)lout)[px] = static_cast<uint8_t>(label);
}

// Adds the posterior rows of leaves[0 .. n - 1] (n <= K, the trees in order) into the float
// accumulator.  A row (postStride(C) doubles, 16-byte aligned) is read in 16-byte pairs; all pairs of the
// K rows go out before the first is summed; pairs past a row repeat its last pair and rows past n repeat
// row 0 (in bounds, unused).
template<int CM, int K>
__device__ __forceinline__ void addLeaves(float (&acc)[CM], const DevForest& forest, const int* leaves, int n) {
    const int C = forest.C;
    double2 pair[K][CM / 2];
#pragma unroll
    for (int j = 0; j < K; j++) {
        const double* row = forest.post + static_cast<size_t>(leaves[j < n ? j : 0]) * postStride(C);
#pragma unroll
        for (int k = 0; k < CM / 2; k++) pair[j][k] = pairAt(row, k < (C + 1) / 2 ? k : (C + 1) / 2 - 1);
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
#pragma unroll
        for (int k = 0; k < CM / 2; k++) __asm__ volatile("" : : "v"(pair[j][k].x), "v"(pair[j][k].y));
    }
#pragma unroll
    for (int j = 0; j < K; j++) {
        if (j < n) {
#pragma unroll
            for (int c = 0; c < CM; c++) {
                const double v = (c & 1) ? pair[j][c / 2].y : pair[j][c / 2].x;
                if (c < C) acc[c] = static_cast<float>(static_cast<double>(acc[c]) + v);
            }
        }
    }
}

// One thread per pixel, the trees one after another (walking several trees in lockstep per thread
// was measured slower): per tree, the leaf's double posterior is added into the float accumulator.
template<int CM>
__device__ __forceinline__ void classifyPixel(const DevImage& im, int px, const DevForest& forest,
        float* __restrict__ pout, uint8_t* __restrict__ lout) {
    const int W = im.W, H = im.H;
    const int x = px % W, y = px / W;
    const int C = forest.C;
    const size_t plane = static_cast<size_t>(W) * H;
    int32_t depth;
    const bool depthOk = pixelDepthAt(im, x, y, depth);
    const double rinv = depthOk ? depthReciprocal(depthToFloat(depth)) : 0.0;
    float acc[CM];
#pragma unroll
    for (int c = 0; c < CM; c++) acc[c] = 0.0f;
    for (int t = 0; t < forest.numTrees; t++) {
        const int leaf = walkTree(im, x, y, depthOk, rinv, forest, forest.treeRoot[t]);
        addLeaves<CM, 1>(acc, forest, &leaf, 1);
    }
    finishPixel<CM>(acc, C, plane, px, pout, lout);
}

// The (frame, pixel block) work of a launch, frame-major, is cut into 8 contiguous ranges, one per
// XCD (workgroup i runs on XCD i % 8): an XCD works through its frames one at a time and a frame's
// blocks top to bottom, so its L2 holds the integral rows that its resident pixels gather from instead
// of those of every frame in flight.  False for the padding workgroups past the work.
struct ClassifyTile {
    int frame, block;
};
__device__ __forceinline__ bool classifyTile(int blocksPerFrame, int numFrames, ClassifyTile& t) {
    const int total = blocksPerFrame * numFrames;
    const int per = (total + kNumXcds - 1) / kNumXcds;
    const int j = static_cast<int>(blockIdx.x % kNumXcds) * per + static_cast<int>(blockIdx.x / kNumXcds);
    if (j >= total) return false;
    t.frame = j / blocksPerFrame;
    t.block = j - t.frame * blocksPerFrame;
    return true;
}

// A wave classifies a tile of 16 x 4 pixels, a workgroup of 4 waves 16 x 16: neighbouring pixels take
// the same branches longer than along a row, so a wave's gathers and node reads touch fewer lines
// (measured on 64-frame batches of 640x480 with 10 trees of depth 20: 16 x 4 3778 frames/s, 32 x 2 and
// 8 x 8 3602, 64 x 1 3420).
constexpr int kTileW = 16;
constexpr int kTileH = 64 / kTileW;

__host__ __device__ __forceinline__ int tileBlocks(int W, int H) {
    return ((W + kTileW - 1) / kTileW) * ((H + 4 * kTileH - 1) / (4 * kTileH));
}
// the pixel of this lane in tile `block` of the rows rowBegin .. rowEnd - 1 of a frame W wide, or -1
// outside them
__device__ __forceinline__ int tilePixel(int block, int W, int rowBegin, int rowEnd) {
    constexpr int TW = kTileW, TH = kTileH;
    const int tilesX = (W + TW - 1) / TW;
    const int bx = block % tilesX, by = block / tilesX;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = bx * TW + lane % TW, y = rowBegin + (by * 4 + wave) * TH + lane / TW;
    return (x < W && y < rowEnd) ? y * W + x : -1;
}

inline dim3 classifyGrid(int blocksPerFrame, int numFrames) {
    const int64_t total = static_cast<int64_t>(blocksPerFrame) * numFrames;
    const int64_t per = (total + kNumXcds - 1) / kNumXcds;
    if (per * kNumXcds > INT32_MAX) throw Error(CF_ERR_UNSUPPORTED, "too many pixels in one classify launch");
    return dim3(static_cast<unsigned>(per * kNumXcds));
}

// frames of one image set: frame f of the batch is image frameIdx[f]; outputs stacked per frame
template<int CM>
__global__ void __launch_bounds__(256) k_classify(const DevImage* __restrict__ imgs, const int* __restrict__ frameIdx,
        int numFrames, int W, int H, DevForest forest, float* __restrict__ probs, uint8_t* __restrict__ labels) {
    ClassifyTile tile;
    if (!classifyTile(tileBlocks(W, H), numFrames, tile)) return;
    const int frame = tile.frame;
    const int px = tilePixel(tile.block, W, 0, H);
    if (px < 0) return;
    const size_t plane = static_cast<size_t>(W) * H;
    classifyPixel<CM>(imgs[frameIdx[frame]], px, forest,
            probs ? probs + static_cast<size_t>(frame) * forest.C * plane : nullptr,
            labels ? labels + static_cast<size_t>(frame) * plane : nullptr);
}

// the job of tile j of a launch: the last job whose first tile is <= j
__device__ __forceinline__ int findJob(const FrameJob* __restrict__ jobs, int numJobs, int j) {
    int lo = 0, hi = numJobs - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (jobs[mid].firstTile <= j) {
            lo = mid;
        } else {
            hi = mid - 1;
        }
    }
    return lo;
}

// independent frames of any sizes (or row bands of them), each with its own images and outputs; the
// launch's tiles are cut into XCD ranges as a batch's are, and a workgroup finds its job by binary search
// over the jobs' first tiles
template<int CM>
__global__ void __launch_bounds__(256) k_classify_jobs(const FrameJob* __restrict__ jobs, int numJobs, int totalTiles,
        DevForest forest) {
    ClassifyTile tile;
    if (!classifyTile(totalTiles, 1, tile)) return;
    const int j = tile.block;
    const FrameJob job = jobs[findJob(jobs, numJobs, j)];
    const int px = tilePixel(j - job.firstTile, job.im.W, job.rowBegin, job.rowEnd);
    if (px < 0) return;
    classifyPixel<CM>(job.im, px, forest, job.probs, job.labels);
}

// Few pixels (a single frame, a few small frames): with one thread per pixel most SIMDs hold a single
// wave, and a wave's time is its walk of all trees one after another.  Here a workgroup takes one 16 x 4
// tile and its waves walk the trees t = w, w + kSplitWaves, ... (w the wave); the leaf numbers meet in
// LDS and wave 0 adds the posteriors in tree order, the same float sums as one thread walking all trees.
constexpr int kSplitWaves = 4;

__host__ __device__ __forceinline__ int splitTiles(int W, int rows) {
    return ((W + kTileW - 1) / kTileW) * ((rows + kTileH - 1) / kTileH);
}

template<int CM>
__device__ __forceinline__ void classifyTileSplit(const DevImage& im, int tile, int rowBegin, int rowEnd,
        const DevForest& forest, float* __restrict__ pout, uint8_t* __restrict__ lout) {
    extern __shared__ int leafLds[];   // [numTrees][64]
    const int W = im.W;
    const int tilesX = (W + kTileW - 1) / kTileW;
    const int bx = tile % tilesX, by = tile / tilesX;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int x = bx * kTileW + lane % kTileW, y = rowBegin + by * kTileH + lane / kTileW;
    const bool in = x < W && y < rowEnd;
    const int T = forest.numTrees;
    if (in) {
        int32_t depth;
        const bool depthOk = pixelDepthAt(im, x, y, depth);
        const double rinv = depthOk ? depthReciprocal(depthToFloat(depth)) : 0.0;
        for (int t = wave; t < T; t += kSplitWaves) {
            leafLds[t * 64 + lane] = walkTree(im, x, y, depthOk, rinv, forest, forest.treeRoot[t]);
        }
    }
    __syncthreads();
    if (wave != 0 || !in) return;
    constexpr int K = CM / 2 >= 8 ? 1 : 8 / (CM / 2);   // leaf rows read together
    float acc[CM];
#pragma unroll
    for (int c = 0; c < CM; c++) acc[c] = 0.0f;
    for (int t0 = 0; t0 < T; t0 += K) {
        int leaves[K];
#pragma unroll
        for (int j = 0; j < K; j++) leaves[j] = leafLds[(t0 + j < T ? t0 + j : t0) * 64 + lane];
        addLeaves<CM, K>(acc, forest, leaves, T - t0 < K ? T - t0 : K);
    }
    finishPixel<CM>(acc, forest.C, static_cast<size_t>(W) * im.H, y * W + x, pout, lout);
}

template<int CM>
__global__ void __launch_bounds__(256) k_classify_split(const DevImage* __restrict__ imgs,
        const int* __restrict__ frameIdx, int numFrames, int W, int H, DevForest forest, float* __restrict__ probs,
        uint8_t* __restrict__ labels) {
    ClassifyTile tile;
    if (!classifyTile(splitTiles(W, H), numFrames, tile)) return;
    const int frame = tile.frame;
    const size_t plane = static_cast<size_t>(W) * H;
    classifyTileSplit<CM>(imgs[frameIdx[frame]], tile.block, 0, H, forest,
            probs ? probs + static_cast<size_t>(frame) * forest.C * plane : nullptr,
            labels ? labels + static_cast<size_t>(frame) * plane : nullptr);
}

template<int CM>
__global__ void __launch_bounds__(256) k_classify_jobs_split(const FrameJob* __restrict__ jobs, int numJobs,
        int totalTiles, DevForest forest) {
    ClassifyTile tile;
    if (!classifyTile(totalTiles, 1, tile)) return;
    const int j = tile.block;
    const FrameJob job = jobs[findJob(jobs, numJobs, j)];
    classifyTileSplit<CM>(job.im, j - job.firstTile, job.rowBegin, job.rowEnd, forest, job.probs, job.labels);
}

// the smallest instance whose class count CM covers C
template<class Launch>
void dispatchClasses(int C, Launch&& launch) {
    if (C <= 6) {
        launch(std::integral_constant<int, 6>{});
    } else if (C <= 8) {
        launch(std::integral_constant<int, 8>{});
    } else if (C <= kMaxClasses) {
        launch(std::integral_constant<int, kMaxClasses>{});
    } else {
        throw Error(CF_ERR_ARG, "too many classes (max 32)");
    }
}

// Launches of at most kSplitPixels pixels (CURFIL_CLASSIFY_SPLIT_MPIX overrides it, read once per process
// for single and batched predictions and when a serving loop is created; 0: never) split each
// tile's trees over the waves of a workgroup, given at least kSplitWaves trees and at most 128 (their
// leaf numbers, 256 bytes per tree, fill the workgroup's LDS).  Measured with 10 trees of depth 20: one
// 640x480 frame 0.80 against 0.89 ms to the host; 64 frames (19.7 Mpix) 2.7% slower split.  The bound
// covers every launch of the serving loop (at most 1.25 Mpix), which served the 320x240 / 640x480 /
// 1280x960 mix at 3000 frames/s with p99 12 ms instead of 30 ms; a bound of 0.5 or 0.8 Mpix, which
// leaves its larger launches to the other kernel, did worse than no split at all.
constexpr double kSplitPixels = 1.3e6;
constexpr int kSplitMaxTrees = 128;
double classifySplitPixels() {
    const char* e = std::getenv("CURFIL_CLASSIFY_SPLIT_MPIX");
    return e ? std::atof(e) * 1e6 : kSplitPixels;
}
static bool splitLaunch(double pixels, double maxPx, const DevForest& forest) {
    return pixels <= maxPx && forest.numTrees >= kSplitWaves && forest.numTrees <= kSplitMaxTrees;
}
static size_t splitLds(const DevForest& forest) {
    return static_cast<size_t>(forest.numTrees) * 64 * sizeof(int);
}

void launchClassify(hipStream_t s, const DevImage* imgs, const int* frameIdx, int numFrames, int W, int H,
        const DevForest& forest, float* probs, uint8_t* labels) {
    static const double maxPx = classifySplitPixels();   // once per process
    const double pixels = static_cast<double>(numFrames) * W * H;
    if (splitLaunch(pixels, maxPx, forest)) {
        const dim3 grid = classifyGrid(splitTiles(W, H), numFrames);
        dispatchClasses(forest.C, [&](auto cm) {
            hipLaunchKernelGGL(k_classify_split<decltype(cm)::value>, grid, dim3(64 * kSplitWaves), splitLds(forest),
                    s, imgs, frameIdx, numFrames, W, H, forest, probs, labels);
        });
        return;
    }
    const dim3 grid = classifyGrid(tileBlocks(W, H), numFrames);
    dispatchClasses(forest.C, [&](auto cm) {
        hipLaunchKernelGGL(k_classify<decltype(cm)::value>, grid, dim3(256), 0, s, imgs, frameIdx, numFrames, W, H,
                forest, probs, labels);
    });
}

ClassifyPlan planClassifyJobs(FrameJob* jobs, int numJobs, const DevForest& forest, double splitMaxPixels) {
    double pixels = 0;
    for (int i = 0; i < numJobs; i++) pixels += static_cast<double>(jobs[i].im.W) * (jobs[i].rowEnd - jobs[i].rowBegin);
    ClassifyPlan plan;
    plan.split = splitLaunch(pixels, splitMaxPixels, forest);
    int tiles = 0;
    for (int i = 0; i < numJobs; i++) {
        const int rows = jobs[i].rowEnd - jobs[i].rowBegin;
        const int t = plan.split ? splitTiles(jobs[i].im.W, rows) : tileBlocks(jobs[i].im.W, rows);
        if (tiles > INT32_MAX - t) throw Error(CF_ERR_UNSUPPORTED, "too many pixels in one launch");
        jobs[i].firstTile = tiles;
        tiles += t;
    }
    plan.totalTiles = tiles;
    return plan;
}

void launchClassifyJobs(hipStream_t s, const FrameJob* jobs, int numJobs, const ClassifyPlan& plan,
        const DevForest& forest) {
    if (numJobs == 0) return;
    const dim3 grid = classifyGrid(plan.totalTiles, 1);
    dispatchClasses(forest.C, [&](auto cm) {
        if (plan.split) {
            hipLaunchKernelGGL(k_classify_jobs_split<decltype(cm)::value>, grid, dim3(64 * kSplitWaves),
                    splitLds(forest), s, jobs, numJobs, plan.totalTiles, forest);
        } else {
            hipLaunchKernelGGL(k_classify_jobs<decltype(cm)::value>, grid, dim3(256), 0, s, jobs, numJobs,
                    plan.totalTiles, forest);
        }
    });
}

}  // namespace cf
