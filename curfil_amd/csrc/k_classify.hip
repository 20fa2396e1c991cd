// Dense per-pixel forest prediction with the CPU branch's arithmetic of
// RandomForestImage::predict (random_forest_image.cpp:307-345):
//   * pixel depth by inclusion-exclusion on the integrals, invalid -> NaN
//     responses (PixelInstance, random_tree_image.h:140-165); every split
//     sends NaN right (SplitFunction::split, random_tree.h:122-144);
//   * per tree, the leaf's double posterior is added into a float
//     accumulator; the double sum normalises; the label is the first class
//     with the largest probability (strict >, starting from -1).
// One thread per pixel; the frames of a batch are the grid's y dimension, so
// one launch classifies a whole batch against all trees.  The forest holds
// split nodes only (see uploadForest): a child that is a leaf is referenced by
// its leaf index, so a walk ends at the last split without reading a leaf
// record.
#include "cf_internal.h"

namespace cf {

// One node record: the feature (16 bytes) and {threshold, left, right, -}; a child reference >= 0 is
// a node, < 0 the leaf -(ref)-1.
struct NodeRec {
    uint4 feat;
    uint4 tail;
};

__device__ __forceinline__ Feat unpackFeat(const uint4& w) {
    Feat f;
    f.type = sbyte(w.x, 0);
    f.o1x = sbyte(w.x, 1);
    f.o1y = sbyte(w.x, 2);
    f.o2x = sbyte(w.x, 3);
    f.o2y = sbyte(w.y, 0);
    f.r1x = sbyte(w.y, 1);
    f.r1y = sbyte(w.y, 2);
    f.r2x = sbyte(w.y, 3);
    f.r2y = sbyte(w.z, 0);
    f.c1 = static_cast<int>((w.z >> 8) & 0xffu);
    f.c2 = static_cast<int>((w.z >> 16) & 0xffu);
    return f;
}

__device__ __forceinline__ NodeRec loadNode(const SplitRec* nodes, int i) {
    const uint4* p = reinterpret_cast<const uint4*>(nodes + i);
    return NodeRec{p[0], p[1]};
}

// PixelInstance depth (random_tree_image.h:140-165) by inclusion-exclusion on the integral: true and
// the depth in mm when the pixel's depth is valid.
__device__ __forceinline__ bool pixelDepthAt(const DevImage& im, int x, int y, int32_t& depth) {
    const int W = im.W;
    auto at = [&](int xx, int yy) -> int2 { return im.dep[yy * W + xx]; };
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
// (invalid depth, a box outside the image, a depth box without valid pixels) goes right.
__device__ __forceinline__ int walkTree(const DevImage& im, int x, int y, bool depthOk, double rinv,
        const DevForest& forest, int root) {
    const int W = im.W, H = im.H;
    const size_t plane = static_cast<size_t>(W) * H;
    int ref = root;
    while (ref >= 0) {
        const NodeRec rec = loadNode(forest.nodes, ref);
        const Feat f = unpackFeat(rec.feat);
        int idx[8];
        const bool isDepth = f.type == T_DEPTH;
        const bool ok = depthOk && responseCorners(f, x, y, W, H, rinv, false, idx);
        // one 8-byte load per depth corner and one 4-byte load per colour corner, in separate branches
        // (a shared corner array makes the compiler split the depth loads in two)
        double r = cf_nan();
        if (ok && isDepth) {
            int2 dv[8];
#pragma unroll
            for (int c = 0; c < 8; c++) dv[c] = planeAt(im.dep, idx[c]);
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
            for (int c = 0; c < 4; c++) cv[c] = planeAt(colA, idx[c]);
#pragma unroll
            for (int c = 4; c < 8; c++) cv[c] = planeAt(colB, idx[c]);
            r = colorResponseFromCorners(cv);
        }
        const float thr = __uint_as_float(rec.tail.x);
        ref = static_cast<int>((r <= static_cast<double>(thr)) ? rec.tail.y : rec.tail.z);
    }
    return -ref - 1;
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
            if (pout) pout[c * plane + px] = p;
            if (p > bestProb) {
                label = c;
                bestProb = p;
            }
        }
    }
    if (lout) lout[px] = static_cast<uint8_t>(label);
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
        const double* post = forest.post + static_cast<size_t>(leaf) * C;
        // the leaf's row is read whole (indices past C repeat the last class, in bounds and unused) so the
        // loads go out together instead of one conditional load after another
        double pv[CM];
#pragma unroll
        for (int c = 0; c < CM; c++) pv[c] = post[c < C ? c : C - 1];
#pragma unroll
        for (int c = 0; c < CM; c++) {
            if (c < C) acc[c] = static_cast<float>(static_cast<double>(acc[c]) + pv[c]);
        }
    }
    finishPixel<CM>(acc, C, plane, px, pout, lout);
}

// frames of one image set: frame f of the batch is image frameIdx[f]; outputs stacked per frame
template<int CM>
__global__ void __launch_bounds__(256) k_classify(const DevImage* __restrict__ imgs, const int* __restrict__ frameIdx,
        int W, int H, DevForest forest, float* __restrict__ probs, uint8_t* __restrict__ labels) {
    const int frame = blockIdx.y;
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= W * H) return;
    const size_t plane = static_cast<size_t>(W) * H;
    classifyPixel<CM>(imgs[frameIdx[frame]], px, forest,
            probs ? probs + static_cast<size_t>(frame) * forest.C * plane : nullptr,
            labels ? labels + static_cast<size_t>(frame) * plane : nullptr);
}

// independent frames of one size, each with its own images and outputs
template<int CM>
__global__ void __launch_bounds__(256) k_classify_jobs(const FrameJob* __restrict__ jobs, int W, int H,
        DevForest forest) {
    const FrameJob job = jobs[blockIdx.y];
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= W * H) return;
    classifyPixel<CM>(job.im, px, forest, job.probs, job.labels);
}

void launchClassify(hipStream_t s, const DevImage* imgs, const int* frameIdx, int numFrames, int W, int H,
        const DevForest& forest, float* probs, uint8_t* labels) {
    const dim3 grid((W * H + 255) / 256, numFrames);
    if (forest.C <= 8) {
        hipLaunchKernelGGL(k_classify<8>, grid, dim3(256), 0, s, imgs, frameIdx, W, H, forest, probs, labels);
    } else if (forest.C <= kMaxClasses) {
        hipLaunchKernelGGL(k_classify<kMaxClasses>, grid, dim3(256), 0, s, imgs, frameIdx, W, H, forest, probs, labels);
    } else {
        throw Error(CF_ERR_ARG, "too many classes (max 32)");
    }
}

void launchClassifyJobs(hipStream_t s, const FrameJob* jobs, int numJobs, int W, int H, const DevForest& forest) {
    if (numJobs == 0) return;
    const dim3 grid((W * H + 255) / 256, numJobs);
    if (forest.C <= 8) {
        hipLaunchKernelGGL(k_classify_jobs<8>, grid, dim3(256), 0, s, jobs, W, H, forest);
    } else if (forest.C <= kMaxClasses) {
        hipLaunchKernelGGL(k_classify_jobs<kMaxClasses>, grid, dim3(256), 0, s, jobs, W, H, forest);
    } else {
        throw Error(CF_ERR_ARG, "too many classes (max 32)");
    }
}

}  // namespace cf
