// Dense per-pixel forest prediction with the CPU branch's arithmetic of
// RandomForestImage::predict (random_forest_image.cpp:307-345):
//   * pixel depth by inclusion-exclusion on the integrals, invalid -> NaN
//     responses (PixelInstance, random_tree_image.h:140-165); every split
//     sends NaN right (SplitFunction::split, random_tree.h:122-144);
//   * per tree, the leaf's double posterior is added into a float
//     accumulator; the double sum normalises; the label is the first class
//     with the largest probability (strict >, starting from -1).
// One thread per pixel; the frames of a batch are the grid's y dimension, so
// one launch classifies a whole batch against all trees.
#include "cf_internal.h"

namespace cf {

template<int CM>
__global__ void __launch_bounds__(256) k_classify(const DevImage* __restrict__ imgs, const int* __restrict__ frameIdx,
        int W, int H, DevForest forest, float* __restrict__ probs, uint8_t* __restrict__ labels) {
    const int frame = blockIdx.y;
    const int px = blockIdx.x * blockDim.x + threadIdx.x;
    if (px >= W * H) return;
    const int x = px % W, y = px / W;
    const DevImage im = imgs[frameIdx[frame]];
    const int C = forest.C;
    // PixelInstance depth
    auto at = [&](int xx, int yy) -> int2 { return im.dep[yy * W + xx]; };
    const int2 c00 = at(x, y);
    const int2 cl = x > 0 ? at(x - 1, y) : make_int2(0, 0);
    const int2 ca = y > 0 ? at(x, y - 1) : make_int2(0, 0);
    const int2 cal = (x > 0 && y > 0) ? at(x - 1, y - 1) : make_int2(0, 0);
    const int valid = static_cast<int>(static_cast<uint32_t>(c00.y) -
            (static_cast<uint32_t>(cl.y) + static_cast<uint32_t>(ca.y) - static_cast<uint32_t>(cal.y)));
    int32_t depth = -1;
    if (valid == 1) {
        depth = static_cast<int32_t>(static_cast<uint32_t>(c00.x) -
                (static_cast<uint32_t>(cl.x) + static_cast<uint32_t>(ca.x) - static_cast<uint32_t>(cal.x)));
    }
    float acc[CM];
#pragma unroll
    for (int c = 0; c < CM; c++) acc[c] = 0.0f;
    for (int t = 0; t < forest.numTrees; t++) {
        int node = forest.treeRoot[t];
        int ch = forest.child[node];
        while (ch >= 0) {
            const Feat f = loadFeat(&forest.nodes[node].feat);
            const float threshold = forest.nodes[node].threshold;
            const double r = featureResponse(im, f, x, y, depth, false);
            node = (r <= static_cast<double>(threshold)) ? ch : ch + 1;
            ch = forest.child[node];
        }
        const double* post = forest.post + static_cast<size_t>(-ch - 1) * C;
#pragma unroll
        for (int c = 0; c < CM; c++) {
            if (c < C) acc[c] = static_cast<float>(static_cast<double>(acc[c]) + post[c]);
        }
    }
    double sum = 0.0;
#pragma unroll
    for (int c = 0; c < CM; c++) {
        if (c < C) sum += static_cast<double>(acc[c]);
    }
    float bestProb = -1.0f;
    int label = 0;
    const size_t plane = static_cast<size_t>(W) * H;
    float* pout = probs ? probs + static_cast<size_t>(frame) * C * plane : nullptr;
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
    if (labels) labels[static_cast<size_t>(frame) * plane + px] = static_cast<uint8_t>(label);
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

}  // namespace cf
