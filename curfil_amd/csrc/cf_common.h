// Shared definitions of the curfil_amd library: device image descriptors,
// packed feature records and the bit-exact feature response.
//
// The response restates ImageFeatureFunction::calculateFeatureResponse and
// PixelInstance::averageRegionColor/Depth (random_tree_image.h:222-316,
// 485-623) operation by operation: float offset normalisation with correctly
// rounded division, double box sums of float integral corners, and
// ((double)sum / 1000.0) / numValid for depth.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/curfil_hip.h"

namespace cf {

constexpr int kFeatParams = 11;
constexpr int kMaxClasses = 32;     // label bins per node (reference caps at 32 on the GPU path)
constexpr int kMaxThresholds = 256;

enum FeatParam { P_TYPE = 0, P_O1X, P_O1Y, P_O2X, P_O2Y, P_R1X, P_R1Y, P_R2X, P_R2Y, P_C1, P_C2 };
enum FeatType { T_DEPTH = 0, T_COLOR = 1 };

// One feature as 16 bytes (11 parameters + padding) so a uniform feature is
// one scalar 16-byte load.
struct __align__(16) FeatRec {
    int8_t p[16];
};

// Device view of one integral image.  col: [3][H][W] float; dep: [H][W]
// int2 {sum of depth mm (mod 2^32), count of valid pixels}.
struct DevImage {
    const float* col;
    const int2* dep;
    int W, H;
    int pad0, pad1;
};

__host__ __device__ inline double cf_nan() { return __builtin_nan(""); }

// Depth::getFloatValue (image.h:143-146)
__device__ __forceinline__ float depthToFloat(int32_t mm) {
    return static_cast<float>(mm) / 1000.0f;
}

// XY::normalize (random_tree_image.h:82-87): int -> float division -> truncation.
__device__ __forceinline__ int normalizeCoord(int v, float d) {
    return static_cast<int>(static_cast<float>(v) / d);
}

// averageRegionColor (random_tree_image.h:222-262): box (lx,rx] x (uy,ly].
__device__ __forceinline__ double boxColor(const DevImage& im, int ch, int px, int py, int ox, int oy, int rx,
        int ry) {
    const int w = rx > 1 ? rx : 1;
    const int h = ry > 1 ? ry : 1;
    const int x = px + ox, y = py + oy;
    const int lx = x - w, hx = x + w, uy = y - h, ly = y + h;
    if (lx < 0 || hx >= im.W || uy < 0 || ly >= im.H) return cf_nan();
    const float* c = im.col + static_cast<size_t>(ch) * im.W * im.H;
    const double lowerRight = c[ly * im.W + hx];
    const double lowerLeft = c[ly * im.W + lx];
    const double upperRight = c[uy * im.W + hx];
    const double upperLeft = c[uy * im.W + lx];
    return (lowerRight - upperRight) + (upperLeft - lowerLeft);
}

// averageRegionDepth (random_tree_image.h:268-316).  Sums use wrapping
// 32-bit arithmetic: the box sum is exact whenever it fits in int32, even if
// single integral entries wrapped (they are stored modulo 2^32).
__device__ __forceinline__ double boxDepth(const DevImage& im, int px, int py, int ox, int oy, int rx, int ry) {
    const int w = rx > 1 ? rx : 1;
    const int h = ry > 1 ? ry : 1;
    const int x = px + ox, y = py + oy;
    const int lx = x - w, hx = x + w, uy = y - h, ly = y + h;
    if (lx < 0 || hx >= im.W || uy < 0 || ly >= im.H) return cf_nan();
    const int2 lr = im.dep[ly * im.W + hx];
    const int2 ll = im.dep[ly * im.W + lx];
    const int2 ur = im.dep[uy * im.W + hx];
    const int2 ul = im.dep[uy * im.W + lx];
    const int numValid = static_cast<int>((static_cast<uint32_t>(lr.y) - static_cast<uint32_t>(ur.y)) +
            (static_cast<uint32_t>(ul.y) - static_cast<uint32_t>(ll.y)));
    if (numValid == 0) return cf_nan();
    const int sum = static_cast<int>((static_cast<uint32_t>(lr.x) - static_cast<uint32_t>(ur.x)) +
            (static_cast<uint32_t>(ul.x) - static_cast<uint32_t>(ll.x)));
    const double feat = static_cast<double>(sum) / 1000.0;
    return feat / static_cast<double>(numValid);
}

// calculateColorFeature / calculateDepthFeature (random_tree_image.h:564-623).
__device__ __forceinline__ double featureResponse(const DevImage& im, const int8_t* p, int px, int py,
        int32_t depthMm, bool flip) {
    if (!(depthMm > 0)) return cf_nan();
    const float d = depthToFloat(depthMm);
    const int o1x = flip ? -p[P_O1X] : p[P_O1X];
    const int o2x = flip ? -p[P_O2X] : p[P_O2X];
    double a, b;
    if (p[P_TYPE] == T_DEPTH) {
        a = boxDepth(im, px, py, normalizeCoord(o1x, d), normalizeCoord(p[P_O1Y], d),
                normalizeCoord(p[P_R1X], d), normalizeCoord(p[P_R1Y], d));
        if (__builtin_isnan(a)) return a;
        b = boxDepth(im, px, py, normalizeCoord(o2x, d), normalizeCoord(p[P_O2Y], d),
                normalizeCoord(p[P_R2X], d), normalizeCoord(p[P_R2Y], d));
        if (__builtin_isnan(b)) return b;
    } else {
        a = boxColor(im, static_cast<uint8_t>(p[P_C1]), px, py, normalizeCoord(o1x, d), normalizeCoord(p[P_O1Y], d),
                normalizeCoord(p[P_R1X], d), normalizeCoord(p[P_R1Y], d));
        if (__builtin_isnan(a)) return a;
        b = boxColor(im, static_cast<uint8_t>(p[P_C2]), px, py, normalizeCoord(o2x, d), normalizeCoord(p[P_O2Y], d),
                normalizeCoord(p[P_R2X], d), normalizeCoord(p[P_R2Y], d));
        if (__builtin_isnan(b)) return b;
    }
    return a - b;
}

// Smallest float >= r (NaN stays NaN).  For a float threshold t:
//   r <= (double)t  <=>  roundUpToFloat(r) <= t
// so the split test `r <= t` of SplitFunction::split (random_tree.h:128) can be
// evaluated in float without changing any decision.
__device__ __forceinline__ float roundUpToFloat(double r) {
    float f = static_cast<float>(r);
    if (static_cast<double>(f) < r) {
        uint32_t u = __float_as_uint(f);
        if (f == 0.0f) u = 1u;                 // +0 or -0 -> smallest positive denormal
        else if (u & 0x80000000u) u -= 1u;     // negative: toward zero
        else u += 1u;                          // positive: away from zero
        f = __uint_as_float(u);
    }
    return f;
}

// Node split record used by partition and prediction.
struct SplitRec {
    FeatRec feat;
    float threshold;
    int32_t flags;
    int32_t pad0, pad1;
};

}  // namespace cf
