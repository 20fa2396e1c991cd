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

// A feature's parameters unpacked to ints.  Kernels read a FeatRec as one
// aligned uint4 and unpack it with shifts: byte-indexed access to the record
// let the compiler form a misaligned scalar base address (record + 5) with
// signed immediate offsets, which the scalar unit does not add up as intended.
struct Feat {
    int type, o1x, o1y, o2x, o2y, r1x, r1y, r2x, r2y, c1, c2;
};

__host__ __device__ __forceinline__ int sbyte(uint32_t w, int k) {
    return static_cast<int>(static_cast<int8_t>((w >> (8 * k)) & 0xffu));
}

__device__ __forceinline__ Feat loadFeat(const FeatRec* rec) {
    const uint4 w = *reinterpret_cast<const uint4*>(rec);
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

// XY::normalize (random_tree_image.h:82-87) is (int)((float)v / d).  For the
// operands that occur here (|v| <= 127, d = mm / 1000.0f with mm <= 50000)
// that equals (int)(float)(v * rinv) with rinv = 1.0 / (double)d: the product
// is within 2^-52 (relative) of v/d, while a quotient v/d of a 7-bit integer by
// a 24-bit float lies at least 2^-49 (relative) away from every float rounding
// midpoint and is never one itself, so both round to the same float.
// tests/test_numerics.py checks all 12.75M (v, mm) pairs.
__host__ __device__ __forceinline__ double depthReciprocal(float d) {
    return 1.0 / static_cast<double>(d);
}
__device__ __forceinline__ int normalizeCoord(int v, double rinv) {
    return static_cast<int>(static_cast<float>(static_cast<double>(v) * rinv));
}

// averageRegionColor (random_tree_image.h:222-262): box (lx,rx] x (uy,ly] of
// one colour plane c (W x H).
__device__ __forceinline__ double boxColor(const float* __restrict__ c, int W, int H, int px, int py, int ox, int oy,
        int rx, int ry) {
    const int w = rx > 1 ? rx : 1;
    const int h = ry > 1 ? ry : 1;
    const int x = px + ox, y = py + oy;
    const int lx = x - w, hx = x + w, uy = y - h, ly = y + h;
    if (lx < 0 || hx >= W || uy < 0 || ly >= H) return cf_nan();
    const double lowerRight = c[ly * W + hx];
    const double lowerLeft = c[ly * W + lx];
    const double upperRight = c[uy * W + hx];
    const double upperLeft = c[uy * W + lx];
    return (lowerRight - upperRight) + (upperLeft - lowerLeft);
}

// averageRegionDepth (random_tree_image.h:268-316).  Sums use wrapping
// 32-bit arithmetic: the box sum is exact whenever it fits in int32, even if
// single integral entries wrapped (they are stored modulo 2^32).
// (double)sum / 1000.0, correctly rounded, without a division: the product with
// RN(1/1000) corrected once by the exact residual fma(-q, 1000, a).  Checked
// against the division for every 32-bit integer sum (tests/test_numerics.py);
// box sums are 32-bit integers.
__host__ __device__ __forceinline__ double depthSumToMetres(int sum) {
    constexpr double kInv1000 = 1.0 / 1000.0;
    const double a = static_cast<double>(sum);
    const double q = a * kInv1000;
    return fma(fma(-q, 1000.0, a), kInv1000, q);
}

__device__ __forceinline__ double boxDepth(const int2* __restrict__ dep, int W, int H, int px, int py, int ox, int oy,
        int rx, int ry) {
    const int w = rx > 1 ? rx : 1;
    const int h = ry > 1 ? ry : 1;
    const int x = px + ox, y = py + oy;
    const int lx = x - w, hx = x + w, uy = y - h, ly = y + h;
    if (lx < 0 || hx >= W || uy < 0 || ly >= H) return cf_nan();
    const int2 lr = dep[ly * W + hx];
    const int2 ll = dep[ly * W + lx];
    const int2 ur = dep[uy * W + hx];
    const int2 ul = dep[uy * W + lx];
    const int numValid = static_cast<int>((static_cast<uint32_t>(lr.y) - static_cast<uint32_t>(ur.y)) +
            (static_cast<uint32_t>(ul.y) - static_cast<uint32_t>(ll.y)));
    if (numValid == 0) return cf_nan();
    const int sum = static_cast<int>((static_cast<uint32_t>(lr.x) - static_cast<uint32_t>(ur.x)) +
            (static_cast<uint32_t>(ul.x) - static_cast<uint32_t>(ll.x)));
    return depthSumToMetres(sum) / static_cast<double>(numValid);
}

// calculateColorFeature / calculateDepthFeature (random_tree_image.h:564-623)
// for a pixel with valid depth; rinv = depthReciprocal(Depth::getFloatValue()).
__device__ __forceinline__ double featureResponseR(const float* __restrict__ col, const int2* __restrict__ dep, int W,
        int H, const Feat& f, int px, int py, double rinv, bool flip) {
    const int o1x = flip ? -f.o1x : f.o1x;
    const int o2x = flip ? -f.o2x : f.o2x;
    double a, b;
    if (f.type == T_DEPTH) {
        a = boxDepth(dep, W, H, px, py, normalizeCoord(o1x, rinv), normalizeCoord(f.o1y, rinv),
                normalizeCoord(f.r1x, rinv), normalizeCoord(f.r1y, rinv));
        if (__builtin_isnan(a)) return a;
        b = boxDepth(dep, W, H, px, py, normalizeCoord(o2x, rinv), normalizeCoord(f.o2y, rinv),
                normalizeCoord(f.r2x, rinv), normalizeCoord(f.r2y, rinv));
        if (__builtin_isnan(b)) return b;
    } else {
        const size_t plane = static_cast<size_t>(W) * H;
        a = boxColor(col + static_cast<size_t>(f.c1) * plane, W, H, px, py, normalizeCoord(o1x, rinv),
                normalizeCoord(f.o1y, rinv), normalizeCoord(f.r1x, rinv), normalizeCoord(f.r1y, rinv));
        if (__builtin_isnan(a)) return a;
        b = boxColor(col + static_cast<size_t>(f.c2) * plane, W, H, px, py, normalizeCoord(o2x, rinv),
                normalizeCoord(f.o2y, rinv), normalizeCoord(f.r2x, rinv), normalizeCoord(f.r2y, rinv));
        if (__builtin_isnan(b)) return b;
    }
    return a - b;
}

// The same from the pixel's depth in mm, NaN if the depth is invalid.
__device__ __forceinline__ double featureResponse(const DevImage& im, const Feat& f, int px, int py,
        int32_t depthMm, bool flip) {
    if (!(depthMm > 0)) return cf_nan();
    return featureResponseR(im.col, im.dep, im.W, im.H, f, px, py, depthReciprocal(depthToFloat(depthMm)), flip);
}

// Gather-first form of the response, used where latency matters: the 8 corner
// indices (box a: lower-right, lower-left, upper-right, upper-left, then box b)
// are formed first so all corner loads can be in flight together.  Returns
// false if either box leaves the image (the reference's NaN).  Computing box
// b even when box a is already NaN changes nothing: the response is NaN then.
__device__ __forceinline__ bool boxCorners(int px, int py, int ox, int oy, int rx, int ry, int W, int H, int* idx) {
    const int w = rx > 1 ? rx : 1;
    const int h = ry > 1 ? ry : 1;
    const int x = px + ox, y = py + oy;
    const int lx = x - w, hx = x + w, uy = y - h, ly = y + h;
    idx[0] = ly * W + hx;
    idx[1] = ly * W + lx;
    idx[2] = uy * W + hx;
    idx[3] = uy * W + lx;
    return !(lx < 0 || hx >= W || uy < 0 || ly >= H);
}

__device__ __forceinline__ bool responseCorners(const Feat& f, int px, int py, int W, int H, double rinv, bool flip,
        int* idx) {
    const int o1x = flip ? -f.o1x : f.o1x;
    const int o2x = flip ? -f.o2x : f.o2x;
    const bool a = boxCorners(px, py, normalizeCoord(o1x, rinv), normalizeCoord(f.o1y, rinv),
            normalizeCoord(f.r1x, rinv), normalizeCoord(f.r1y, rinv), W, H, idx);
    const bool b = boxCorners(px, py, normalizeCoord(o2x, rinv), normalizeCoord(f.o2y, rinv),
            normalizeCoord(f.r2x, rinv), normalizeCoord(f.r2y, rinv), W, H, idx + 4);
    return a && b;
}

// Element i of an image plane through a 32-bit byte offset (a plane is far
// below 4 GiB), so the load can take the uniform base in scalar registers.
template<typename V>
__device__ __forceinline__ V planeAt(const V* base, int i) {
    const uint32_t off = static_cast<uint32_t>(i) * static_cast<uint32_t>(sizeof(V));
    return *reinterpret_cast<const V*>(reinterpret_cast<const char*>(base) + off);
}

// averageRegionColor of both boxes and their difference from gathered corners.
__device__ __forceinline__ double colorResponseFromCorners(const float* c) {
    const double a = (static_cast<double>(c[0]) - static_cast<double>(c[2])) +
            (static_cast<double>(c[3]) - static_cast<double>(c[1]));
    const double b = (static_cast<double>(c[4]) - static_cast<double>(c[6])) +
            (static_cast<double>(c[7]) - static_cast<double>(c[5]));
    return a - b;
}

// The mean is formed whether or not the box holds a valid pixel and then
// selected, with no branch: a branch lets the compiler sink the loads of the
// sums behind the count check and issue a corner's two halves apart.
__device__ __forceinline__ double depthBoxFromCorners(const int2* d) {
    const int numValid = static_cast<int>((static_cast<uint32_t>(d[0].y) - static_cast<uint32_t>(d[2].y)) +
            (static_cast<uint32_t>(d[3].y) - static_cast<uint32_t>(d[1].y)));
    const int sum = static_cast<int>((static_cast<uint32_t>(d[0].x) - static_cast<uint32_t>(d[2].x)) +
            (static_cast<uint32_t>(d[3].x) - static_cast<uint32_t>(d[1].x)));
    const double mean = depthSumToMetres(sum) / static_cast<double>(numValid);
    return numValid == 0 ? cf_nan() : mean;
}

// averageRegionDepth of both boxes and their difference (NaN if a box has no
// valid pixel).
__device__ __forceinline__ double depthResponseFromCorners(const int2* d) {
    const double a = depthBoxFromCorners(d);
    const double b = depthBoxFromCorners(d + 4);
    return __builtin_isnan(a) ? a : (__builtin_isnan(b) ? b : a - b);
}

// Smallest float >= r (NaN stays NaN).  For a float threshold t:
//   r <= (double)t  <=>  roundUpToFloat(r) <= t
// so the split test `r <= t` of SplitFunction::split (random_tree.h:128) can be
// evaluated in float without changing any decision.
__device__ __forceinline__ float roundUpToFloat(double r) {
    const float f = static_cast<float>(r);
    const uint32_t u = __float_as_uint(f);
    // the next float up: +0 or -0 -> smallest positive denormal, negative -> toward zero, positive -> away
    const uint32_t next = (f == 0.0f) ? 1u : ((u & 0x80000000u) ? u - 1u : u + 1u);
    return (static_cast<double>(f) < r) ? __uint_as_float(next) : f;
}

// Node split record used by partition and prediction.
struct SplitRec {
    FeatRec feat;
    float threshold;
    int32_t flags;       // packed forests: left child (global index), or -(leaf index)-1
    int32_t pad0, pad1;
};

}  // namespace cf
