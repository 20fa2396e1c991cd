// CPU ORACLE FOR PARITY TESTS ONLY.
//
// This file is test infrastructure: a from-scratch restatement of the CPU
// (AccelerationMode::CPU_ONLY) path of CURFIL's random-forest training and
// prediction.  Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
// leg may load it, and only as the checker or as the timed CPU baseline.  The
// product library (curfil_amd) never links or calls it.
//
// Every function cites the reference file:line whose behaviour it restates.
// It is written in C++ (not plain C) because the root-sample ordering of the
// reference depends on libstdc++'s std::sort (random_tree_image.cpp:776-802),
// which this file reuses with the same comparator, and on a stable key sort of
// the features (random_tree_image.cpp:517, thrust::sort_by_key on int keys).
//
// Pinning (what checks each part; tests/test_oracle_*.py):
//   * MT19937 + Boost uniform_int + sampleFeature draw order + threshold
//     sampling + stable sort: golden hash 4763250482941632261 of
//     feature_generation_test.cpp:310-387, plus the standard MT19937 value
//     4123659995 for the 10000th output of seed 5489.
//   * integrals: image_test.cpp:87-279 known answers, brute force.
//   * responses / counters: random_tree_image_gpu_test.cu:318-866 known answers.
//   * scores: random_tree_image_gpu_test.cu:106-240 known answers and a
//     50-digit decimal recomputation (tests/test_oracle_pins.py).
//   * argmax tie rule: constructed tie / near-tie cases.
//   * normalizeHistogram: random_tree_test.cpp:46-154 known answers.
//   * level loop: forced-outcome cases (maxDepth, minSampleCount, node-id
//     rollback) whose tree does not depend on the features drawn.
//   * prediction: one-hot forests over the known-answer responses of
//     random_tree_image_gpu_test.cu:653-766 (votes, probabilities, labels).
//   * class-uniform merge order with several threads: parity unpinned (the
//     reference's own order is nondeterministic there); single-thread order.
//
// Node counters and prediction run over host threads (ORC_THREADS, default the
// hardware threads, at most 16), as the reference's TBB loops do; both are
// integer sums or independent rows, so the result does not depend on it.
//
// Build: oracle/Makefile (g++ -O2 -ffp-contract=off, no fast-math).

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace orc {

// ---------------------------------------------------------------------------
// RNG: boost::mt19937 + boost::uniform_int<int>, RandomSource seed counter.
// random_tree.h:1009-1063 (Sampler), random_tree.cpp:107-112 (getNext),
// random_tree.h:1129-1156 (RandomSource).  Boost's mt19937 is the standard
// MT19937; uniform_int is Boost's bucket-division rejection method.
// ---------------------------------------------------------------------------
struct MT19937 {
    uint32_t mt[624];
    int idx;
    explicit MT19937(uint32_t s) { seed(s); }
    void seed(uint32_t s) {
        mt[0] = s;
        for (int i = 1; i < 624; i++) {
            mt[i] = 1812433253u * (mt[i - 1] ^ (mt[i - 1] >> 30)) + static_cast<uint32_t>(i);
        }
        idx = 624;
    }
    void twist() {
        for (int i = 0; i < 624; i++) {
            uint32_t y = (mt[i] & 0x80000000u) | (mt[(i + 1) % 624] & 0x7fffffffu);
            uint32_t v = mt[(i + 397) % 624] ^ (y >> 1);
            if (y & 1u) v ^= 0x9908b0dfu;
            mt[i] = v;
        }
        idx = 0;
    }
    uint32_t next() {
        if (idx >= 624) twist();
        uint32_t y = mt[idx++];
        y ^= (y >> 11);
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= (y >> 18);
        return y;
    }
};

// Sampler(seed, lower, upper): random_tree.h:1015-1018.
struct Sampler {
    int seed, lower, upper;
    MT19937 rng;
    Sampler(int seed, int lower, int upper) : seed(seed), lower(lower), upper(upper),
            rng(static_cast<uint32_t>(seed)) {}
    // Boost uniform_int<int> over the 32-bit engine range (brange = 2^32-1 > range).
    int getNext() {
        const uint32_t range = static_cast<uint32_t>(upper) - static_cast<uint32_t>(lower);
        if (range == 0) return lower;  // no engine call
        const uint32_t brange = 0xFFFFFFFFu;
        uint32_t bucket = brange / (range + 1u);
        if (brange % (range + 1u) == range) ++bucket;
        for (;;) {
            uint32_t r = rng.next() / bucket;
            if (r <= range) return static_cast<int>(static_cast<uint32_t>(lower) + r);
        }
    }
};

// RandomSource: random_tree.h:1129-1156 -- every sampler takes seed++.
struct RandomSource {
    int seed;
    explicit RandomSource(int s) : seed(s) {}
    Sampler uniformSampler(int val) { return uniformSampler(0, val - 1); }
    Sampler uniformSampler(int lo, int hi) { return Sampler(seed++, lo, hi); }
};

// ReservoirSampler: random_tree.h:1069-1122.
template<class T>
struct Reservoir {
    size_t count = 0, samples = 0;
    std::vector<T> reservoir;
    explicit Reservoir(size_t n = 0) : samples(n) { reservoir.reserve(n); }
    void sample(Sampler& s, const T& v) {
        if (samples == 0) throw std::runtime_error("no samples to sample");
        if (reservoir.size() < samples) {
            reservoir.push_back(v);
        } else {
            size_t r = static_cast<size_t>(s.getNext()) % (count + 1);
            if (r < samples) reservoir[r] = v;
        }
        count++;
    }
};

// ---------------------------------------------------------------------------
// Images: RGBDImage integral in double with one running Kahan term per
// channel (image.cpp:236-277), stored as float (colour) / int (depth, valid).
// ---------------------------------------------------------------------------
struct Image {
    int W = 0, H = 0;
    std::vector<float> col;   // [3][H][W] integral
    std::vector<int32_t> dep; // [H][W] integral of mm
    std::vector<int32_t> val; // [H][W] integral of validity
    std::vector<uint8_t> lab; // [H][W] labels

    float color(int c, int x, int y) const { return col[(static_cast<size_t>(c) * H + y) * W + x]; }
    int32_t depthI(int x, int y) const { return dep[static_cast<size_t>(y) * W + x]; }
    int32_t validI(int x, int y) const { return val[static_cast<size_t>(y) * W + x]; }
};

// Colour channel: the reference copies floats into a double array, integrates
// in double (T=double), copies back to float (image.cpp:236-248).
static void integrateColorChannel(const float* in, int W, int H, float* out) {
    std::vector<double> d(static_cast<size_t>(W) * H);
    for (size_t i = 0; i < d.size(); i++) d[i] = in[i];
    // T = double: static_casts are identities.
    double c = 0.0;
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            double above = (y > 0) ? d[(y - 1) * W + x] : 0.0;
            double left = (x > 0) ? d[y * W + x - 1] : 0.0;
            double aboveLeft = (x > 0 && y > 0) ? d[(y - 1) * W + x - 1] : 0.0;
            double dy = (left + above - aboveLeft) - c;
            double t = d[y * W + x] + dy;
            c = (t - d[y * W + x]) - dy;
            d[y * W + x] = t;
        }
    }
    for (size_t i = 0; i < d.size(); i++) out[i] = static_cast<float>(d[i]);
}

// Int channel, T = int: image.cpp:250-277 instantiated for int (no overflow
// for the sizes the oracle is run at).
static void integrateIntChannel(const int32_t* in, int W, int H, int32_t* out) {
    std::vector<int32_t> d(in, in + static_cast<size_t>(W) * H);
    double c = 0.0;
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            int above = (y > 0) ? d[(y - 1) * W + x] : 0;
            int left = (x > 0) ? d[y * W + x - 1] : 0;
            int aboveLeft = (x > 0 && y > 0) ? d[(y - 1) * W + x - 1] : 0;
            double dy = (left + above - aboveLeft) - c;
            double t = d[y * W + x] + dy;
            c = (t - d[y * W + x]) - dy;
            d[y * W + x] = static_cast<int>(t);
        }
    }
    std::memcpy(out, d.data(), d.size() * sizeof(int32_t));
}

struct Dataset {
    std::vector<Image> images;
};

// ---------------------------------------------------------------------------
// Samples / pixel depth: PixelInstance (random_tree_image.h:140-181).
// ---------------------------------------------------------------------------
struct Sample {       // layout shared with the C ABI's cf_sample
    uint16_t x, y;
    uint32_t image;
    int32_t depth_mm; // <= 0: invalid (Depth::INVALID is -1)
    uint8_t label, flip, pad0, pad1;
};
static_assert(sizeof(Sample) == 16, "sample record is 16 bytes");

// random_tree_image.h:148-164: depth by inclusion-exclusion on the integrals.
static int32_t pixelDepth(const Image& im, int x, int y) {
    int aboveValid = (y > 0) ? im.validI(x, y - 1) : 0;
    int leftValid = (x > 0) ? im.validI(x - 1, y) : 0;
    int aboveLeftValid = (x > 0 && y > 0) ? im.validI(x - 1, y - 1) : 0;
    int valid = im.validI(x, y) - (leftValid + aboveValid - aboveLeftValid);
    if (valid == 1) {
        int above = (y > 0) ? im.depthI(x, y - 1) : 0;
        int left = (x > 0) ? im.depthI(x - 1, y) : 0;
        int aboveLeft = (x > 0 && y > 0) ? im.depthI(x - 1, y - 1) : 0;
        return im.depthI(x, y) - (left + above - aboveLeft);
    }
    return -1;
}

// ---------------------------------------------------------------------------
// Features and responses: XY::normalize (random_tree_image.h:82-87),
// averageRegionColor/Depth (:222-316), calculateColor/DepthFeature (:564-623).
// Feature as 11 int8 in the SoA order of ImageFeaturesAndThresholds
// (random_tree_image.h:702-773): type, o1x, o1y, o2x, o2y, r1x, r1y, r2x, r2y,
// c1, c2.
// ---------------------------------------------------------------------------
enum { F_TYPE = 0, F_O1X, F_O1Y, F_O2X, F_O2Y, F_R1X, F_R1Y, F_R2X, F_R2Y, F_C1, F_C2 };
enum { DEPTH = 0, COLOR = 1 };

struct Feature {
    int type, o1x, o1y, r1x, r1y, c1, o2x, o2y, r2x, r2y, c2;
};

static inline int normalizeCoord(int v, float depthF) {
    return static_cast<int>(v / depthF);
}

static double avgRegionColor(const Image& im, int px, int py, int ox, int oy, int rx, int ry, int ch) {
    const int w = std::max(1, rx);
    const int h = std::max(1, ry);
    int x = px + ox, y = py + oy;
    int lx = x - w, rxx = x + w, uy = y - h, ly = y + h;
    if (lx < 0 || rxx >= im.W || uy < 0 || ly >= im.H) return std::numeric_limits<double>::quiet_NaN();
    double lowerRight = im.color(ch, rxx, ly);
    double lowerLeft = im.color(ch, lx, ly);
    double upperRight = im.color(ch, rxx, uy);
    double upperLeft = im.color(ch, lx, uy);
    return (lowerRight - upperRight) + (upperLeft - lowerLeft);
}

static double avgRegionDepth(const Image& im, int px, int py, int ox, int oy, int rx, int ry) {
    const int w = std::max(1, rx);
    const int h = std::max(1, ry);
    int x = px + ox, y = py + oy;
    int lx = x - w, rxx = x + w, uy = y - h, ly = y + h;
    if (lx < 0 || rxx >= im.W || uy < 0 || ly >= im.H) return std::numeric_limits<double>::quiet_NaN();
    int numValid = (im.validI(rxx, ly) - im.validI(rxx, uy)) + (im.validI(lx, uy) - im.validI(lx, ly));
    if (numValid == 0) return std::numeric_limits<double>::quiet_NaN();
    int sum = (im.depthI(rxx, ly) - im.depthI(rxx, uy)) + (im.depthI(lx, uy) - im.depthI(lx, ly));
    double feat = sum / static_cast<double>(1000);
    return feat / numValid;
}

static double response(const Image& im, const Feature& f, int px, int py, int32_t depthMm, bool flip) {
    if (!(depthMm > 0)) return std::numeric_limits<double>::quiet_NaN();
    const float d = depthMm / 1000.0f;  // Depth::getFloatValue, image.h:143-146
    const int o1x = flip ? -f.o1x : f.o1x;
    const int o2x = flip ? -f.o2x : f.o2x;
    double a, b;
    if (f.type == DEPTH) {
        a = avgRegionDepth(im, px, py, normalizeCoord(o1x, d), normalizeCoord(f.o1y, d),
                normalizeCoord(f.r1x, d), normalizeCoord(f.r1y, d));
        if (std::isnan(a)) return a;
        b = avgRegionDepth(im, px, py, normalizeCoord(o2x, d), normalizeCoord(f.o2y, d),
                normalizeCoord(f.r2x, d), normalizeCoord(f.r2y, d));
        if (std::isnan(b)) return b;
    } else {
        a = avgRegionColor(im, px, py, normalizeCoord(o1x, d), normalizeCoord(f.o1y, d),
                normalizeCoord(f.r1x, d), normalizeCoord(f.r1y, d), f.c1);
        if (std::isnan(a)) return a;
        b = avgRegionColor(im, px, py, normalizeCoord(o2x, d), normalizeCoord(f.o2y, d),
                normalizeCoord(f.r2x, d), normalizeCoord(f.r2y, d), f.c2);
        if (std::isnan(b)) return b;
    }
    return a - b;
}

static Feature featureFromSoA(const int8_t* feats, int F, int i) {
    Feature f;
    f.type = feats[F_TYPE * F + i];
    f.o1x = feats[F_O1X * F + i];
    f.o1y = feats[F_O1Y * F + i];
    f.o2x = feats[F_O2X * F + i];
    f.o2y = feats[F_O2Y * F + i];
    f.r1x = feats[F_R1X * F + i];
    f.r1y = feats[F_R1Y * F + i];
    f.r2x = feats[F_R2X * F + i];
    f.r2y = feats[F_R2Y * F + i];
    f.c1 = static_cast<uint8_t>(feats[F_C1 * F + i]);
    f.c2 = static_cast<uint8_t>(feats[F_C2 * F + i]);
    return f;
}

static void featureToSoA(const Feature& f, int8_t* feats, int F, int i) {
    feats[F_TYPE * F + i] = static_cast<int8_t>(f.type);
    feats[F_O1X * F + i] = static_cast<int8_t>(f.o1x);
    feats[F_O1Y * F + i] = static_cast<int8_t>(f.o1y);
    feats[F_O2X * F + i] = static_cast<int8_t>(f.o2x);
    feats[F_O2Y * F + i] = static_cast<int8_t>(f.o2y);
    feats[F_R1X * F + i] = static_cast<int8_t>(f.r1x);
    feats[F_R1Y * F + i] = static_cast<int8_t>(f.r1y);
    feats[F_R2X * F + i] = static_cast<int8_t>(f.r2x);
    feats[F_R2Y * F + i] = static_cast<int8_t>(f.r2y);
    feats[F_C1 * F + i] = static_cast<int8_t>(f.c1);
    feats[F_C2 * F + i] = static_cast<int8_t>(f.c2);
}

// getSortKey: random_tree_image.h:444-452 (values re-read through int8 storage).
static int sortKey(const Feature& f) {
    int32_t k = 0;
    k |= static_cast<uint8_t>(f.type & 0x03) << 30;
    k |= static_cast<uint8_t>(f.c1 & 0x0F) << 26;
    k |= static_cast<uint8_t>(f.c2 & 0x0F) << 22;
    k |= static_cast<uint8_t>((f.o1y + 127) & 0xFF) << 14;
    k |= static_cast<uint8_t>((f.o1x + 127) & 0xFF) << 6;
    return k;
}

// sampleFeature: random_tree_image.cpp:648-700.  `argOrder` selects the
// evaluation order of the two getNext() calls inside one constructor call
// (bit0: offsets, bit1: regions, bit2: colour channels; 1 = second argument is
// drawn first, which is what g++ on x86-64 does).  The golden hash decides it.
static Feature sampleFeature(RandomSource& rs, int boxRadius, int regionSize, bool useDepthImages, int argOrder) {
    Sampler rgen_ft = rs.uniformSampler(2);
    Sampler rgen_xy = rs.uniformSampler(-boxRadius, boxRadius);
    Sampler rgen_region = rs.uniformSampler(1, regionSize);
    Sampler rgen_cc = rs.uniformSampler(3);

    auto pair = [](Sampler& s, bool reversed, int& a, int& b) {
        if (reversed) { b = s.getNext(); a = s.getNext(); }
        else { a = s.getNext(); b = s.getNext(); }
    };
    Feature f{};
    pair(rgen_xy, argOrder & 1, f.o1x, f.o1y);
    do {
        pair(rgen_xy, argOrder & 1, f.o2x, f.o2y);
    } while (f.o1x == f.o2x && f.o1y == f.o2y);
    pair(rgen_region, argOrder & 2, f.r1x, f.r1y);
    pair(rgen_region, argOrder & 2, f.r2x, f.r2y);
    if (useDepthImages) {
        int t = rgen_ft.getNext();
        if (t == 0) {
            f.type = DEPTH;
            f.c1 = 0;
            f.c2 = 0;
            return f;
        }
    }
    f.type = COLOR;
    pair(rgen_cc, argOrder & 4, f.c1, f.c2);
    return f;
}

// generateRandomFeatures: random_tree_image.cpp:577-646; sortFeatures:
// random_tree_image.cpp:504-566 (stable sort by int key, then gather).
// Returns 0 on success, -1 when the reference throws "max loops exceeded".
static int generateFeatures(const Dataset& ds, const std::vector<Sample>& samples, int seed, int F, int T,
        int boxRadius, int regionSize, bool useDepthImages, int argOrder, bool sort,
        int8_t* outFeats, float* outThr) {
    RandomSource rs(seed);
    Sampler sampleGen = rs.uniformSampler(static_cast<int>(samples.size()));
    std::vector<Feature> feats;
    std::vector<float> thr(static_cast<size_t>(T) * F);
    std::vector<int> keys;
    int generated = 0;
    int maxLoops = 3 * F;
    while (generated < F) {
        Feature f = sampleFeature(rs, boxRadius, regionSize, useDepthImages, argOrder);
        if (--maxLoops <= 0) return -1;
        bool valid = true;
        for (int t = 0; t < T; t++) {
            int maxTries = 10;
            float threshold;
            do {
                const Sample& s = samples.at(sampleGen.getNext());
                threshold = static_cast<float>(response(ds.images[s.image], f, s.x, s.y, s.depth_mm, false));
            } while (std::isnan(threshold) && --maxTries > 0);
            if (std::isnan(threshold)) { valid = false; break; }
            thr[static_cast<size_t>(t) * F + generated] = threshold;
        }
        if (!valid) continue;
        feats.push_back(f);
        keys.push_back(sortKey(f));
        generated++;
    }
    std::vector<int> order(F);
    for (int i = 0; i < F; i++) order[i] = i;
    if (sort) {
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    }
    for (int i = 0; i < F; i++) {
        featureToSoA(feats[order[i]], outFeats, F, i);
        for (int t = 0; t < T; t++) outThr[static_cast<size_t>(t) * F + i] = thr[static_cast<size_t>(t) * F + order[i]];
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Counters, scores, argmax.
// FeatureEvaluationCPU::operator(): random_tree_image.cpp:68-157 (counter index
// [label][feature][threshold][lr], lr = !(r <= t)).
// ---------------------------------------------------------------------------
static void nodeCountersRange(const Dataset& ds, const Sample* samples, size_t n, const std::vector<Feature>& ff,
        const float* thr, int F, int T, uint32_t* counters) {
    for (size_t k = 0; k < n; k++) {
        const Sample& s = samples[k];
        const Image& im = ds.images[s.image];
        for (int f = 0; f < F; f++) {
            double v1 = 0, v2 = 0;
            const int flipSetting = s.flip;  // 0 NoFlip, 1 Flip, 2 Both
            if (flipSetting == 1) v1 = response(im, ff[f], s.x, s.y, s.depth_mm, true);
            else v1 = response(im, ff[f], s.x, s.y, s.depth_mm, false);
            if (flipSetting == 2) v2 = response(im, ff[f], s.x, s.y, s.depth_mm, true);
            for (int t = 0; t < T; t++) {
                const float th = thr[static_cast<size_t>(t) * F + f];
                int lr = static_cast<int>(!(v1 <= th));
                counters[((static_cast<size_t>(s.label) * F + f) * T + t) * 2 + lr] += 1;
                if (flipSetting == 2) {
                    lr = static_cast<int>(!(v2 <= th));
                    counters[((static_cast<size_t>(s.label) * F + f) * T + t) * 2 + lr] += 1;
                }
            }
        }
    }
}

// Host threads for the counters (ORC_THREADS, default: the hardware threads, at most 16).
static int oracleThreads() {
    if (const char* e = std::getenv("ORC_THREADS")) return std::max(1, std::atoi(e));
    const unsigned hw = std::thread::hardware_concurrency();
    return static_cast<int>(std::min(16u, std::max(1u, hw)));
}

// The samples are split into chunks of at least 100 (the reference's grain,
// random_tree_image.cpp:343), each thread counts into its own counters and the
// counters are summed afterwards (random_tree_image.cpp:362-371).  Integer
// sums: the result does not depend on the split.
static void nodeCounters(const Dataset& ds, const std::vector<Sample>& samples, const int8_t* feats,
        const float* thr, int F, int T, int C, uint32_t* counters) {
    const size_t total = static_cast<size_t>(C) * F * T * 2;
    std::fill(counters, counters + total, 0u);
    std::vector<Feature> ff(F);
    for (int i = 0; i < F; i++) ff[i] = featureFromSoA(feats, F, i);
    const size_t n = samples.size();
    const size_t threads = static_cast<size_t>(oracleThreads());
    const size_t grain = std::max<size_t>(100, (n + threads - 1) / threads);
    const size_t parts = std::max<size_t>(1, (n + grain - 1) / grain);
    if (parts == 1 || static_cast<size_t>(F) * n < 200000) {
        nodeCountersRange(ds, samples.data(), n, ff, thr, F, T, counters);
        return;
    }
    std::vector<std::vector<uint32_t>> local(parts, std::vector<uint32_t>(total, 0u));
    std::vector<std::thread> pool;
    for (size_t p = 0; p < parts; p++) {
        const size_t b = p * grain, e = std::min(n, b + grain);
        pool.emplace_back([&, p, b, e] { nodeCountersRange(ds, samples.data() + b, e - b, ff, thr, F, T, local[p].data()); });
    }
    for (auto& t : pool) t.join();
    for (size_t p = 0; p < parts; p++) {
        for (size_t k = 0; k < total; k++) counters[k] += local[p][k];
    }
}

// score.h:53-58
static double entropy(double p) {
    if (p == 0.0) return 0.0;
    return -p * std::log2(p);
}

// score.h:64-80
static double normalizeScore(double s) {
    if (s > 1.0) return 1;
    if (s < 0.0) return 0;
    return s;
}

// InformationGainScore::calculateScore, score.h:87-139.
template<class W>
static double informationGain(int C, const W* left, const W* right, const W* all, double totalLeft,
        double totalRight) {
    const double total = totalLeft + totalRight;
    const double leftProb = totalLeft / total;
    const double rightProb = totalRight / total;
    double score = 0;
    for (int c = 0; c < C; c++) {
        const double classProb = all[c] / total;
        if (left[c] > 0) {
            double p = left[c] / total;
            score += p * std::log2(p / (leftProb * classProb));
        }
        if (right[c] > 0) {
            double p = right[c] / total;
            score += p * std::log2(p / (rightProb * classProb));
        }
    }
    return normalizeScore(score);
}

// NormalizedInformationGainScore::calculateScore, score.h:191-215.
template<class W>
static double normalizedInformationGain(int C, const W* left, const W* right, const W* all, double totalLeft,
        double totalRight) {
    const double ig = informationGain(C, left, right, all, totalLeft, totalRight);
    if (ig == 0) return 0;
    const double total = totalLeft + totalRight;
    const double Hs = (entropy(totalLeft) + entropy(totalRight) - entropy(total)) / total;
    double Hc = 0;
    for (int c = 0; c < C; c++) {
        if (all[c] > 0) Hc += entropy(all[c]);
    }
    Hc -= entropy(total);
    Hc /= total;
    const double score = (2 * ig) / (Hs + Hc);
    return normalizeScore(score);
}

// ImageFeatureEvaluation::calculateScores (CPU): random_tree_image.cpp:188-237.
static void scoresFromCounters(const uint32_t* counters, const uint32_t* hist, int C, int F, int T,
        double* scores) {
    std::vector<uint32_t> l(C), r(C);
    for (int f = 0; f < F; f++) {
        for (int t = 0; t < T; t++) {
            double sl = 0, sr = 0;
            for (int c = 0; c < C; c++) {
                l[c] = counters[((static_cast<size_t>(c) * F + f) * T + t) * 2 + 0];
                r[c] = counters[((static_cast<size_t>(c) * F + f) * T + t) * 2 + 1];
                sl += l[c];
                sr += r[c];
            }
            scores[static_cast<size_t>(t) * F + f] = normalizedInformationGain(C, l.data(), r.data(), hist, sl, sr);
        }
    }
}

// detail::isScoreBetter, random_tree.cpp:36-52.
static bool isScoreBetter(double best, double score) {
    double diff = std::fabs(score - best);
    if (diff < 1e-13) return false;
    return score > best;
}

// argmax scan, random_tree_image.cpp:445-457.
static void argmax(const double* scores, int T, int F, int& bestT, int& bestF, double& bestScore) {
    bestScore = -std::numeric_limits<double>::infinity();
    bestT = 0;
    bestF = 0;
    for (int t = 0; t < T; t++) {
        for (int f = 0; f < F; f++) {
            const double s = scores[static_cast<size_t>(t) * F + f];
            if (std::isnan(bestScore) || isScoreBetter(bestScore, s)) {
                bestF = f;
                bestT = t;
                bestScore = s;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Training configuration (the subset the hot path reads), random_tree.h:187-493.
// ---------------------------------------------------------------------------
struct Config {
    int randomSeed;
    int samplesPerImage;
    int featureCount;
    int minSampleCount;
    int maxDepth;
    int boxRadius;
    int regionSize;
    int thresholds;
    int numThreads;
    int maxImages;
    int useDepthImages;
    int horizontalFlipping;
    int subsamplingPixelUniform;  // 0 = classUniform
    int numIgnored;
    int ignored[8];               // label ids to ignore
};

// Node table record shared with the C ABI's cf_tree_node.
struct NodeOut {
    int32_t id, level, parent, left, right;
    int32_t numSamples;   // number of training samples that reached the node
    int32_t featureId;    // index of the split feature in the level's sorted list
    float threshold;
    double score;
    int8_t feature[11];
    uint8_t isLeaf;
    int32_t pad;
};
static_assert(sizeof(NodeOut) == 56, "node record layout");

struct TreeBuild {
    std::vector<NodeOut> nodes;
    std::vector<uint32_t> hist;  // [nodes][C]
};

// RandomTree node histogram, random_tree.h:521-527 (Both counted twice).
static void nodeHistogram(const std::vector<Sample>& s, int C, uint32_t* h) {
    std::fill(h, h + C, 0u);
    for (const Sample& x : s) {
        h[x.label] += 1;
        if (x.flip == 2) h[x.label] += 1;
    }
}

static bool hasPureHistogram(const uint32_t* h, int C) {
    int nz = 0;
    for (int c = 0; c < C; c++) {
        if (h[c] == 0) continue;
        if (++nz > 1) return false;
    }
    return nz == 1;
}

struct FrontierNode {
    int index;                    // index in the node table
    std::vector<Sample> samples;
};

// One tree: RandomTreeImage::doTrain (random_tree_image.cpp:714-730) +
// RandomTreeTrain::train level loop (random_tree.h:1249-1415) +
// evaluateBestSplits CPU branch (random_tree_image.cpp:239-502), horizontal
// flipping included.
static int trainTree(const Dataset& ds, int treeId, const std::vector<Sample>& subsamples, int C,
        const Config& cfg, RandomSource& rs, int argOrder, TreeBuild& out) {
    out.nodes.clear();
    out.hist.clear();
    NodeOut root{};
    root.id = treeId;
    root.level = 1;
    root.parent = -1;
    root.left = root.right = -1;
    root.numSamples = static_cast<int32_t>(subsamples.size());
    root.isLeaf = 1;
    out.nodes.push_back(root);
    out.hist.resize(C);
    nodeHistogram(subsamples, C, out.hist.data());

    std::vector<FrontierNode> frontier;
    frontier.push_back(FrontierNode{0, subsamples});
    int idNode = treeId;
    const int F = cfg.featureCount, T = cfg.thresholds;
    std::vector<int8_t> feats(static_cast<size_t>(11) * F);
    std::vector<float> thr(static_cast<size_t>(T) * F);
    std::vector<uint32_t> counters(static_cast<size_t>(C) * F * T * 2);
    std::vector<double> scores(static_cast<size_t>(T) * F);

    for (int level = 1; !frontier.empty(); level++) {
        if (level == cfg.maxDepth) break;
        // evaluateBestSplits: all samples of the level, node order, then one
        // seed draw from the tree's random source.
        std::vector<Sample> all;
        for (const auto& n : frontier) all.insert(all.end(), n.samples.begin(), n.samples.end());
        const int seed = rs.uniformSampler(0xFFFFFF).getNext();
        if (generateFeatures(ds, all, seed, F, T, cfg.boxRadius, cfg.regionSize, cfg.useDepthImages != 0,
                argOrder, true, feats.data(), thr.data()) != 0) {
            return -1;
        }
        struct Best { int t, f; double s; };
        std::vector<Best> best(frontier.size());
        for (size_t i = 0; i < frontier.size(); i++) {
            const auto& n = frontier[i];
            nodeCounters(ds, n.samples, feats.data(), thr.data(), F, T, C, counters.data());
            scoresFromCounters(counters.data(), &out.hist[static_cast<size_t>(n.index) * C], C, F, T, scores.data());
            argmax(scores.data(), T, F, best[i].t, best[i].f, best[i].s);
        }
        std::vector<FrontierNode> next;
        for (size_t i = 0; i < frontier.size(); i++) {
            auto& n = frontier[i];
            const Feature f = featureFromSoA(feats.data(), F, best[i].f);
            const float th = thr[static_cast<size_t>(best[i].t) * F + best[i].f];
            // SplitFunction::split and the partition loop (random_tree.h:122-144, 1295-1360): a Both
            // sample whose flipped response falls on the other side continues as NoFlip on its side,
            // and a Flip copy of it joins the other side at the same point of that side's order
            std::vector<Sample> L, R;
            for (const Sample& s : n.samples) {
                const Image& im = ds.images[s.image];
                const bool left = response(im, f, s.x, s.y, s.depth_mm, s.flip == 1) <= th;
                if (s.flip == 2 && (response(im, f, s.x, s.y, s.depth_mm, true) <= th) != left) {
                    Sample self = s, copy = s;
                    self.flip = 0;
                    copy.flip = 1;
                    (left ? L : R).push_back(self);
                    (left ? R : L).push_back(copy);
                    continue;
                }
                (left ? L : R).push_back(s);
            }
            const int leftId = ++idNode;
            const int rightId = ++idNode;
            if (!L.empty() && !R.empty()) {
                NodeOut& p = out.nodes[n.index];
                p.isLeaf = 0;
                p.featureId = best[i].f;
                p.threshold = th;
                p.score = best[i].s;
                for (int k = 0; k < 11; k++) p.feature[k] = feats[static_cast<size_t>(k) * F + best[i].f];
                const int li = static_cast<int>(out.nodes.size());
                const int childLevel = p.level + 1;  // read before push_back invalidates `p`
                for (int side = 0; side < 2; side++) {
                    NodeOut c{};
                    c.id = side == 0 ? leftId : rightId;
                    c.level = childLevel;
                    c.parent = n.index;
                    c.left = c.right = -1;
                    c.isLeaf = 1;
                    c.numSamples = static_cast<int32_t>(side == 0 ? L.size() : R.size());
                    out.nodes.push_back(c);
                    out.hist.resize(out.hist.size() + C);
                    nodeHistogram(side == 0 ? L : R, C, &out.hist[out.hist.size() - C]);
                }
                out.nodes[n.index].left = li;
                out.nodes[n.index].right = li + 1;
                for (int side = 0; side < 2; side++) {
                    const int ci = li + side;
                    const uint32_t* h = &out.hist[static_cast<size_t>(ci) * C];
                    const auto& S = side == 0 ? L : R;
                    if (!hasPureHistogram(h, C) && S.size() >= static_cast<size_t>(cfg.minSampleCount)) {
                        next.push_back(FrontierNode{ci, S});
                    }
                }
            } else {
                idNode -= 2;
            }
        }
        frontier.swap(next);
    }
    return 0;
}

// calculateLabelPriorDistribution: random_tree_image.cpp:885-909, over the
// images `ids`.  The loop `for (label = 0; label < map.size(); label++)` reads
// map[label], which inserts missing labels and so runs up to the largest label.
static std::vector<uint32_t> labelPrior(const Dataset& ds, const std::vector<int>& ids, int numLabels) {
    std::map<int, size_t> m;
    for (int i : ids) {
        const Image& im = ds.images[i];
        for (int y = 0; y < im.H; ++y)
            for (int x = 0; x < im.W; ++x) m[im.lab[static_cast<size_t>(y) * im.W + x]]++;
    }
    std::vector<uint32_t> prior(std::max<size_t>(numLabels, m.size()), 0);
    for (size_t l = 0; l < m.size(); l++) {
        if (l >= prior.size()) throw std::runtime_error("label does not fit the classes");
        prior[l] = static_cast<uint32_t>(m[static_cast<int>(l)]);
    }
    return prior;
}

static std::vector<uint32_t> labelPrior(const Dataset& ds, int numLabels) {
    std::vector<int> ids(ds.images.size());
    for (size_t i = 0; i < ids.size(); i++) ids[i] = static_cast<int>(i);
    return labelPrior(ds, ids, numLabels);
}

// RandomForestImage::train's maxImages subsampling (random_forest_image.cpp:161-171).
static std::vector<int> sampleImages(int count, int maxImages, RandomSource& rs) {
    std::vector<int> ids;
    if (maxImages > 0 && count > maxImages) {
        Reservoir<int> r(static_cast<size_t>(maxImages));
        Sampler sampler = rs.uniformSampler(0, 10 * count);
        for (int i = 0; i < count; i++) r.sample(sampler, i);
        ids = r.reservoir;
    } else {
        for (int i = 0; i < count; i++) ids.push_back(i);
    }
    return ids;
}

// block sort of RandomTreeImage::train (random_tree_image.cpp:776-802), image
// index instead of image address
static void sortRoots(std::vector<Sample>& all) {
    const size_t BLOCK = 200000;
    for (size_t b = 0; b < all.size(); b += BLOCK) {
        auto e = std::min(all.size(), b + BLOCK);
        std::sort(all.begin() + b, all.begin() + e, [](const Sample& a, const Sample& c) {
            if (a.image != c.image) return a.image < c.image;
            if (a.label != c.label) return a.label < c.label;
            if (a.y / 10 != c.y / 10) return a.y / 10 < c.y / 10;
            if (a.x / 10 != c.x / 10) return a.x / 10 < c.x / 10;
            return false;
        });
    }
}

// subsampleTrainingDataPixelUniform: random_tree_image.cpp:840-883, then the block sort.
static int samplePixelUniform(const Dataset& ds, const std::vector<int>& ids, const Config& cfg, RandomSource& rs,
        size_t subsampleCount, std::vector<Sample>& out) {
    auto ignored = [&](int l) {
        for (int i = 0; i < cfg.numIgnored; i++) if (cfg.ignored[i] == l) return true;
        return false;
    };
    const uint8_t flip = cfg.horizontalFlipping ? 2 : 0;
    Sampler rgenImage = rs.uniformSampler(static_cast<int>(ids.size()));
    std::vector<Sample> all;
    for (size_t n = 0; n < subsampleCount * ids.size(); ++n) {
        const int imageId = ids[rgenImage.getNext()];
        const Image& im = ds.images[imageId];
        Sampler rgenX = rs.uniformSampler(im.W);
        Sampler rgenY = rs.uniformSampler(im.H);
        for (int tries = 0;; tries++) {
            if (tries > 100000000) return -5;
            const int x = rgenX.getNext(), y = rgenY.getNext();
            const int label = im.lab[static_cast<size_t>(y) * im.W + x];
            if (ignored(label)) continue;
            Sample s{};
            s.x = static_cast<uint16_t>(x);
            s.y = static_cast<uint16_t>(y);
            s.image = static_cast<uint32_t>(imageId);
            s.depth_mm = pixelDepth(im, x, y);
            s.label = static_cast<uint8_t>(label);
            s.flip = flip;
            all.push_back(s);
            break;
        }
    }
    sortRoots(all);
    out.swap(all);
    return 0;
}

// subsampleTrainingDataClassUniform with one TBB chunk:
// random_tree_image.cpp:911-1059; then the block sort by
// (image, label, y/10, x/10) of random_tree_image.cpp:776-802 (image index
// instead of image address).
static int sampleClassUniform(const Dataset& ds, const std::vector<int>& ids, const std::vector<uint32_t>& prior,
        const Config& cfg, RandomSource& rs, size_t subsampleCount, std::vector<Sample>& out) {
    const size_t numLabelsAll = prior.size();
    size_t numLabels = numLabelsAll - static_cast<size_t>(cfg.numIgnored);
    const size_t spc = static_cast<size_t>(std::ceil(ids.size() * static_cast<double>(subsampleCount) /
            static_cast<double>(numLabels)));
    const int randomSeed = rs.uniformSampler(0xFFFFFF).getNext();
    auto ignored = [&](int l) {
        for (int i = 0; i < cfg.numIgnored; i++) if (cfg.ignored[i] == l) return true;
        return false;
    };
    const uint8_t flip = cfg.horizontalFlipping ? 2 : 0;
    std::vector<Reservoir<Sample>> chunk;
    for (size_t l = 0; l < numLabelsAll; l++) chunk.emplace_back(spc);
    {
        RandomSource crs(randomSeed);
        Sampler sampler = crs.uniformSampler(0x7FFFFFFF);
        for (int i : ids) {
            const Image& im = ds.images[i];
            for (int y = 0; y < im.H; y++) {
                for (int x = 0; x < im.W; x++) {
                    const int label = im.lab[static_cast<size_t>(y) * im.W + x];
                    if (ignored(label)) continue;
                    const int32_t d = pixelDepth(im, x, y);
                    if (!(d > 0)) continue;
                    Sample s{};
                    s.x = static_cast<uint16_t>(x);
                    s.y = static_cast<uint16_t>(y);
                    s.image = static_cast<uint32_t>(i);
                    s.depth_mm = d;
                    s.label = static_cast<uint8_t>(label);
                    s.flip = flip;
                    chunk[label].sample(sampler, s);
                }
            }
        }
    }
    std::vector<Sample> all;
    Sampler sampler = rs.uniformSampler(0x7FFFFFFF);
    for (size_t l = 0; l < numLabelsAll; l++) {
        Reservoir<Sample> merged(spc);
        if (!ignored(static_cast<int>(l))) {
            for (const Sample& s : chunk[l].reservoir) merged.sample(sampler, s);
        }
        all.insert(all.begin(), merged.reservoir.begin(), merged.reservoir.end());
    }
    if (all.size() != spc * numLabels) return -2;
    sortRoots(all);
    out.swap(all);
    return 0;
}

// normalizeHistogram: random_tree.cpp:54-104.
static void normalizeHistogram(const uint32_t* h, const uint32_t* prior, int C, double bias, bool usePrior,
        double* out) {
    std::vector<double> w(C);
    double sum = 0;
    for (int i = 0; i < C; i++) sum += prior[i];
    for (int i = 0; i < C; i++) w[i] = prior[i] / sum;
    sum = 0;
    for (int i = 0; i < C; i++) sum += h[i];
    for (int i = 0; i < C; i++) {
        out[i] = h[i] - bias * sum;
        if (out[i] < 0.0) out[i] = 0.0;
    }
    sum = 0;
    for (int i = 0; i < C; i++) sum += out[i];
    if (sum > 0) {
        for (int i = 0; i < C; i++) out[i] = usePrior ? out[i] / sum * w[i] : out[i] / sum;
    }
}

}  // namespace orc

using namespace orc;

extern "C" {

// -- RNG pins ---------------------------------------------------------------
uint32_t orc_mt19937_nth(uint32_t seed, int n) {
    MT19937 m(seed);
    uint32_t v = 0;
    for (int i = 0; i < n; i++) v = m.next();
    return v;
}

void orc_sampler_draws(int seed, int lo, int hi, int n, int* out) {
    Sampler s(seed, lo, hi);
    for (int i = 0; i < n; i++) out[i] = s.getNext();
}

// -- integrals ----------------------------------------------------------------
// ReservoirSampler<int> (random_tree.h:1069-1122) of the values 0..n-1 with
// RandomSource(seed).uniformSampler(lo, hi), as random_tree_test.cpp:156-209 runs it.
int orc_reservoir(int seed, int lo, int hi, int n, int k, int* out) {
    RandomSource rs(seed);
    Sampler sampler = rs.uniformSampler(lo, hi);
    Reservoir<int> r(static_cast<size_t>(k));
    for (int i = 0; i < n; i++) r.sample(sampler, i);
    std::copy(r.reservoir.begin(), r.reservoir.end(), out);
    return static_cast<int>(r.reservoir.size());
}

void orc_integral_color(const float* in, int W, int H, float* out) { integrateColorChannel(in, W, H, out); }
void orc_integral_int(const int32_t* in, int W, int H, int32_t* out) { integrateIntChannel(in, W, H, out); }

// -- dataset ----------------------------------------------------------------
// color: [N][3][H][W] float (raw), depth: [N][H][W] int32 mm (<=0 invalid),
// labels: [N][H][W] uint8 (may be NULL).
void* orc_dataset_create(int N, int W, int H, const float* color, const int32_t* depth, const uint8_t* labels) {
    auto* ds = new Dataset();
    ds->images.resize(N);
    const size_t px = static_cast<size_t>(W) * H;
    std::vector<int32_t> d(px), v(px);
    for (int i = 0; i < N; i++) {
        Image& im = ds->images[i];
        im.W = W;
        im.H = H;
        im.col.resize(3 * px);
        for (int c = 0; c < 3; c++) {
            integrateColorChannel(color + (static_cast<size_t>(i) * 3 + c) * px, W, H, &im.col[c * px]);
        }
        for (size_t p = 0; p < px; p++) {
            const int32_t mm = depth[static_cast<size_t>(i) * px + p];
            d[p] = mm > 0 ? mm : 0;   // RGBDImage::setDepth, image.h:363-369
            v[p] = mm > 0 ? 1 : 0;
        }
        im.dep.resize(px);
        im.val.resize(px);
        integrateIntChannel(d.data(), W, H, im.dep.data());
        integrateIntChannel(v.data(), W, H, im.val.data());
        im.lab.assign(px, 0);
        if (labels) std::memcpy(im.lab.data(), labels + static_cast<size_t>(i) * px, px);
    }
    return ds;
}

void orc_dataset_free(void* p) { delete static_cast<Dataset*>(p); }

void orc_dataset_integrals(void* p, int img, float* col, int32_t* dep, int32_t* val) {
    const Image& im = static_cast<Dataset*>(p)->images[img];
    std::memcpy(col, im.col.data(), im.col.size() * sizeof(float));
    std::memcpy(dep, im.dep.data(), im.dep.size() * sizeof(int32_t));
    std::memcpy(val, im.val.data(), im.val.size() * sizeof(int32_t));
}

int32_t orc_pixel_depth(void* p, int img, int x, int y) {
    return pixelDepth(static_cast<Dataset*>(p)->images[img], x, y);
}

// responses for n samples x F features -> out[F][n] (the reference GPU test layout)
void orc_responses(void* p, const Sample* samples, int n, const int8_t* feats, int F, double* out) {
    const Dataset& ds = *static_cast<Dataset*>(p);
    for (int f = 0; f < F; f++) {
        const Feature ft = featureFromSoA(feats, F, f);
        for (int i = 0; i < n; i++) {
            const Sample& s = samples[i];
            out[static_cast<size_t>(f) * n + i] = response(ds.images[s.image], ft, s.x, s.y, s.depth_mm, s.flip == 1);
        }
    }
}

int orc_generate_features(void* p, const Sample* samples, int n, int seed, int F, int T, int boxRadius,
        int regionSize, int useDepthImages, int argOrder, int sort, int8_t* feats, float* thr) {
    std::vector<Sample> s(samples, samples + n);
    return generateFeatures(*static_cast<Dataset*>(p), s, seed, F, T, boxRadius, regionSize, useDepthImages != 0,
            argOrder, sort != 0, feats, thr);
}

void orc_node_counters(void* p, const Sample* samples, int n, const int8_t* feats, const float* thr, int F, int T,
        int C, uint32_t* counters) {
    std::vector<Sample> s(samples, samples + n);
    nodeCounters(*static_cast<Dataset*>(p), s, feats, thr, F, T, C, counters);
}

void orc_scores(const uint32_t* counters, const uint32_t* hist, int C, int F, int T, double* scores) {
    scoresFromCounters(counters, hist, C, F, T, scores);
}

double orc_score(int C, const uint32_t* left, const uint32_t* right, const uint32_t* all) {
    double sl = 0, sr = 0;
    for (int c = 0; c < C; c++) { sl += left[c]; sr += right[c]; }
    return normalizedInformationGain(C, left, right, all, sl, sr);
}

double orc_information_gain(int C, const uint32_t* left, const uint32_t* right, const uint32_t* all) {
    double sl = 0, sr = 0;
    for (int c = 0; c < C; c++) { sl += left[c]; sr += right[c]; }
    return informationGain(C, left, right, all, sl, sr);
}

void orc_argmax(const double* scores, int T, int F, int* bestT, int* bestF, double* bestScore) {
    argmax(scores, T, F, *bestT, *bestF, *bestScore);
}

void orc_normalize_histogram(const uint32_t* h, const uint32_t* prior, int C, double bias, int usePrior, double* out) {
    normalizeHistogram(h, prior, C, bias, usePrior != 0, out);
}

int orc_label_prior(void* p, int numLabels, uint32_t* prior, int cap) {
    auto v = labelPrior(*static_cast<Dataset*>(p), numLabels);
    for (int i = 0; i < static_cast<int>(v.size()) && i < cap; i++) prior[i] = v[i];
    return static_cast<int>(v.size());
}

// -- forest training ------------------------------------------------------------
// RandomForestImage::train (random_forest_image.cpp:132-182), trees trained
// sequentially.  Results: per tree a node table and histograms, plus the
// per-tree root samples (for the GPU path to start from identical samples).
struct OrcForest {
    std::vector<TreeBuild> trees;
    std::vector<std::vector<Sample>> roots;
    std::vector<uint32_t> prior;                     // over all images
    std::vector<std::vector<uint32_t>> treePriors;   // over each tree's images
    int C = 0;
};

void* orc_train_forest(void* p, const Config* cfg, int numTrees, int numLabels, int argOrder, int* status) {
    const Dataset& ds = *static_cast<Dataset*>(p);
    auto* fo = new OrcForest();
    *status = 0;
    fo->prior = labelPrior(ds, numLabels);
    fo->C = static_cast<int>(fo->prior.size());
    RandomSource master(cfg->randomSeed);
    const int SEED = master.uniformSampler(0xFFFF).getNext();
    fo->trees.resize(numTrees);
    fo->roots.resize(numTrees);
    fo->treePriors.resize(numTrees);
    for (int t = 0; t < numTrees; t++) {
        RandomSource rs(SEED + t);
        const std::vector<int> ids = sampleImages(static_cast<int>(ds.images.size()), cfg->maxImages, rs);
        fo->treePriors[t] = labelPrior(ds, ids, numLabels);
        if (fo->treePriors[t].size() != fo->prior.size()) { *status = -6; return fo; }
        std::vector<Sample> sub;
        const size_t perImage = cfg->samplesPerImage / numTrees;
        int st = cfg->subsamplingPixelUniform ? samplePixelUniform(ds, ids, *cfg, rs, perImage, sub)
                                              : sampleClassUniform(ds, ids, fo->treePriors[t], *cfg, rs, perImage, sub);
        if (st != 0) { *status = st; return fo; }
        fo->roots[t] = sub;
        st = trainTree(ds, t, sub, fo->C, *cfg, rs, argOrder, fo->trees[t]);
        if (st != 0) { *status = st; return fo; }
    }
    return fo;
}

void orc_forest_free(void* f) { delete static_cast<OrcForest*>(f); }
int orc_forest_num_classes(void* f) { return static_cast<OrcForest*>(f)->C; }
int orc_forest_tree_size(void* f, int t) { return static_cast<int>(static_cast<OrcForest*>(f)->trees[t].nodes.size()); }
void orc_forest_tree(void* f, int t, NodeOut* nodes, uint32_t* hist) {
    const auto& tb = static_cast<OrcForest*>(f)->trees[t];
    std::memcpy(nodes, tb.nodes.data(), tb.nodes.size() * sizeof(NodeOut));
    std::memcpy(hist, tb.hist.data(), tb.hist.size() * sizeof(uint32_t));
}
int orc_forest_root_count(void* f, int t) { return static_cast<int>(static_cast<OrcForest*>(f)->roots[t].size()); }
void orc_forest_roots(void* f, int t, Sample* out) {
    const auto& r = static_cast<OrcForest*>(f)->roots[t];
    std::memcpy(out, r.data(), r.size() * sizeof(Sample));
}
void orc_forest_prior(void* f, uint32_t* out) {
    const auto& p = static_cast<OrcForest*>(f)->prior;
    std::memcpy(out, p.data(), p.size() * sizeof(uint32_t));
}
void orc_forest_tree_prior(void* f, int t, uint32_t* out) {
    const auto& p = static_cast<OrcForest*>(f)->treePriors[t];
    std::memcpy(out, p.data(), p.size() * sizeof(uint32_t));
}

// -- single-tree training from given root samples (for level-by-level parity) --
int orc_train_tree(void* p, const Config* cfg, int treeId, int C, int rsSeed, const Sample* roots, int n,
        int argOrder, int* numNodes, NodeOut* nodes, uint32_t* hist, int cap) {
    RandomSource rs(rsSeed);
    std::vector<Sample> sub(roots, roots + n);
    TreeBuild tb;
    int st = trainTree(*static_cast<Dataset*>(p), treeId, sub, C, *cfg, rs, argOrder, tb);
    *numNodes = static_cast<int>(tb.nodes.size());
    if (st != 0) return st;
    if (static_cast<int>(tb.nodes.size()) > cap) return -4;
    std::memcpy(nodes, tb.nodes.data(), tb.nodes.size() * sizeof(NodeOut));
    std::memcpy(hist, tb.hist.data(), tb.hist.size() * sizeof(uint32_t));
    return 0;
}

// -- prediction: RandomForestImage::predict CPU branch (random_forest_image.cpp:307-345)
// trees: concatenated node tables (numNodes[t] each) with posteriors [node][C]
// (normalized histograms; only leaves are read).
void orc_predict(void* p, int img, int numTrees, const int* numNodes, const NodeOut* nodes, const double* post,
        int C, float* probs /*[C][H][W]*/, uint8_t* labels /*[H][W]*/) {
    const Image& im = static_cast<Dataset*>(p)->images[img];
    const int W = im.W, H = im.H;
    // rows are independent: split over host threads as the reference splits them over TBB tasks
    // (random_forest_image.cpp:310)
    auto rows = [&](int y0, int y1) {
        for (int y = y0; y < y1; y++) {
            for (int x = 0; x < W; x++) {
                std::vector<float> acc(C, 0.0f);
                const int32_t d = pixelDepth(im, x, y);
                size_t base = 0;
                for (int t = 0; t < numTrees; t++) {
                    const NodeOut* tn = nodes + base;
                    int i = 0;
                    while (!tn[i].isLeaf) {
                        Feature f;
                        const int8_t* q = tn[i].feature;
                        f.type = q[F_TYPE]; f.o1x = q[F_O1X]; f.o1y = q[F_O1Y]; f.o2x = q[F_O2X]; f.o2y = q[F_O2Y];
                        f.r1x = q[F_R1X]; f.r1y = q[F_R1Y]; f.r2x = q[F_R2X]; f.r2y = q[F_R2Y];
                        f.c1 = static_cast<uint8_t>(q[F_C1]); f.c2 = static_cast<uint8_t>(q[F_C2]);
                        const double r = response(im, f, x, y, d, false);
                        i = (r <= tn[i].threshold) ? tn[i].left : tn[i].right;
                    }
                    const double* h = post + (base + i) * C;
                    for (int c = 0; c < C; c++) acc[c] += h[c];   // float += double
                    base += numNodes[t];
                }
                double sum = 0.0f;
                for (int c = 0; c < C; c++) sum += acc[c];
                float bestProb = -1.0f;
                uint8_t lab = 0;
                for (int c = 0; c < C; c++) {
                    acc[c] /= sum;
                    const float prob = acc[c];
                    if (prob > bestProb) { lab = static_cast<uint8_t>(c); bestProb = prob; }
                }
                for (int c = 0; c < C; c++) probs[(static_cast<size_t>(c) * H + y) * W + x] = acc[c];
                labels[static_cast<size_t>(y) * W + x] = lab;
            }
        }
    };
    const int threads = std::min(oracleThreads(), H);
    const int per = (H + threads - 1) / threads;
    std::vector<std::thread> pool;
    for (int t = 0; t < threads; t++) {
        const int y0 = t * per, y1 = std::min(H, y0 + per);
        if (y0 < y1) pool.emplace_back(rows, y0, y1);
    }
    for (auto& t : pool) t.join();
}

// -- golden hash of feature_generation_test.cpp:310-387 ----------------------
static void hashUpdate(uint64_t& hash, uint64_t v) {
    hash ^= v + 0x9e3779b9ull + (hash << 6) + (hash >> 2);
}
// boost::hash<size_t> applied to a float: implicit float -> size_t conversion,
// which g++ on x86-64 emits as a signed 64-bit truncation for |v| < 2^63.
static uint64_t floatToSizeT(float v) {
    return static_cast<uint64_t>(static_cast<int64_t>(v));
}

uint64_t orc_feature_generation_hash(int argOrder, int8_t* featsOut, float* thrOut) {
    const int W = 400, H = 400, NIMG = 10, SPI = 500, F = 500, T = 50, NUM_LABELS = 10;
    RandomSource rs(4711);
    Sampler widthSampler = rs.uniformSampler(W);
    Sampler heightSampler = rs.uniformSampler(H);
    Sampler pixelSampler = rs.uniformSampler(1000);
    Sampler depthSampler = rs.uniformSampler(1, 90);
    const size_t px = static_cast<size_t>(W) * H;
    std::vector<float> color(static_cast<size_t>(NIMG) * 3 * px);
    std::vector<int32_t> depth(static_cast<size_t>(NIMG) * px);
    std::vector<Sample> samples;
    for (int i = 0; i < NIMG; i++) {
        for (int y = 0; y < H; y++) {
            for (int x = 0; x < W; x++) {
                for (int c = 0; c < 3; c++) {
                    color[(static_cast<size_t>(i) * 3 + c) * px + static_cast<size_t>(y) * W + x] =
                            pixelSampler.getNext() / 100.0f;
                }
                const float m = depthSampler.getNext() / 10.0f;
                const int mm = static_cast<int>(1000.0 * static_cast<double>(m));  // Depth(double), image.h:123-131
                depth[static_cast<size_t>(i) * px + static_cast<size_t>(y) * W + x] = mm;
            }
        }
    }
    Dataset* ds = static_cast<Dataset*>(orc_dataset_create(NIMG, W, H, color.data(), depth.data(), nullptr));
    // samples are drawn image by image after that image's pixels; the width and
    // height samplers are independent streams, so drawing them here is equivalent.
    for (int i = 0; i < NIMG; i++) {
        for (int s = 0; s < SPI; s++) {
            Sample smp{};
            smp.x = static_cast<uint16_t>(widthSampler.getNext());
            smp.y = static_cast<uint16_t>(heightSampler.getNext());
            smp.image = static_cast<uint32_t>(i);
            smp.depth_mm = pixelDepth(ds->images[i], smp.x, smp.y);
            smp.label = static_cast<uint8_t>(i % NUM_LABELS);
            samples.push_back(smp);
        }
    }
    std::vector<int8_t> feats(static_cast<size_t>(11) * F);
    std::vector<float> thr(static_cast<size_t>(T) * F);
    generateFeatures(*ds, samples, 9, F, T, 127, 16, true, argOrder, true, feats.data(), thr.data());
    delete ds;
    if (featsOut) std::memcpy(featsOut, feats.data(), feats.size());
    if (thrOut) std::memcpy(thrOut, thr.data(), thr.size() * sizeof(float));
    uint64_t hash = 0;
    const int order[11] = {F_TYPE, F_C1, F_C2, F_O1X, F_O1Y, F_O2X, F_O2Y, F_R1X, F_R1Y, F_R2X, F_R2Y};
    for (int k = 0; k < 11; k++) {
        for (int f = 0; f < F; f++) {
            int v = feats[static_cast<size_t>(order[k]) * F + f];
            if (order[k] == F_C1 || order[k] == F_C2) v = static_cast<uint8_t>(v);
            hashUpdate(hash, static_cast<uint64_t>(static_cast<int64_t>(v)));
        }
    }
    for (int f = 0; f < F; f++) {
        for (int t = 0; t < T; t++) hashUpdate(hash, floatToSizeT(thr[static_cast<size_t>(t) * F + f]));
    }
    return hash;
}

}  // extern "C"
