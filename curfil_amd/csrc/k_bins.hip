// Two-pass split evaluation of a whole level (batched training):
//
//   k_eval_bins   pass 1, image-major: the level's samples ordered by
//                 (image, tree, row) are cut into blocks of <= 256 within
//                 each (image, tree) segment; one workgroup per
//                 (block, group of 16 features), one sample per lane,
//                 computes every response's interval among the feature's
//                 thresholds (the bin used by the interval histograms) and
//                 stores the 16 bins of a sample as one 16-byte row,
//                 bins[group][level index].
//   k_hist_bins   pass 2, node-major: one workgroup per (node chunk, group)
//                 streams the chunk's bin rows and labels into LDS interval
//                 histograms and adds them to the node's global bins.
//
// Counting is order-independent, so the histograms equal those of the direct
// path (k_eval_hist), whose single kernel gathers the integrals in node order:
// at deep levels a node's samples come from every image and the gathers miss
// in L2.  In pass 1 a workgroup's samples come from one image, the
// workgroups of one XCD sweep one image (the segments of all trees, tree by
// tree within each feature group) for all feature groups of a batch before
// the next, and every tree visits its features in an order grouped by the
// integral planes they read (depth, then colour channel pairs), so the
// trees' workgroups of one feature group read the same planes of one image
// and an XCD's L2 holds them.
//
// Pass 1 is bound by the vector memory path (L1 -> L2 requests times their
// latency, about 70 in flight per CU), not by arithmetic or HBM.  Two
// workgroups of four waves per CU, each lane with its one sample's 8 corner
// loads in flight, beat every configuration with more samples in flight
// (more samples per lane, more waves, larger blocks): with 256-sample blocks
// the workgroups an XCD runs at once cover a few feature groups of a
// segment rather than many, the planes they read fit the XCD's L2 (1.3%
// misses instead of 3.8%) and requests return sooner.  The launch reserves
// LDS the kernel does not use to hold the CU at two workgroups.
#include <map>
#include <mutex>

#include "cf_internal.h"

namespace cf {

namespace {

struct BinSample {
    double rinv;   // depthReciprocal(depth in metres)
    int xy;        // x | y << 16
    int lab;       // flip << 8 | (invalid depth) << 10 | (flip entry) << 12;  < 0: empty slot
};

// #{t : th[t] < q} over a row padded with +inf (never below q), T for NaN.
// One compare and one add-with-carry per threshold: the compare's lane mask is
// the carry-in of the count.
template<int TM>
__device__ __forceinline__ int binOfT(float q, const float (&th)[TM], int T) {
    uint32_t b = 0;
#pragma unroll
    for (int t = 0; t < TM; t++) b += (th[t] < q) ? 1u : 0u;
    return (q != q) ? T : static_cast<int>(b);
}

typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));

// Shifts a sample's 16-byte bin row right by one byte and inserts b as the
// last byte: after 16 inserts the first inserted bin is byte 0.
__device__ __forceinline__ void pushBin(uint32_t (&w)[4], uint32_t b) {
    w[0] = (w[0] >> 8) | (w[1] << 24);
    w[1] = (w[1] >> 8) | (w[2] << 24);
    w[2] = (w[2] >> 8) | (w[3] << 24);
    w[3] = (w[3] >> 8) | (b << 24);
}

// The 8 corners of one (sample, feature) response: depth corners as {sum, count}, colour corners'
// float bits in .x.  Returns false (and issues no load) when the sample has no valid depth or a box
// leaves the image.
__device__ __forceinline__ bool gatherCorners(const Feat& fr, const BinSample& rs, const float* col,
        const int2* dep, int W, int H, size_t plane, int2 (&cv)[8]) {
    const int lab = rs.lab;
    int idx[8];
    const bool ok = lab >= 0 && ((lab >> 10) & 1) == 0 &&
            responseCorners(fr, rs.xy & 0xffff, rs.xy >> 16, W, H, rs.rinv, ((lab >> 8) & 3) == 1, idx);
    if (ok) {
        if (fr.type == T_DEPTH) {
#pragma unroll
            for (int c = 0; c < 8; c++) cv[c] = planeAt(dep, idx[c]);
        } else {
            const float* colA = col + static_cast<size_t>(fr.c1) * plane;
            const float* colB = col + static_cast<size_t>(fr.c2) * plane;
#pragma unroll
            for (int c = 0; c < 4; c++) cv[c].x = __float_as_int(planeAt(colA, idx[c]));
#pragma unroll
            for (int c = 4; c < 8; c++) cv[c].x = __float_as_int(planeAt(colB, idx[c]));
        }
    }
    return ok;
}

template<int TM>
__device__ __forceinline__ uint32_t binFromCorners(const Feat& fr, bool ok, const int2 (&cv)[8],
        const float (&th)[TM], int T) {
    if (!ok) return static_cast<uint32_t>(T);
    double r;
    if (fr.type == T_DEPTH) {
        r = depthResponseFromCorners(cv);
    } else {
        float c8[8];
#pragma unroll
        for (int c = 0; c < 8; c++) c8[c] = __int_as_float(cv[c].x);
        r = colorResponseFromCorners(c8);
    }
    return static_cast<uint32_t>(binOfT<TM>(roundUpToFloat(r), th, T));
}

}  // namespace

// Pass 1 with the launch shape <LANES lanes per workgroup, SPL samples per
// lane, FPS features whose corner gathers are issued together>: a workgroup
// covers one block of LANES * SPL samples.
template<int TM, int LANES, int SPL, int FPS>
__global__ void __launch_bounds__(LANES) k_eval_bins(const float* __restrict__ colBase,
        const int2* __restrict__ depBase, int W, int H, const cf_sample* __restrict__ samples,
        const uint32_t* __restrict__ perm, const BinBlock* __restrict__ blocks, const int2* __restrict__ work,
        int g0, const FeatRec* __restrict__ gFeat, const float* __restrict__ gThr, int F, int T, int N,
        uint4* __restrict__ bins, size_t flipOffset) {
    static_assert(kBinGroup % FPS == 0, "features per step divide the group");
    // work[blockIdx] = (sample block, feature group within the batch); the
    // XCDs' lists are interleaved (workgroup w runs on XCD w % 8) so each XCD
    // sweeps one image (the segments of all trees) for every group of the
    // batch before the next image.  Padding entries have a negative block.
    const int2 wk = work[blockIdx.x];
    if (wk.x < 0) return;
    const int g = wk.y;
    const BinBlock bb = blocks[wk.x];
    const size_t plane = static_cast<size_t>(W) * H;
    // all samples of a block belong to image bb.image: the plane bases are uniform
    const int2* dep = depBase + static_cast<size_t>(bb.image) * plane;
    const float* col = colBase + static_cast<size_t>(bb.image) * 3 * plane;

    const int pos0 = (g0 + g) * kBinGroup;
    uint4* out = bins + static_cast<size_t>(g) * N;
    uint4* outFlip = bins + flipOffset + static_cast<size_t>(g) * N;
    for (int base = 0; base < bb.count; base += LANES * SPL) {
        BinSample rs[SPL];
        int outIdx[SPL];
#pragma unroll
        for (int k = 0; k < SPL; k++) {
            const int i = base + k * LANES + threadIdx.x;
            // empty lanes keep in-range fields (they issue no gathers)
            rs[k].lab = -1;
            rs[k].xy = 0;
            rs[k].rinv = 0.0;
            outIdx[k] = -1;
            if (i < bb.count) {
                const uint32_t v = perm[bb.begin + i];
                const int s = static_cast<int>(v & ~kFlipEntry);
                const cf_sample x = samples[s];
                const bool valid = x.depth_mm > 0;
                // a flip entry evaluates a Both sample flipped; other entries in the sample's own
                // orientation (Both: unflipped); bit 12 selects the flipped bin rows
                const bool flipEntry = (v & kFlipEntry) != 0;
                outIdx[k] = s;
                rs[k].rinv = valid ? depthReciprocal(depthToFloat(x.depth_mm)) : 0.0;
                rs[k].xy = x.x | (x.y << 16);
                rs[k].lab = ((flipEntry ? 1 : x.flip) << 8) | ((valid ? 0 : 1) << 10) | (flipEntry ? 1 << 12 : 0);
            }
        }
        uint32_t row[SPL][4];
#pragma unroll
        for (int k = 0; k < SPL; k++) row[k][0] = row[k][1] = row[k][2] = row[k][3] = 0u;

        // thresholds are uniform (scalar loads): with few of them in flight they are loaded together with
        // the features, before the gathers
        constexpr bool kThFirst = FPS * TM <= 64;
        for (int fi = 0; fi < kBinGroup; fi += FPS) {
            Feat fr[FPS];
            bool live[FPS];
            int2 cv[FPS][SPL][8];
            bool ok[FPS][SPL];
            float th[kThFirst ? FPS : 1][TM];
            // the gathers of FPS features x SPL samples are all issued before any response is formed
#pragma unroll
            for (int j = 0; j < FPS; j++) {
                const size_t tp = static_cast<size_t>(bb.tree) * F + pos0 + fi + j;
                live[j] = pos0 + fi + j < F;   // uniform: the last group of the table may be partial
                fr[j] = live[j] ? loadFeat(gFeat + tp) : Feat{};
                if (kThFirst && live[j]) {
#pragma unroll
                    for (int t = 0; t < TM; t++) th[kThFirst ? j : 0][t] = gThr[tp * TM + t];
                }
#pragma unroll
                for (int k = 0; k < SPL; k++) {
                    ok[j][k] = live[j] && gatherCorners(fr[j], rs[k], col, dep, W, H, plane, cv[j][k]);
                }
            }
#pragma unroll
            for (int j = 0; j < FPS; j++) {
                if (!live[j]) {
#pragma unroll
                    for (int k = 0; k < SPL; k++) pushBin(row[k], 0u);
                    continue;
                }
                const int tj = kThFirst ? j : 0;
                if (!kThFirst) {
                    const size_t tp = static_cast<size_t>(bb.tree) * F + pos0 + fi + j;
#pragma unroll
                    for (int t = 0; t < TM; t++) th[tj][t] = gThr[tp * TM + t];
                }
#pragma unroll
                for (int k = 0; k < SPL; k++) {
                    pushBin(row[k], binFromCorners<TM>(fr[j], ok[j][k], cv[j][k], th[tj], T));
                }
            }
        }
#pragma unroll
        for (int k = 0; k < SPL; k++) {
            if (outIdx[k] >= 0) {
                // rows are read once, by pass 2: streaming stores keep them from evicting image planes from L2
                uint4* o = (rs[k].lab & (1 << 12)) ? outFlip : out;
                const u32x4 v = {row[k][0], row[k][1], row[k][2], row[k][3]};
                __builtin_nontemporal_store(v, reinterpret_cast<u32x4*>(o + outIdx[k]));
            }
        }
    }
}

// Feature records and thresholds in evaluation order, thresholds padded to
// TM per feature with +inf: gFeat[tree][pos], gThr[tree][pos][TM].  A NaN
// threshold becomes +inf and -0 becomes +0; neither changes any comparison
// th < q (NaN < q and +inf < q are false for every q, and -0 and +0 compare
// equal).
__global__ void __launch_bounds__(256) k_bin_tables(const FeatRec* __restrict__ feat, const float* __restrict__ thrFT,
        const uint16_t* __restrict__ evalOrder, int total, int F, int T, int TM, FeatRec* __restrict__ gFeat,
        float* __restrict__ gThr) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int tree = i / F;
    const size_t tf = static_cast<size_t>(tree) * F + evalOrder[i];
    gFeat[i] = feat[tf];
    for (int t = 0; t < TM; t++) {
        const float v = t < T ? thrFT[tf * T + t] : __builtin_inff();
        gThr[static_cast<size_t>(i) * TM + t] = (v != v) ? __builtin_inff() : v + 0.0f;
    }
}

// Pass 2: interval histograms of (node chunk, feature group) from the bin rows.
// G holds the batch's groups: G[slot][ng * 16][T + 1][C].  Lane l counts
// feature l % 16 of sample l / 16, so one LDS atomic instruction covers 4
// samples x 16 features: the 16 features' histograms lie in different banks
// and at most 4 lanes can hit one address (all lanes counting one feature, as
// for samples of an almost pure node, would collide on few bins).
__global__ void __launch_bounds__(256) k_hist_bins(const uint4* __restrict__ bins, const uint8_t* __restrict__ labels,
        const Chunk* __restrict__ chunks, int g0, int ng, int F, int T, int C, int N, uint32_t* __restrict__ G,
        size_t flipOffset) {
    extern __shared__ uint32_t h[];
    const int chunkIdx = blockIdx.x / ng;
    const int g = blockIdx.x % ng;
    const Chunk ch = chunks[chunkIdx];
    const int nf = min(kBinGroup, F - (g0 + g) * kBinGroup);
    const int binsPerF = (T + 1) * C;
    const int nb = nf * binsPerF;
    for (int k = threadIdx.x; k < nb; k += 256) h[k] = 0;
    __syncthreads();
    const int fi = threadIdx.x & (kBinGroup - 1);
    const int sub = threadIdx.x >> 4;   // 16 samples per pass of the workgroup
    const uint8_t* row = reinterpret_cast<const uint8_t*>(bins + static_cast<size_t>(g) * N + ch.begin);
    const uint8_t* rowFlip = reinterpret_cast<const uint8_t*>(bins + flipOffset + static_cast<size_t>(g) * N + ch.begin);
    const uint8_t* lab = labels + ch.begin;
    uint32_t* hf = h + fi * binsPerF;
    if (fi < nf) {
        constexpr int kUnroll = 16;   // bins loaded ahead of their atomics (8: 4% slower, 4: 19% slower)
        for (int base = 0; base < ch.count; base += 16 * kUnroll) {
            int b[kUnroll], l[kUnroll];
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                const int i = base + u * 16 + sub;
                l[u] = -1;
                if (i < ch.count) {
                    b[u] = row[static_cast<size_t>(i) * kBinGroup + fi];
                    l[u] = lab[i];
                }
            }
#pragma unroll
            for (int u = 0; u < kUnroll; u++) {
                // bins are 0..T; the clamp keeps the LDS index inside the histogram whatever a row holds
                if (l[u] >= 0) {
                    atomicAdd(&hf[min(b[u], T) * C + (l[u] & 0x7f)], 1u);
                    if (l[u] & 0x80) {   // still Both: its flipped evaluation counts too
                        const int i = base + u * 16 + sub;
                        atomicAdd(&hf[min(static_cast<int>(rowFlip[static_cast<size_t>(i) * kBinGroup + fi]), T) * C +
                                (l[u] & 0x7f)], 1u);
                    }
                }
            }
        }
    }
    __syncthreads();
    uint32_t* out = G + (static_cast<size_t>(ch.slot) * ng + g) * kBinGroup * binsPerF;
    for (int k = threadIdx.x; k < nb; k += 256) {
        const uint32_t v = h[k];
        if (v) atomicAdd(&out[k], v);
    }
}

// Per tree: feature evaluation order grouped by the planes a feature reads
// (depth first, then colour by (channel1, channel2)), stable in feature index.
// One wave per table: a stable counting sort over the 10 keys, 64 features at a time -- a lane's place
// among the lanes of its key is the count of its key's ballot bits below it.
__global__ void __launch_bounds__(64) k_feature_eval_order(const FeatRec* __restrict__ feat, int F,
        uint16_t* __restrict__ evalOrder) {
    constexpr int kKeys = 10;
    const FeatRec* ft = feat + static_cast<size_t>(blockIdx.x) * F;
    uint16_t* eo = evalOrder + static_cast<size_t>(blockIdx.x) * F;
    const int lane = threadIdx.x;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    auto keyOf = [&](int f) {
        if (f >= F) return -1;
        const Feat x = loadFeat(ft + f);
        return x.type == T_DEPTH ? 0 : 1 + x.c1 * 3 + x.c2;
    };
    int start[kKeys];
#pragma unroll
    for (int k = 0; k < kKeys; k++) start[k] = 0;
    for (int f0 = 0; f0 < F; f0 += 64) {
        const int key = keyOf(f0 + lane);
#pragma unroll
        for (int k = 0; k < kKeys; k++) start[k] += __popcll(__ballot(key == k));
    }
    int acc = 0;
#pragma unroll
    for (int k = 0; k < kKeys; k++) {
        const int c = start[k];
        start[k] = acc;
        acc += c;
    }
    for (int f0 = 0; f0 < F; f0 += 64) {
        const int key = keyOf(f0 + lane);
#pragma unroll
        for (int k = 0; k < kKeys; k++) {
            const uint64_t m = __ballot(key == k);
            if (key == k) eo[start[k] + __popcll(m & below)] = static_cast<uint16_t>(f0 + lane);
            start[k] += __popcll(m);
        }
    }
}

// Sort keys ((tree * images + image) << yBits | row) and values (level index) of one tree's level
// segment.
__global__ void __launch_bounds__(256) k_level_keys(const cf_sample* __restrict__ samples, int begin, int n,
        uint32_t keyBase, uint32_t keyMul, uint32_t* __restrict__ keys, uint32_t* __restrict__ vals, int flipN,
        uint32_t absentKey, int yBits) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const cf_sample x = samples[begin + i];
    // (tree, image) segment, then the sample's row: the samples of a row sit in one wave, where
    // corner reads of a feature that fall on one cache line are merged
    const uint32_t row = yBits > 0 ? x.y : 0u;
    const uint32_t seg = static_cast<uint32_t>(x.image) * keyMul + keyBase;
    keys[begin + i] = (seg << yBits) | row;
    vals[begin + i] = static_cast<uint32_t>(begin + i);
    if (flipN > 0) {
        keys[flipN + begin + i] = x.flip == 2 ? ((seg << yBits) | row) : (absentKey << yBits);
        vals[flipN + begin + i] = static_cast<uint32_t>(begin + i) | kFlipEntry;
    }
}

// One thread per (XCD list position, XCD): the segment holding the position by binary search over
// the XCD's block prefix counts, then group and block within the segment.
__global__ void __launch_bounds__(256) k_build_work(const int* __restrict__ segFirst, const int* __restrict__ segCount,
        const int* __restrict__ xcdPrefix, const int* __restrict__ xcdBlocks, int numSegs, int ngb, int maxLen,
        int2* __restrict__ work) {
    constexpr int kXcds = 8;
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<size_t>(maxLen) * kXcds) return;
    const int x = static_cast<int>(i % kXcds);
    const long long pos = static_cast<long long>(i / kXcds);
    int2 item = make_int2(-1, 0);
    if (x < numSegs && pos < static_cast<long long>(xcdBlocks[x]) * ngb) {
        // segments x, x + 8, ...: the last m with xcdPrefix[x + 8 m] * ngb <= pos
        int lo = 0, hi = (numSegs - x + kXcds - 1) / kXcds;   // m in [lo, hi)
        while (hi - lo > 1) {
            const int mid = (lo + hi) / 2;
            if (static_cast<long long>(xcdPrefix[x + kXcds * mid]) * ngb <= pos) lo = mid;
            else hi = mid;
        }
        const int seg = x + kXcds * lo;
        const long long local = pos - static_cast<long long>(xcdPrefix[seg]) * ngb;
        const int nb = segCount[seg];
        item = make_int2(segFirst[seg] + static_cast<int>(local % nb), static_cast<int>(local / nb));
    }
    work[i] = item;
}

__global__ void __launch_bounds__(256) k_level_labels(const cf_sample* __restrict__ samples, int n,
        uint8_t* __restrict__ labels) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) labels[i] = static_cast<uint8_t>(samples[i].label | (samples[i].flip == 2 ? 0x80 : 0));
}

size_t deviceLdsLimit() {
    static std::mutex mu;
    static std::map<int, size_t> limits;
    int dev = 0;
    CF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto it = limits.find(dev);
    if (it != limits.end()) return it->second;
    int v = 0;
    CF_HIP(hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
    return limits[dev] = static_cast<size_t>(v);
}

void allowLds(const void* kernel, size_t bytes) {
    if (bytes <= 64 * 1024) return;
    if (bytes > deviceLdsLimit()) {
        throw Error(CF_ERR_UNSUPPORTED, "(thresholds + 1) x classes too large for the LDS histograms of 16 features (" +
                std::to_string(bytes) + " bytes)");
    }
    CF_HIP(hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(bytes)));
}

int binThresholdWidth(int T) {
    static const int widths[] = {8, 12, 16, 20, 24, 32, 64, kMaxBinThresholds};
    for (int w : widths) {
        if (T <= w) return w;
    }
    throw Error(CF_ERR_ARG, "too many thresholds for 8-bit bins");
}

void launchBinTables(hipStream_t s, const FeatTables& tabs, int trees, const uint16_t* evalOrder, FeatRec* gFeat,
        float* gThr) {
    const int total = trees * tabs.F;
    if (total == 0) return;
    hipLaunchKernelGGL(k_bin_tables, dim3((total + 255) / 256), dim3(256), 0, s, tabs.feat.get(), tabs.thrFT.get(),
            evalOrder, total, tabs.F, tabs.T, binThresholdWidth(tabs.T), gFeat, gThr);
}

void launchFeatureEvalOrder(hipStream_t s, const FeatTables& tabs, int trees, uint16_t* evalOrder) {
    hipLaunchKernelGGL(k_feature_eval_order, dim3(trees), dim3(64), 0, s, tabs.feat.get(), tabs.F, evalOrder);
}

void launchLevelKeys(hipStream_t s, const cf_sample* samples, int begin, int n, uint32_t keyBase, uint32_t keyMul,
        uint32_t* keys, uint32_t* vals, int flipN, uint32_t absentKey, int yBits) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_level_keys, dim3((n + 255) / 256), dim3(256), 0, s, samples, begin, n, keyBase, keyMul, keys,
            vals, flipN, absentKey, yBits);
}

void launchBuildWork(hipStream_t s, const int* segFirst, const int* segCount, const int* xcdPrefix,
        const int* xcdBlocks, int numSegs, int ngb, int maxLen, int2* work) {
    const size_t total = static_cast<size_t>(maxLen) * 8;
    if (total == 0) return;
    hipLaunchKernelGGL(k_build_work, dim3(static_cast<unsigned>((total + 255) / 256)), dim3(256), 0, s, segFirst,
            segCount, xcdPrefix, xcdBlocks, numSegs, ngb, maxLen, work);
}

void launchLevelLabels(hipStream_t s, const cf_sample* samples, int n, uint8_t* labels) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_level_labels, dim3((n + 255) / 256), dim3(256), 0, s, samples, n, labels);
}

// Launch shapes of pass 1 the tuner chooses from.  Shape 0 is the default:
// two workgroups of four waves per CU, one sample per lane, one feature's
// gathers at a time (the fastest of a manual sweep on the full-frame
// workload).  The others trade samples in flight per CU (waves, samples per
// lane, workgroups per CU) against L2 working set, and issue two features'
// gathers together.
constexpr EvalShape kEvalShapes[] = {
    {256, 1, 1, 2},
    {256, 1, 2, 2},
    {256, 1, 1, 3},
    {128, 1, 1, 4},
    {128, 2, 1, 2},
    {256, 2, 1, 1},
    {512, 1, 1, 1},
    {128, 1, 2, 4},
    {128, 1, 4, 4},
    {256, 1, 4, 2},
};
constexpr int kNumEvalShapes = static_cast<int>(sizeof(kEvalShapes) / sizeof(kEvalShapes[0]));
static_assert(kNumEvalShapes <= 16, "TrainStats counts launches of at most 16 shapes");

int numEvalShapes() { return kNumEvalShapes; }

const EvalShape& evalShape(int i) {
    if (i < 0 || i >= kNumEvalShapes) throw Error(CF_ERR_ARG, "unknown pass-1 launch shape");
    return kEvalShapes[i];
}

// Dynamic LDS that k_eval_bins reserves (and does not use) so that exactly
// w workgroups fit a CU: more than 1 / (w + 1) and at most 1 / w of the CU's
// LDS; 64 KB where that lies inside (w = 2 on 160 KB CUs), else the middle of
// the interval.
static size_t evalBinsLdsBytes(int w) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, size_t> bytes;
    int dev = 0;
    CF_HIP(hipGetDevice(&dev));
    std::lock_guard<std::mutex> g(mu);
    auto it = bytes.find({dev, w});
    if (it != bytes.end()) return it->second;
    int perCu = 0;
    CF_HIP(hipDeviceGetAttribute(&perCu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev));
    const size_t lo = static_cast<size_t>(perCu) / (w + 1) + 1;
    const size_t hi = static_cast<size_t>(perCu) / w;
    size_t b = 64 * 1024;
    if (b < lo || b > hi) b = ((lo + hi) / 2) & ~static_cast<size_t>(255);
    return bytes[{dev, w}] = b;
}

template<int LANES, int SPL, int FPS>
static void launchEvalBinsShape(hipStream_t s, int TM, size_t lds, const ImageSet& imgs, const cf_sample* samples,
        const uint32_t* perm, const BinBlock* blocks, const int2* work, int numWork, int g0, const FeatRec* gFeat,
        const float* gThr, const FeatTables& tabs, int N, uint4* bins, size_t flipOffset) {
    const dim3 grid(static_cast<unsigned>(numWork)), block(LANES);
#define CF_EVAL_BINS(TMV)                                                                                        \
    allowLds(reinterpret_cast<const void*>(&k_eval_bins<TMV, LANES, SPL, FPS>), lds);                            \
    hipLaunchKernelGGL((k_eval_bins<TMV, LANES, SPL, FPS>), grid, block, lds, s, imgs.col, imgs.dep, imgs.W,     \
            imgs.H, samples, perm, blocks, work, g0, gFeat, gThr, tabs.F, tabs.T, N, bins, flipOffset)
    switch (TM) {
        case 8: CF_EVAL_BINS(8); break;
        case 12: CF_EVAL_BINS(12); break;
        case 16: CF_EVAL_BINS(16); break;
        case 20: CF_EVAL_BINS(20); break;
        case 24: CF_EVAL_BINS(24); break;
        case 32: CF_EVAL_BINS(32); break;
        case 64: CF_EVAL_BINS(64); break;
        default: CF_EVAL_BINS(kMaxBinThresholds); break;
    }
#undef CF_EVAL_BINS
}

void launchEvalBins(hipStream_t s, const ImageSet& imgs, const cf_sample* samples, const uint32_t* perm,
        const BinBlock* blocks, const int2* work, int numWork, int g0, const FeatRec* gFeat, const float* gThr,
        const FeatTables& tabs, int N, uint4* bins, size_t flipOffset, int shape) {
    if (numWork == 0) return;
    const int TM = binThresholdWidth(tabs.T);
    const EvalShape& sh = evalShape(shape);
    const size_t lds = evalBinsLdsBytes(sh.wgPerCu);
#define CF_SHAPE(L, S, F)                                                                                         \
    if (sh.lanes == L && sh.perLane == S && sh.featPerStep == F) {                                               \
        launchEvalBinsShape<L, S, F>(s, TM, lds, imgs, samples, perm, blocks, work, numWork, g0, gFeat, gThr, tabs, \
                N, bins, flipOffset);                                                                            \
        return;                                                                                                  \
    }
    CF_SHAPE(256, 1, 1)
    CF_SHAPE(256, 1, 2)
    CF_SHAPE(128, 1, 1)
    CF_SHAPE(128, 2, 1)
    CF_SHAPE(256, 2, 1)
    CF_SHAPE(512, 1, 1)
    CF_SHAPE(128, 1, 2)
    CF_SHAPE(128, 1, 4)
    CF_SHAPE(256, 1, 4)
#undef CF_SHAPE
    throw Error(CF_ERR_RUNTIME, "pass-1 launch shape without a kernel");
}

void launchHistBins(hipStream_t s, const uint4* bins, const uint8_t* labels, const Chunk* chunks, int numChunks,
        int g0, int ng, const FeatTables& tabs, int C, int N, uint32_t* G, size_t flipOffset) {
    if (numChunks == 0 || ng == 0) return;
    const size_t lds = static_cast<size_t>(kBinGroup) * (tabs.T + 1) * C * sizeof(uint32_t);
    allowLds(reinterpret_cast<const void*>(&k_hist_bins), lds);
    hipLaunchKernelGGL(k_hist_bins, dim3(static_cast<unsigned>(numChunks) * ng), dim3(256), lds, s, bins, labels,
            chunks, g0, ng, tabs.F, tabs.T, C, N, G, flipOffset);
}

}  // namespace cf
