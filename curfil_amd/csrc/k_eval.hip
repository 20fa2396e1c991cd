// Split evaluation of a frontier of nodes:
//
//   k_eval_hist   one workgroup per (node chunk, feature tile): the double
//                 response of every (sample, feature), mapped to its interval
//                 among the feature's thresholds, counted per class in LDS
//                 ("interval histogram", T+1 bins x C classes per feature),
//                 then added to the node's global bins.
//                 Replaces featureResponseKernel + aggregateHistogramsKernel
//                 (random_tree_image_gpu.cu:1455-1720) and the CPU counter
//                 loop (random_tree_image.cpp:97-153): the response is never
//                 stored and never re-read per threshold.
//   k_scores      one thread per (node, feature): prefix sums over the bins
//                 give every threshold's left/right class counts; FP64
//                 NormalizedInformationGainScore (score.h:45-216) in the
//                 reference's operation order; per-feature best / second best.
//   k_reduce      per node: the reference's threshold-major sequential argmax
//                 with its 1e-13 tie rule (random_tree_image.cpp:445-457,
//                 random_tree.cpp:36-52) as max + first occurrence, valid
//                 whenever no score lies within 1e-13 below the maximum; such
//                 nodes are flagged and replayed sequentially (k_seq_argmax).
//
// Left at threshold t  <=>  r <= t  <=>  q := roundUpToFloat(r) <= t
//                      <=>  bin(q) := #{k : thr[k] < q} <= rank[t]
// with NaN responses in bin T (always right), so left(t, c) is the prefix sum
// of the bins up to rank[t].
#include <cfloat>

#include "cf_internal.h"

namespace cf {

constexpr int kEvalThreads = 256;
constexpr int kFeatTile = 16;
constexpr int kSamplesPerThread = 4;   // samples held in registers per pass
constexpr int kPass = kEvalThreads * kSamplesPerThread;
constexpr int kXcds = 8;                  // MI355X: 8 XCDs, workgroups dispatched round-robin

// One training pixel held in registers for the feature loop.
struct RegSample {
    double rinv;   // depthReciprocal(depth in metres)
    int img;
    int xy;        // x | y << 16
    int lab;       // label | flip << 8 | (invalid depth) << 10;  < 0: empty slot
};

// Interval of q among the T thresholds (thresholds held in registers; the
// loop bound is a compile-time constant so the compare chain unrolls).
template<int TM>
__device__ __forceinline__ int binOf(float q, const float (&th)[TM], int T) {
    if (__builtin_isnan(q)) return T;
    int b = 0;
#pragma unroll
    for (int t = 0; t < TM; t++) b += (t < T && th[t] < q) ? 1 : 0;
    return b;
}

// All images of a set share W x H and are stored back to back, so a sample's
// integrals are found from its image index.
//
// Work mapping: workgroups are dispatched round-robin over the 8 XCDs, each
// with its own L2.  Workgroup L runs on XCD L % 8; the j-th workgroup of an
// XCD takes feature tile j % tiles of chunk (j / tiles) * 8 + L % 8, so an XCD
// sweeps all feature tiles of one chunk (one image region) before the next
// and keeps that image's integrals cache-resident.
template<int TM>
__global__ void __launch_bounds__(kEvalThreads) k_eval_hist(const float* __restrict__ colBase,
        const int2* __restrict__ depBase, int W, int H, const cf_sample* __restrict__ samples,
        const Chunk* __restrict__ chunks, int numChunks, int tiles, const FeatRec* __restrict__ feat,
        const float* __restrict__ thrFT, int F, int T, int C, uint32_t* __restrict__ G) {
    extern __shared__ uint32_t h[];
    const int xcd = blockIdx.x % kXcds;
    const int j = blockIdx.x / kXcds;
    const int chunkIdx = (j / tiles) * kXcds + xcd;
    if (chunkIdx >= numChunks) return;
    const Chunk ch = chunks[chunkIdx];
    const int f0 = (j % tiles) * kFeatTile;
    const int nf = min(kFeatTile, F - f0);
    const int binsPerF = (T + 1) * C;
    const int nb = nf * binsPerF;
    const size_t plane = static_cast<size_t>(W) * H;
    for (int k = threadIdx.x; k < nb; k += kEvalThreads) h[k] = 0;
    __syncthreads();
    const FeatRec* fp = feat + static_cast<size_t>(ch.tree) * F + f0;
    const float* tp = thrFT + (static_cast<size_t>(ch.tree) * F + f0) * T;
    for (int base = 0; base < ch.count; base += kPass) {
        RegSample rs[kSamplesPerThread];
#pragma unroll
        for (int k = 0; k < kSamplesPerThread; k++) {
            const int i = base + k * kEvalThreads + threadIdx.x;
            // lanes past the chunk's end gather (unused) corners at index 0
            // of image 0, so every field must hold an in-range value
            rs[k].lab = -1;
            rs[k].img = 0;
            rs[k].xy = 0;
            rs[k].rinv = 0.0;
            if (i < ch.count) {
                const cf_sample s = samples[ch.begin + i];
                const bool valid = s.depth_mm > 0;
                rs[k].rinv = valid ? depthReciprocal(depthToFloat(s.depth_mm)) : 0.0;
                rs[k].img = static_cast<int>(s.image);
                rs[k].xy = s.x | (s.y << 16);
                rs[k].lab = s.label | (s.flip << 8) | ((valid ? 0 : 1) << 10);
            }
        }
        for (int fi = 0; fi < nf; fi++) {
            const Feat fr = loadFeat(fp + fi);
            float th[TM];
            const float* tf = tp + static_cast<size_t>(fi) * T;
#pragma unroll
            for (int t = 0; t < TM; t++) th[t] = (t < T) ? tf[t] : 0.0f;
            const int hoff = fi * binsPerF;
            // flipped responses of flip=Both samples take the generic path below
            if (fr.type == T_DEPTH) {
                int2 cv[kSamplesPerThread][8];
                bool ok[kSamplesPerThread];
#pragma unroll
                for (int k = 0; k < kSamplesPerThread; k++) {
                    const int lab = rs[k].lab;
                    int idx[8];
                    ok[k] = lab >= 0 && (lab >> 10) == 0 &&
                            responseCorners(fr, rs[k].xy & 0xffff, rs[k].xy >> 16, W, H, rs[k].rinv,
                                    ((lab >> 8) & 3) == 1, idx);
                    const int2* dep = depBase + static_cast<size_t>(rs[k].img) * plane;
#pragma unroll
                    for (int c = 0; c < 8; c++) cv[k][c] = dep[ok[k] ? idx[c] : 0];
                }
#pragma unroll
                for (int k = 0; k < kSamplesPerThread; k++) {
                    const int lab = rs[k].lab;
                    if (lab < 0) continue;
                    int b = T;
                    if (ok[k]) b = binOf<TM>(roundUpToFloat(depthResponseFromCorners(cv[k])), th, T);
                    atomicAdd(&h[hoff + b * C + (lab & 0xff)], 1u);
                }
            } else {
                float cv[kSamplesPerThread][8];
                bool ok[kSamplesPerThread];
#pragma unroll
                for (int k = 0; k < kSamplesPerThread; k++) {
                    const int lab = rs[k].lab;
                    int idx[8];
                    ok[k] = lab >= 0 && (lab >> 10) == 0 &&
                            responseCorners(fr, rs[k].xy & 0xffff, rs[k].xy >> 16, W, H, rs[k].rinv,
                                    ((lab >> 8) & 3) == 1, idx);
                    const float* colA = colBase + static_cast<size_t>(rs[k].img) * 3 * plane + fr.c1 * plane;
                    const float* colB = colBase + static_cast<size_t>(rs[k].img) * 3 * plane + fr.c2 * plane;
#pragma unroll
                    for (int c = 0; c < 4; c++) cv[k][c] = colA[ok[k] ? idx[c] : 0];
#pragma unroll
                    for (int c = 4; c < 8; c++) cv[k][c] = colB[ok[k] ? idx[c] : 0];
                }
#pragma unroll
                for (int k = 0; k < kSamplesPerThread; k++) {
                    const int lab = rs[k].lab;
                    if (lab < 0) continue;
                    int b = T;
                    if (ok[k]) b = binOf<TM>(roundUpToFloat(colorResponseFromCorners(cv[k])), th, T);
                    atomicAdd(&h[hoff + b * C + (lab & 0xff)], 1u);
                }
            }
            // the second (flipped) response of samples with flip setting Both
#pragma unroll
            for (int k = 0; k < kSamplesPerThread; k++) {
                const int lab = rs[k].lab;
                if (lab < 0 || ((lab >> 8) & 3) != 2) continue;
                int b = T;
                if ((lab >> 10) == 0) {
                    const float* col = colBase + static_cast<size_t>(rs[k].img) * 3 * plane;
                    const int2* dep = depBase + static_cast<size_t>(rs[k].img) * plane;
                    const double r = featureResponseR(col, dep, W, H, fr, rs[k].xy & 0xffff, rs[k].xy >> 16,
                            rs[k].rinv, true);
                    b = binOf<TM>(roundUpToFloat(r), th, T);
                }
                atomicAdd(&h[hoff + b * C + (lab & 0xff)], 1u);
            }
        }
    }
    __syncthreads();
    uint32_t* g = G + (static_cast<size_t>(ch.slot) * F + f0) * binsPerF;
    for (int k = threadIdx.x; k < nb; k += kEvalThreads) {
        const uint32_t v = h[k];
        if (v) atomicAdd(&g[k], v);
    }
}

// NormalizedInformationGainScore::calculateScore (score.h:191-215) over
// InformationGainScore::calculateScore (score.h:87-139); the operation order
// follows the reference so host and device round identically apart from log2.
__device__ __forceinline__ double entropyOf(double p) {
    if (p == 0.0) return 0.0;
    return -p * log2(p);
}
__device__ __forceinline__ double clamp01(double s) {
    if (s > 1.0) return 1.0;
    if (s < 0.0) return 0.0;
    return s;
}

template<int CM>
__device__ double nigScore(int C, const uint32_t* L, const uint32_t* R, const uint32_t* A) {
    double totalLeft = 0, totalRight = 0;
#pragma unroll
    for (int c = 0; c < CM; c++) {
        if (c < C) {
            totalLeft += static_cast<double>(L[c]);
            totalRight += static_cast<double>(R[c]);
        }
    }
    const double total = totalLeft + totalRight;
    const double leftProb = totalLeft / total;
    const double rightProb = totalRight / total;
    double score = 0;
#pragma unroll
    for (int c = 0; c < CM; c++) {
        if (c < C) {
            const uint32_t l = L[c], r = R[c];
            const double classProb = static_cast<double>(A[c]) / total;
            if (l > 0) {
                const double p = static_cast<double>(l) / total;
                score = __dadd_rn(score, __dmul_rn(p, log2(p / (leftProb * classProb))));
            }
            if (r > 0) {
                const double p = static_cast<double>(r) / total;
                score = __dadd_rn(score, __dmul_rn(p, log2(p / (rightProb * classProb))));
            }
        }
    }
    const double ig = clamp01(score);
    if (ig == 0) return 0;
    const double Hs = (entropyOf(totalLeft) + entropyOf(totalRight) - entropyOf(total)) / total;
    double Hc = 0;
#pragma unroll
    for (int c = 0; c < CM; c++) {
        if (c < C && A[c] > 0) Hc += entropyOf(static_cast<double>(A[c]));
    }
    Hc -= entropyOf(total);
    Hc /= total;
    return clamp01((2 * ig) / (Hs + Hc));
}

template<int CM>
__global__ void __launch_bounds__(256) k_scores(const uint32_t* __restrict__ G, const uint32_t* __restrict__ nodeHist,
        const int* __restrict__ slotTree, int numSlots, const uint16_t* __restrict__ rank,
        const uint16_t* __restrict__ order, int F, int T, int C, FeatSummary* __restrict__ summ,
        double* __restrict__ allScores, int onlySlot, int fOffset, int Fb, const uint16_t* __restrict__ evalOrder) {
    // G holds Fb features per slot starting at evaluation position fOffset;
    // position p is feature evalOrder[tree][p] (p itself without an order).
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    int slot, j;
    if (onlySlot >= 0) {
        slot = onlySlot;
        j = gid;
        if (j >= Fb) return;
    } else {
        if (gid >= numSlots * Fb) return;
        slot = gid / Fb;
        j = gid % Fb;
    }
    const int pos = fOffset + j;
    if (pos >= F) return;
    const int tree = slotTree[slot];
    const int f = evalOrder ? static_cast<int>(evalOrder[static_cast<size_t>(tree) * F + pos]) : pos;
    uint32_t A[CM], run[CM];
#pragma unroll
    for (int c = 0; c < CM; c++) {
        A[c] = (c < C) ? nodeHist[static_cast<size_t>(slot) * C + c] : 0u;
        run[c] = 0;
    }
    const size_t tf = static_cast<size_t>(tree) * F + f;
    const uint32_t* g = G + (static_cast<size_t>(slot) * Fb + j) * (T + 1) * C;
    const uint16_t* rk = rank + tf * T;
    const uint16_t* od = order + tf * T;
    double best = -DBL_MAX, second = -DBL_MAX;
    int bestT = 0x7fffffff;
    int b = 0;
    for (int k = 0; k < T; k++) {
        const int t = od[k];
        const int r = rk[t];
        while (b <= r) {
#pragma unroll
            for (int c = 0; c < CM; c++) {
                if (c < C) run[c] += g[b * C + c];
            }
            b++;
        }
        uint32_t right[CM];
#pragma unroll
        for (int c = 0; c < CM; c++) right[c] = A[c] - run[c];
        const double sc = nigScore<CM>(C, run, right, A);
        if (allScores) allScores[(static_cast<size_t>(onlySlot >= 0 ? 0 : slot) * T + t) * F + f] = sc;
        if (sc > best) {
            second = best;
            best = sc;
            bestT = t;
        } else if (sc == best) {
            bestT = min(bestT, t);
        } else if (sc > second) {
            second = sc;
        }
    }
    FeatSummary s;
    s.best = best;
    s.second = second;
    s.thresh = bestT;
    s.pad = 0;
    summ[static_cast<size_t>(slot) * F + f] = s;
}

// scores[T][F] of one node straight from counters [F][T][C][2]
// (ImageFeatureEvaluation::calculateScores, random_tree_image.cpp:188-237).
template<int CM>
__global__ void __launch_bounds__(256) k_scores_counters(const uint32_t* __restrict__ counters,
        const uint32_t* __restrict__ hist, int F, int T, int C, double* __restrict__ scores) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F * T) return;
    const int f = i / T, t = i % T;
    uint32_t L[CM], R[CM], A[CM];
#pragma unroll
    for (int c = 0; c < CM; c++) {
        const size_t k = ((static_cast<size_t>(f) * T + t) * C + c) * 2;
        L[c] = c < C ? counters[k] : 0u;
        R[c] = c < C ? counters[k + 1] : 0u;
        A[c] = c < C ? hist[c] : 0u;
    }
    scores[static_cast<size_t>(t) * F + f] = nigScore<CM>(C, L, R, A);
}

void launchScoresFromCounters(hipStream_t s, const uint32_t* counters, const uint32_t* hist, int F, int T, int C,
        double* scores) {
    const dim3 grid((F * T + 255) / 256);
    if (C <= 8) {
        hipLaunchKernelGGL(k_scores_counters<8>, grid, dim3(256), 0, s, counters, hist, F, T, C, scores);
    } else {
        hipLaunchKernelGGL(k_scores_counters<kMaxClasses>, grid, dim3(256), 0, s, counters, hist, F, T, C, scores);
    }
}

__device__ __forceinline__ double waveMax(double v) {
    for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ long long waveMinLL(long long v) {
    for (int o = 32; o > 0; o >>= 1) {
        const long long u = __shfl_xor(v, o, 64);
        v = u < v ? u : v;
    }
    return v;
}

// One wave per node.
__global__ void __launch_bounds__(64) k_reduce(const FeatSummary* __restrict__ summ, int numSlots, int F,
        NodeBest* __restrict__ out) {
    const int slot = blockIdx.x;
    const int lane = threadIdx.x;
    const FeatSummary* s = summ + static_cast<size_t>(slot) * F;
    double m = -DBL_MAX;
    for (int f = lane; f < F; f += 64) m = fmax(m, s[f].best);
    m = waveMax(m);
    long long key = 0x7fffffffffffffffll;
    int band = 0;
    for (int f = lane; f < F; f += 64) {
        const FeatSummary x = s[f];
        if (x.best == m) {
            const long long k = static_cast<long long>(x.thresh) * F + f;
            key = k < key ? k : key;
            if (x.second > -DBL_MAX && (m - x.second) < 1e-13) band = 1;
        } else if ((m - x.best) < 1e-13) {
            band = 1;
        }
    }
    key = waveMinLL(key);
    band = __any(band) ? 1 : 0;
    if (lane == 0) {
        NodeBest nb;
        nb.score = m;
        nb.thresh = static_cast<int>(key / F);
        nb.feat = static_cast<int>(key % F);
        nb.flag = band;
        nb.pad = 0;
        out[slot] = nb;
    }
}

// The literal scan of random_tree_image.cpp:445-457 for one node.
// detail::isScoreBetter (random_tree.cpp:36-52) as used by the scan.
__device__ __forceinline__ bool scoreBetter(double best, double score) {
    if (__builtin_isnan(best)) return true;
    return !(fabs(score - best) < 1e-13) && (score > best);
}

// The scan of random_tree_image.cpp:445-457 for one node, one wave: scores
// are read 64 at a time in scan order; a ballot marks the lanes whose score
// beats the current best, and only those are resolved in order with the
// reference's predicate.  A score that does not beat the best cannot beat any
// later (larger) best: fl(score - best) does not grow as best grows.  The
// result is the sequential scan's.
__global__ void __launch_bounds__(64) k_seq_argmax(const double* __restrict__ allScores, int T, int F,
        NodeBest* __restrict__ allOut) {
    constexpr int kWin = 8;   // windows of 64 scores read per round
    const int lane = threadIdx.x;
    const long long n = static_cast<long long>(T) * F;
    // one node per workgroup: scores[T][F] and its result
    const double* scores = allScores + static_cast<size_t>(blockIdx.x) * n;
    NodeBest* out = allOut + blockIdx.x;
    double best = -__builtin_inf();
    long long pos = 0;
    for (long long base = 0; base < n; base += 64 * kWin) {
        double sc[kWin];
#pragma unroll
        for (int u = 0; u < kWin; u++) {
            const long long i = base + u * 64 + lane;
            sc[u] = i < n ? scores[i] : -__builtin_inf();
        }
#pragma unroll
        for (int u = 0; u < kWin; u++) {
            unsigned long long m = __ballot(scoreBetter(best, sc[u]));
            while (m) {
                const int l = __ffsll(static_cast<long long>(m)) - 1;
                const double c = __shfl(sc[u], l, 64);
                if (scoreBetter(best, c)) {
                    best = c;
                    pos = base + u * 64 + l;
                }
                m &= m - 1;
            }
        }
    }
    if (lane == 0) {
        NodeBest nb;
        nb.score = best;
        nb.feat = static_cast<int>(pos % F);
        nb.thresh = static_cast<int>(pos / F);
        nb.flag = 0;
        nb.pad = 0;
        *out = nb;
    }
}

// counters[F][T][C][2] of one node (slot 0) from its bins.
__global__ void __launch_bounds__(256) k_counters_from_bins(const uint32_t* __restrict__ G,
        const uint16_t* __restrict__ rank, int F, int T, int C, uint32_t* __restrict__ counters) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= F * T) return;
    const int f = i / T, t = i % T;
    const uint32_t* g = G + static_cast<size_t>(f) * (T + 1) * C;
    const int r = rank[static_cast<size_t>(f) * T + t];
    for (int c = 0; c < C; c++) {
        uint32_t left = 0, total = 0;
        for (int b = 0; b <= T; b++) {
            const uint32_t v = g[b * C + c];
            total += v;
            if (b <= r) left += v;
        }
        counters[((static_cast<size_t>(f) * T + t) * C + c) * 2 + 0] = left;
        counters[((static_cast<size_t>(f) * T + t) * C + c) * 2 + 1] = total - left;
    }
}

__global__ void __launch_bounds__(256) k_responses(const DevImage* __restrict__ imgs,
        const cf_sample* __restrict__ samples, int n, const FeatRec* __restrict__ feats, int F,
        double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n * F) return;
    const int f = i / n, k = i % n;
    const cf_sample s = samples[k];
    out[i] = featureResponse(imgs[s.image], loadFeat(feats + f), s.x, s.y, s.depth_mm, s.flip == 1);
}

// response of sample i to feature i
__global__ void __launch_bounds__(256) k_responses_paired(const DevImage* __restrict__ imgs,
        const cf_sample* __restrict__ samples, int n, const FeatRec* __restrict__ feats, double* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const cf_sample s = samples[i];
    out[i] = featureResponse(imgs[s.image], loadFeat(feats + i), s.x, s.y, s.depth_mm, s.flip == 1);
}

void launchEvalHistograms(hipStream_t s, const ImageSet& imgs, const cf_sample* samples, const Chunk* chunks,
        int numChunks, const FeatTables& tabs, int C, uint32_t* G) {
    if (numChunks == 0) return;
    const int tiles = (tabs.F + kFeatTile - 1) / kFeatTile;
    const size_t lds = static_cast<size_t>(kFeatTile) * (tabs.T + 1) * C * sizeof(uint32_t);
    const int groups = (numChunks + kXcds - 1) / kXcds;
    const dim3 grid(groups * kXcds * tiles), block(kEvalThreads);
    const void* kernel = tabs.T <= 8 ? reinterpret_cast<const void*>(&k_eval_hist<8>)
            : tabs.T <= 16 ? reinterpret_cast<const void*>(&k_eval_hist<16>)
            : tabs.T <= 24 ? reinterpret_cast<const void*>(&k_eval_hist<24>)
            : tabs.T <= 32 ? reinterpret_cast<const void*>(&k_eval_hist<32>)
            : tabs.T <= 64 ? reinterpret_cast<const void*>(&k_eval_hist<64>)
            : reinterpret_cast<const void*>(&k_eval_hist<kMaxThresholds>);
    allowLds(kernel, lds);
    if (tabs.T <= 8) {
        hipLaunchKernelGGL(k_eval_hist<8>, grid, block, lds, s, imgs.col, imgs.dep, imgs.W, imgs.H, samples,
                chunks, numChunks, tiles, tabs.feat.get(), tabs.thrFT.get(), tabs.F, tabs.T, C, G);
    } else if (tabs.T <= 16) {
        hipLaunchKernelGGL(k_eval_hist<16>, grid, block, lds, s, imgs.col, imgs.dep, imgs.W, imgs.H, samples,
                chunks, numChunks, tiles, tabs.feat.get(), tabs.thrFT.get(), tabs.F, tabs.T, C, G);
    } else if (tabs.T <= 24) {
        hipLaunchKernelGGL(k_eval_hist<24>, grid, block, lds, s, imgs.col, imgs.dep, imgs.W, imgs.H, samples,
                chunks, numChunks, tiles, tabs.feat.get(), tabs.thrFT.get(), tabs.F, tabs.T, C, G);
    } else if (tabs.T <= 32) {
        hipLaunchKernelGGL(k_eval_hist<32>, grid, block, lds, s, imgs.col, imgs.dep, imgs.W, imgs.H, samples,
                chunks, numChunks, tiles, tabs.feat.get(), tabs.thrFT.get(), tabs.F, tabs.T, C, G);
    } else if (tabs.T <= 64) {
        hipLaunchKernelGGL(k_eval_hist<64>, grid, block, lds, s, imgs.col, imgs.dep, imgs.W, imgs.H, samples,
                chunks, numChunks, tiles, tabs.feat.get(), tabs.thrFT.get(), tabs.F, tabs.T, C, G);
    } else {
        hipLaunchKernelGGL(k_eval_hist<kMaxThresholds>, grid, block, lds, s, imgs.col, imgs.dep, imgs.W, imgs.H, samples,
                chunks, numChunks, tiles, tabs.feat.get(), tabs.thrFT.get(), tabs.F, tabs.T, C, G);
    }
}

void launchScores(hipStream_t s, const uint32_t* G, const uint32_t* nodeHist, const int* slotTree, int numSlots,
        const FeatTables& tabs, int C, FeatSummary* summ, double* allScores, int onlySlot, int fOffset, int Fb,
        const uint16_t* evalOrder) {
    if (Fb < 0) Fb = tabs.F;
    const int threads = onlySlot >= 0 ? Fb : numSlots * Fb;
    if (threads == 0) return;
    const dim3 grid((threads + 255) / 256);
    if (C <= 8) {
        hipLaunchKernelGGL(k_scores<8>, grid, dim3(256), 0, s, G, nodeHist, slotTree, numSlots, tabs.rank.get(),
                tabs.order.get(), tabs.F, tabs.T, C, summ, allScores, onlySlot, fOffset, Fb, evalOrder);
    } else {
        hipLaunchKernelGGL(k_scores<kMaxClasses>, grid, dim3(256), 0, s, G, nodeHist, slotTree, numSlots,
                tabs.rank.get(), tabs.order.get(), tabs.F, tabs.T, C, summ, allScores, onlySlot, fOffset, Fb,
                evalOrder);
    }
}

void launchReduce(hipStream_t s, const FeatSummary* summ, int numSlots, int F, NodeBest* best) {
    if (numSlots == 0) return;
    hipLaunchKernelGGL(k_reduce, dim3(numSlots), dim3(64), 0, s, summ, numSlots, F, best);
}

void launchSequentialArgmax(hipStream_t s, const double* scores, int T, int F, NodeBest* best, int numNodes) {
    if (numNodes <= 0) return;
    hipLaunchKernelGGL(k_seq_argmax, dim3(static_cast<unsigned>(numNodes)), dim3(64), 0, s, scores, T, F, best);
}

void launchCountersFromBins(hipStream_t s, const uint32_t* G, const FeatTables& tabs, int C, uint32_t* counters) {
    const int n = tabs.F * tabs.T;
    hipLaunchKernelGGL(k_counters_from_bins, dim3((n + 255) / 256), dim3(256), 0, s, G, tabs.rank.get(), tabs.F,
            tabs.T, C, counters);
}

void launchResponses(hipStream_t s, const DevImage* imgs, const cf_sample* samples, int n, const FeatRec* feats,
        int F, double* out) {
    const int total = n * F;
    if (total == 0) return;
    hipLaunchKernelGGL(k_responses, dim3((total + 255) / 256), dim3(256), 0, s, imgs, samples, n, feats, F, out);
}

void launchResponsesPaired(hipStream_t s, const DevImage* imgs, const cf_sample* samples, int n, const FeatRec* feats,
        double* out) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_responses_paired, dim3((n + 255) / 256), dim3(256), 0, s, imgs, samples, n, feats, out);
}

}  // namespace cf
