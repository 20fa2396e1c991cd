// Stable partition of every split node's samples into its children
// (RandomTreeTrain::train, random_tree.h:1295-1360: a sample goes LEFT iff its
// response is <= the split threshold, children keep the parent's order).
//
// Horizontal flipping (SplitFunction::split, random_tree.h:122-144): a sample
// with flip setting Both whose flipped response falls on the other side is
// split in two: the sample itself, now NoFlip, goes to its side, and a Flip
// copy goes to the other side, at the same position in that child's order.
// So every sample gives at most one entry to each child; an entry weighs 2 in
// the child's histogram iff it is still Both.
//
//   k_part_dir      per chunk: the entries of each sample (bit 0: the sample
//                   goes left, bit 1: it is split), the chunk's left and right
//                   entry counts and the node's left class histogram
//   k_part_scatter  per chunk: order-preserving compaction of left and right
//                   entries to the destinations the host assigned (children
//                   that stop growing get no destination)
#include "cf_internal.h"

namespace cf {

// Split record of every node of the level from its argmax (feature of the sorted table, threshold index).
__global__ void __launch_bounds__(256) k_split_records(const NodeBest* __restrict__ best,
        const int* __restrict__ nodeTree, int numNodes, const FeatRec* __restrict__ feat,
        const float* __restrict__ thrFT, int F, int T, SplitRec* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= numNodes) return;
    const NodeBest b = best[i];
    const size_t tf = static_cast<size_t>(nodeTree[i]) * F + b.feat;
    SplitRec sr;
    sr.feat = feat[tf];
    sr.threshold = thrFT[tf * T + b.thresh];
    sr.flags = 0;
    sr.pad0 = sr.pad1 = 0;
    out[i] = sr;
}

__global__ void __launch_bounds__(256) k_part_dir(const DevImage* __restrict__ imgs,
        const cf_sample* __restrict__ samples, const Chunk* __restrict__ chunks, const SplitRec* __restrict__ splits,
        int C, uint8_t* __restrict__ dir, int2* __restrict__ chunkCounts, uint32_t* __restrict__ leftHist) {
    __shared__ uint32_t hist[kMaxClasses];
    __shared__ int totalL, totalR;
    const Chunk ch = chunks[blockIdx.x];
    const SplitRec sp = splits[ch.slot];
    const Feat f = loadFeat(&splits[ch.slot].feat);
    if (threadIdx.x < kMaxClasses) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) {
        totalL = 0;
        totalR = 0;
    }
    __syncthreads();
    int cntL = 0, cntR = 0;
    const double thr = static_cast<double>(sp.threshold);
    for (int i = threadIdx.x; i < ch.count; i += blockDim.x) {
        const cf_sample s = samples[ch.begin + i];
        const DevImage im = imgs[s.image];
        const bool left = featureResponse(im, f, s.x, s.y, s.depth_mm, s.flip == 1) <= thr;
        bool split = false;
        if (s.flip == 2) split = (featureResponse(im, f, s.x, s.y, s.depth_mm, true) <= thr) != left;
        dir[ch.begin + i] = static_cast<uint8_t>((left ? 1 : 0) | (split ? 2 : 0));
        // left entry: the sample (weight 2 if still Both) or its Flip copy (weight 1)
        const uint32_t wL = split ? 1u : (left ? (s.flip == 2 ? 2u : 1u) : 0u);
        if (wL) atomicAdd(&hist[s.label], wL);
        cntL += (left || split) ? 1 : 0;
        cntR += (!left || split) ? 1 : 0;
    }
    atomicAdd(&totalL, cntL);
    atomicAdd(&totalR, cntR);
    __syncthreads();
    if (threadIdx.x == 0) chunkCounts[blockIdx.x] = make_int2(totalL, totalR);
    if (threadIdx.x < C && hist[threadIdx.x]) atomicAdd(&leftHist[static_cast<size_t>(ch.slot) * C + threadIdx.x], hist[threadIdx.x]);
}

__global__ void __launch_bounds__(256) k_part_scatter(const cf_sample* __restrict__ samples,
        const Chunk* __restrict__ chunks, const uint8_t* __restrict__ dir, const int2* __restrict__ dst,
        cf_sample* __restrict__ out) {
    __shared__ int waveL[4], waveR[4];
    const Chunk ch = chunks[blockIdx.x];
    const int2 d = dst[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const unsigned long long below = (1ull << lane) - 1ull;
    int doneL = 0, doneR = 0;
    for (int base = 0; base < ch.count; base += 256) {
        const int i = base + threadIdx.x;
        const bool valid = i < ch.count;
        const uint8_t dr = valid ? dir[ch.begin + i] : 0;
        const bool left = dr & 1, split = dr & 2;
        const bool hasL = valid && (left || split);
        const bool hasR = valid && (!left || split);
        const unsigned long long mL = __ballot(hasL), mR = __ballot(hasR);
        if (lane == 0) {
            waveL[wave] = __popcll(mL);
            waveR[wave] = __popcll(mR);
        }
        __syncthreads();
        int preL = __popcll(mL & below), preR = __popcll(mR & below), totL = 0, totR = 0;
        for (int w = 0; w < 4; w++) {
            if (w < wave) {
                preL += waveL[w];
                preR += waveR[w];
            }
            totL += waveL[w];
            totR += waveR[w];
        }
        if (valid) {
            const cf_sample s = samples[ch.begin + i];
            if (hasL && d.x >= 0) {
                cf_sample e = s;
                if (split) e.flip = left ? 0 : 1;   // the sample itself (NoFlip) or its Flip copy
                out[d.x + doneL + preL] = e;
            }
            if (hasR && d.y >= 0) {
                cf_sample e = s;
                if (split) e.flip = left ? 1 : 0;
                out[d.y + doneR + preR] = e;
            }
        }
        doneL += totL;
        doneR += totR;
        __syncthreads();
    }
}

void launchSplitRecords(hipStream_t s, const NodeBest* best, const int* nodeTree, int numNodes, const FeatTables& tabs,
        SplitRec* out) {
    if (numNodes == 0) return;
    hipLaunchKernelGGL(k_split_records, dim3((numNodes + 255) / 256), dim3(256), 0, s, best, nodeTree, numNodes,
            tabs.feat.get(), tabs.thrFT.get(), tabs.F, tabs.T, out);
}

void launchPartitionDirections(hipStream_t s, const DevImage* imgs, const cf_sample* samples, const Chunk* chunks,
        int numChunks, const SplitRec* splits, int C, uint8_t* dir, int2* chunkCounts, uint32_t* leftHist) {
    if (numChunks == 0) return;
    hipLaunchKernelGGL(k_part_dir, dim3(numChunks), dim3(256), 0, s, imgs, samples, chunks, splits, C, dir,
            chunkCounts, leftHist);
}

void launchPartitionScatter(hipStream_t s, const cf_sample* samples, const Chunk* chunks, int numChunks,
        const uint8_t* dir, const int2* dst, cf_sample* out) {
    if (numChunks == 0) return;
    hipLaunchKernelGGL(k_part_scatter, dim3(numChunks), dim3(256), 0, s, samples, chunks, dir, dst, out);
}

}  // namespace cf
