// Stable partition of every split node's samples into its children
// (RandomTreeTrain::train, random_tree.h:1307-1352: a sample goes LEFT iff its
// response is <= the split threshold, children keep the parent's order).
//
//   k_part_dir      per chunk: direction of each sample, the chunk's left
//                   count and the node's left class histogram
//   k_part_scatter  per chunk: order-preserving compaction of left and right
//                   samples to the destinations the host assigned (children
//                   that stop growing get no destination)
#include "cf_internal.h"

namespace cf {

__global__ void __launch_bounds__(256) k_part_dir(const DevImage* __restrict__ imgs,
        const cf_sample* __restrict__ samples, const Chunk* __restrict__ chunks, const SplitRec* __restrict__ splits,
        int C, uint8_t* __restrict__ dir, int* __restrict__ chunkLeft, uint32_t* __restrict__ leftHist) {
    __shared__ uint32_t hist[kMaxClasses];
    __shared__ int total;
    const Chunk ch = chunks[blockIdx.x];
    const SplitRec sp = splits[ch.slot];
    const Feat f = loadFeat(&splits[ch.slot].feat);
    if (threadIdx.x < kMaxClasses) hist[threadIdx.x] = 0;
    if (threadIdx.x == 0) total = 0;
    __syncthreads();
    int cnt = 0;
    for (int i = threadIdx.x; i < ch.count; i += blockDim.x) {
        const cf_sample s = samples[ch.begin + i];
        const double r = featureResponse(imgs[s.image], f, s.x, s.y, s.depth_mm, s.flip == 1);
        const bool left = r <= static_cast<double>(sp.threshold);
        dir[ch.begin + i] = left ? 1 : 0;
        if (left) {
            cnt++;
            atomicAdd(&hist[s.label], 1u);
        }
    }
    atomicAdd(&total, cnt);
    __syncthreads();
    if (threadIdx.x == 0) chunkLeft[blockIdx.x] = total;
    if (threadIdx.x < C && hist[threadIdx.x]) atomicAdd(&leftHist[static_cast<size_t>(ch.slot) * C + threadIdx.x], hist[threadIdx.x]);
}

__global__ void __launch_bounds__(256) k_part_scatter(const cf_sample* __restrict__ samples,
        const Chunk* __restrict__ chunks, const uint8_t* __restrict__ dir, const int2* __restrict__ dst,
        cf_sample* __restrict__ out) {
    __shared__ int waveL[4];
    const Chunk ch = chunks[blockIdx.x];
    const int2 d = dst[blockIdx.x];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int doneL = 0, doneR = 0;
    for (int base = 0; base < ch.count; base += 256) {
        const int i = base + threadIdx.x;
        const bool valid = i < ch.count;
        const bool isLeft = valid && dir[ch.begin + i];
        const unsigned long long m = __ballot(isLeft);
        const int pre = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) waveL[wave] = __popcll(m);
        __syncthreads();
        int wavePre = 0, totL = 0;
        for (int w = 0; w < 4; w++) {
            if (w < wave) wavePre += waveL[w];
            totL += waveL[w];
        }
        const int nValid = min(256, ch.count - base);
        if (valid) {
            const int prefL = wavePre + pre;
            const int prefR = threadIdx.x - prefL;
            const cf_sample s = samples[ch.begin + i];
            if (isLeft) {
                if (d.x >= 0) out[d.x + doneL + prefL] = s;
            } else {
                if (d.y >= 0) out[d.y + doneR + prefR] = s;
            }
        }
        doneL += totL;
        doneR += nValid - totL;
        __syncthreads();
    }
}

void launchPartitionDirections(hipStream_t s, const DevImage* imgs, const cf_sample* samples, const Chunk* chunks,
        int numChunks, const SplitRec* splits, int C, uint8_t* dir, int* chunkLeft, uint32_t* leftHist) {
    if (numChunks == 0) return;
    hipLaunchKernelGGL(k_part_dir, dim3(numChunks), dim3(256), 0, s, imgs, samples, chunks, splits, C, dir, chunkLeft,
            leftHist);
}

void launchPartitionScatter(hipStream_t s, const cf_sample* samples, const Chunk* chunks, int numChunks,
        const uint8_t* dir, const int2* dst, cf_sample* out) {
    if (numChunks == 0) return;
    hipLaunchKernelGGL(k_part_scatter, dim3(numChunks), dim3(256), 0, s, samples, chunks, dir, dst, out);
}

}  // namespace cf
