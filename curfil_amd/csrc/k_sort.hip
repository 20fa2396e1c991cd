// Image-major order of a batched level: a counting sort of the level
// entries by their key ((tree, image) segment << yBits | row).
//
// The keys are small integers (segments x rows: 4000 x 480 buckets for 5
// trees over 800 frames of 640x480), so one counting pass replaces a radix
// sort:
//   k_sort_count    counts the entries of every bucket (global atomics);
//   k_scan_*        exclusive prefix sum of the counts (per-block scans, a
//                   scan of the block totals, the block offsets added);
//   k_sort_scatter  places every entry at its bucket's cursor (atomic, so
//                   the order inside a bucket follows arrival);
//   k_sort_settle   restores the input order inside every segment bucket of
//                   at most kSettleMax entries by sorting its values (the
//                   entries' level indices, with kFlipEntry on the flipped
//                   evaluations that follow all plain entries in the input),
//                   one lane per bucket;
//   k_sort_starts   the first position of every (tree, image) segment.
// Buckets hold a few entries on average (one sample per 77 pixels per tree
// at the benchmark's sampling density; a bucket holds at most the entries of
// one row of one segment), so the result is the stable sort by key.  Longer
// buckets, and the bucket past the segments that collects the flipped
// entries of samples that are not Both (up to one per sample, never
// evaluated), keep their arrival order: the order inside a bucket only
// decides which lanes of pass 1 evaluate which samples of one row, never a
// result, since every bin row is stored at its entry's level index.
#include "cf_internal.h"

namespace cf {

namespace {

constexpr int kScanBlock = 1024;   // counts scanned per workgroup (256 lanes x 4)
constexpr uint32_t kSettleMax = 64;   // longest bucket put back into input order

__global__ void __launch_bounds__(256) k_fill_u32(uint32_t* __restrict__ p, size_t n, uint32_t v) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void __launch_bounds__(256) k_sort_count(const uint32_t* __restrict__ keys, int n,
        uint32_t* __restrict__ count) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) atomicAdd(&count[keys[i]], 1u);
}

// Exclusive scan of kScanBlock counts per workgroup into `start`, the workgroup's total into blockSum.
__global__ void __launch_bounds__(256) k_scan_blocks(const uint32_t* __restrict__ count, size_t n,
        uint32_t* __restrict__ start, uint32_t* __restrict__ blockSum) {
    __shared__ uint32_t waveSum[4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const size_t base = static_cast<size_t>(blockIdx.x) * kScanBlock + static_cast<size_t>(tid) * 4;
    uint32_t v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) v[k] = base + k < n ? count[base + k] : 0u;
    const uint32_t own = v[0] + v[1] + v[2] + v[3];
    // inclusive scan of `own` over the wave
    uint32_t inc = own;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) waveSum[wave] = inc;
    __syncthreads();
    uint32_t before = inc - own;
    for (int w = 0; w < wave; w++) before += waveSum[w];
#pragma unroll
    for (int k = 0; k < 4; k++) {
        if (base + k < n) start[base + k] = before;
        before += v[k];
    }
    if (tid == 255) blockSum[blockIdx.x] = before;
}

// Exclusive scan of the block totals in place, by one workgroup.
__global__ void __launch_bounds__(256) k_scan_totals(uint32_t* __restrict__ blockSum, int nb) {
    __shared__ uint32_t waveSum[4];
    __shared__ uint32_t carry;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0) carry = 0;
    __syncthreads();
    for (int b0 = 0; b0 < nb; b0 += 256) {
        const int i = b0 + tid;
        const uint32_t own = i < nb ? blockSum[i] : 0u;
        uint32_t inc = own;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        if (lane == 63) waveSum[wave] = inc;
        __syncthreads();
        uint32_t before = carry + inc - own;
        for (int w = 0; w < wave; w++) before += waveSum[w];
        if (i < nb) blockSum[i] = before;
        __syncthreads();
        if (tid == 255) carry = before + own;
        __syncthreads();
    }
}

__global__ void __launch_bounds__(256) k_scan_add(uint32_t* __restrict__ start, size_t n,
        const uint32_t* __restrict__ blockSum) {
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) start[i] += blockSum[i / kScanBlock];
}

// cursor[b] starts at bucket b's first position and ends one past its last.
__global__ void __launch_bounds__(256) k_sort_scatter(const uint32_t* __restrict__ keys,
        const uint32_t* __restrict__ vals, int n, uint32_t* __restrict__ cursor, uint32_t* __restrict__ out) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[atomicAdd(&cursor[keys[i]], 1u)] = vals[i];
}

// One lane per segment bucket of 2..kSettleMax entries: insertion sort of its values.
__global__ void __launch_bounds__(256) k_sort_settle(const uint32_t* __restrict__ count,
        const uint32_t* __restrict__ cursor, size_t buckets, uint32_t* __restrict__ out) {
    const size_t b = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (b >= buckets) return;
    const uint32_t c = count[b];
    if (c < 2 || c > kSettleMax) return;
    const uint32_t e = cursor[b], s = e - c;
    for (uint32_t i = s + 1; i < e; i++) {
        const uint32_t v = out[i];
        uint32_t j = i;
        while (j > s && out[j - 1] > v) {
            out[j] = out[j - 1];
            j--;
        }
        out[j] = v;
    }
}

// starts[seg] = first position of segment seg, or -1 if it is empty.  After the scatter, a bucket's
// first position is cursor - count.
__global__ void __launch_bounds__(256) k_sort_starts(const uint32_t* __restrict__ count,
        const uint32_t* __restrict__ cursor, int numSegs, int yBits, int* __restrict__ starts) {
    const int seg = blockIdx.x * blockDim.x + threadIdx.x;
    if (seg > numSegs) return;
    const size_t b0 = static_cast<size_t>(seg) << yBits;
    const uint32_t first = cursor[b0] - count[b0];
    if (seg == numSegs) {
        starts[seg] = static_cast<int>(first);
        return;
    }
    const size_t b1 = static_cast<size_t>(seg + 1) << yBits;
    const uint32_t next = cursor[b1] - count[b1];
    starts[seg] = next > first ? static_cast<int>(first) : -1;
}

inline unsigned blocksFor(size_t n, int per) { return static_cast<unsigned>((n + per - 1) / per); }

}  // namespace

void launchFillU32(hipStream_t s, uint32_t* p, size_t n, uint32_t v) {
    if (n == 0) return;
    hipLaunchKernelGGL(k_fill_u32, dim3(blocksFor(n, 256)), dim3(256), 0, s, p, n, v);
}

size_t levelSortBuckets(int numSegs, int yBits) {
    // the segments, one past the last for the flipped evaluations of samples that are not Both, and
    // the empty bucket after them that k_sort_starts reads
    return (static_cast<size_t>(numSegs) + 2) << yBits;
}

void levelCountingSort(hipStream_t s, const uint32_t* keys, const uint32_t* vals, int n, int numSegs, int yBits,
        uint32_t* count, uint32_t* cursor, uint32_t* blockSum, uint32_t* out, int* starts) {
    const size_t buckets = levelSortBuckets(numSegs, yBits);
    if (buckets > (static_cast<size_t>(1) << 28)) throw Error(CF_ERR_UNSUPPORTED, "too many level segments to sort");
    const int nb = static_cast<int>(blocksFor(buckets, kScanBlock));
    hipLaunchKernelGGL(k_fill_u32, dim3(blocksFor(buckets, 256)), dim3(256), 0, s, count, buckets, 0u);
    if (n > 0) hipLaunchKernelGGL(k_sort_count, dim3(blocksFor(n, 256)), dim3(256), 0, s, keys, n, count);
    hipLaunchKernelGGL(k_scan_blocks, dim3(nb), dim3(256), 0, s, count, buckets, cursor, blockSum);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(256), 0, s, blockSum, nb);
    hipLaunchKernelGGL(k_scan_add, dim3(blocksFor(buckets, 256)), dim3(256), 0, s, cursor, buckets, blockSum);
    if (n > 0) {
        hipLaunchKernelGGL(k_sort_scatter, dim3(blocksFor(n, 256)), dim3(256), 0, s, keys, vals, n, cursor, out);
        // the segments' buckets only: the bucket past them holds the flipped entries of samples that are not Both
        const size_t segBuckets = static_cast<size_t>(numSegs) << yBits;
        hipLaunchKernelGGL(k_sort_settle, dim3(blocksFor(segBuckets, 256)), dim3(256), 0, s, count, cursor, segBuckets,
                out);
    }
    hipLaunchKernelGGL(k_sort_starts, dim3(blocksFor(static_cast<size_t>(numSegs) + 1, 256)), dim3(256), 0, s,
            count, cursor, numSegs, yBits, starts);
}

size_t levelSortScanBlocks(int numSegs, int yBits) {
    return (levelSortBuckets(numSegs, yBits) + kScanBlock - 1) / kScanBlock;
}

}  // namespace cf
