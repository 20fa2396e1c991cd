// Integral images on the device (RGBDImage::calculateIntegral,
// image.cpp:236-277 and :531-556).
//
// 8-bit colour and integer depth: the reference integrates in double with a
// Kahan term; for integer input every partial sum is an exact integer below
// 2^53, so its result is the exact sum rounded once to float (colour) or the
// exact integer (depth, valid).  Here the sums are exact 32-bit integer prefix
// sums (wrapping for depth, see boxDepth), built in two passes:
//   pass A: column prefix sums, 16 row segments per column (coalesced along x);
//   pass B: one wave per row scans 64-pixel segments with DPP/shuffle scans.
// Float colour (non-integer values): the reference's sequential recurrence is
// replayed exactly, one thread per channel (used for small test images).
#include "cf_internal.h"

namespace cf {

// pass A: column prefix of 5 channels (r, g, b, depth mm, valid) into tmp[5][H][W].
// A workgroup takes 16 columns; its 16 x 16 lanes split the rows into 16
// segments: each lane sums its segment, the segment totals meet in LDS, and
// each lane writes its segment's prefix sums starting from the totals of the
// segments above (integer sums, so the order of additions does not matter).
// Workgroups of 256 lanes: the serving loop integrates frames while the
// classification of others fills the CUs, and a workgroup of 1024 lanes (16
// waves) got a CU only when a classification launch drained, so the next
// launch, which waits for the frame, started late.
constexpr int kColChunk = 16, kRowSegs = 16;

__global__ void __launch_bounds__(kColChunk * kRowSegs) k_integral_cols_u8(const uint8_t* __restrict__ rgb,
        const uint16_t* __restrict__ depth, int W, int H, uint32_t* __restrict__ tmp) {
    __shared__ uint32_t total[kRowSegs][5][kColChunk];
    const int cx = threadIdx.x % kColChunk, seg = threadIdx.x / kColChunk;
    const int x = blockIdx.x * kColChunk + cx;
    const int rowsPer = (H + kRowSegs - 1) / kRowSegs;
    const int y0 = min(H, seg * rowsPer), y1 = min(H, y0 + rowsPer);
    const bool col = x < W;
    uint32_t sum[5] = {0u, 0u, 0u, 0u, 0u};
    if (col) {
#pragma unroll 4
        for (int y = y0; y < y1; y++) {
            const size_t p = static_cast<size_t>(y) * W + x;
            const uint32_t mm = depth[p];
            sum[0] += rgb[3 * p + 0];
            sum[1] += rgb[3 * p + 1];
            sum[2] += rgb[3 * p + 2];
            sum[3] += mm;
            sum[4] += (mm > 0) ? 1u : 0u;
        }
    }
#pragma unroll
    for (int c = 0; c < 5; c++) total[seg][c][cx] = sum[c];
    __syncthreads();
    uint32_t run[5] = {0u, 0u, 0u, 0u, 0u};
    for (int k = 0; k < seg; k++) {
#pragma unroll
        for (int c = 0; c < 5; c++) run[c] += total[k][c][cx];
    }
    if (!col) return;
    const size_t plane = static_cast<size_t>(W) * H;
#pragma unroll 4
    for (int y = y0; y < y1; y++) {
        const size_t p = static_cast<size_t>(y) * W + x;
        const uint32_t mm = depth[p];
        run[0] += rgb[3 * p + 0];
        run[1] += rgb[3 * p + 1];
        run[2] += rgb[3 * p + 2];
        run[3] += mm;
        run[4] += (mm > 0) ? 1u : 0u;
        tmp[p] = run[0];
        tmp[plane + p] = run[1];
        tmp[2 * plane + p] = run[2];
        tmp[3 * plane + p] = run[3];
        tmp[4 * plane + p] = run[4];
    }
}

// same for int32 depth (float-colour path: colour handled by the Kahan kernel)
__global__ void __launch_bounds__(256) k_integral_cols_depth_i32(const int32_t* __restrict__ depth, int W, int H,
        uint32_t* __restrict__ tmp) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= W) return;
    uint32_t d = 0, v = 0;
    const size_t plane = static_cast<size_t>(W) * H;
    for (int y = 0; y < H; y++) {
        const size_t p = static_cast<size_t>(y) * W + x;
        const int32_t mm = depth[p];
        d += static_cast<uint32_t>(mm > 0 ? mm : 0);
        v += (mm > 0) ? 1u : 0u;
        tmp[3 * plane + p] = d;
        tmp[4 * plane + p] = v;
    }
}

__device__ __forceinline__ uint32_t waveInclusiveScan(uint32_t v) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const uint32_t n = __shfl_up(v, o, 64);
        if (lane >= o) v += n;
    }
    return v;
}

// pass B: row prefix.  blockDim = 256 (4 waves), one wave per (row, channel)
// job; channels 0..2 -> float colour, channel 3+4 -> int2 depth.
__global__ void __launch_bounds__(256) k_integral_rows(const uint32_t* __restrict__ tmp, int W, int H,
        int firstChannel, float* __restrict__ col, int2* __restrict__ dep) {
    const int wave = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    const int numCh = 4 - firstChannel;  // jobs: colour channels [firstChannel,3) + one depth job
    if (wave >= H * numCh) return;
    const int y = wave / numCh;
    const int ch = firstChannel + wave % numCh;
    const size_t plane = static_cast<size_t>(W) * H;
    const size_t row = static_cast<size_t>(y) * W;
    if (ch < 3) {
        uint32_t carry = 0;
        for (int x0 = 0; x0 < W; x0 += 64) {
            const int x = x0 + lane;
            const uint32_t v = (x < W) ? tmp[ch * plane + row + x] : 0u;
            const uint32_t s = waveInclusiveScan(v) + carry;
            if (x < W) col[ch * plane + row + x] = __uint2float_rn(s);
            carry = __shfl(s, 63, 64);
        }
    } else {
        uint32_t cd = 0, cv = 0;
        for (int x0 = 0; x0 < W; x0 += 64) {
            const int x = x0 + lane;
            const uint32_t d = (x < W) ? tmp[3 * plane + row + x] : 0u;
            const uint32_t v = (x < W) ? tmp[4 * plane + row + x] : 0u;
            const uint32_t sd = waveInclusiveScan(d) + cd;
            const uint32_t sv = waveInclusiveScan(v) + cv;
            if (x < W) dep[row + x] = make_int2(static_cast<int>(sd), static_cast<int>(sv));
            cd = __shfl(sd, 63, 64);
            cv = __shfl(sv, 63, 64);
        }
    }
}

// Sequential double Kahan integration of one float channel per thread,
// operation by operation as image.cpp:250-277 with T = double.
__global__ void k_integral_kahan_f32(const float* __restrict__ color, int W, int H, double* __restrict__ work,
        float* __restrict__ col) {
    const int ch = threadIdx.x;
    if (ch >= 3) return;
    const size_t plane = static_cast<size_t>(W) * H;
    const float* in = color + ch * plane;
    double* d = work + ch * plane;
    float* out = col + ch * plane;
    double c = 0.0;
    for (int y = 0; y < H; ++y) {
        for (int x = 0; x < W; ++x) {
            const double above = (y > 0) ? d[(y - 1) * W + x] : 0.0;
            const double left = (x > 0) ? d[y * W + x - 1] : 0.0;
            const double aboveLeft = (x > 0 && y > 0) ? d[(y - 1) * W + x - 1] : 0.0;
            const double cur = static_cast<double>(in[y * W + x]);
            const double dy = __dsub_rn(__dadd_rn(left, above), aboveLeft) - c;
            const double t = __dadd_rn(cur, dy);
            c = __dsub_rn(__dsub_rn(t, cur), dy);
            d[y * W + x] = t;
            out[y * W + x] = __double2float_rn(t);
        }
    }
}

// The same recurrence with the previous row of partial sums in LDS and the
// left / above-left terms in registers (the values are the ones the
// recurrence stored, so results are identical); one wave per channel.
__global__ void __launch_bounds__(64) k_integral_kahan_f32_lds(const float* __restrict__ color, int W, int H,
        float* __restrict__ col) {
    extern __shared__ double rows[];   // [3][W]: row y-1 ahead of x, row y behind it
    const int ch = threadIdx.x;
    if (ch >= 3) return;
    const size_t plane = static_cast<size_t>(W) * H;
    const float* in = color + ch * plane;
    float* out = col + ch * plane;
    double* row = rows + static_cast<size_t>(ch) * W;
    double c = 0.0;
    for (int y = 0; y < H; ++y) {
        double left = 0.0, aboveLeft = 0.0;
        double above = y > 0 ? row[0] : 0.0;
        for (int x = 0; x < W; ++x) {
            const double nextAbove = (y > 0 && x + 1 < W) ? row[x + 1] : 0.0;   // not yet overwritten
            const double cur = static_cast<double>(in[static_cast<size_t>(y) * W + x]);
            const double dy = __dsub_rn(__dadd_rn(left, above), aboveLeft) - c;
            const double t = __dadd_rn(cur, dy);
            c = __dsub_rn(__dsub_rn(t, cur), dy);
            row[x] = t;
            out[static_cast<size_t>(y) * W + x] = __double2float_rn(t);
            left = t;
            aboveLeft = above;
            above = nextAbove;
        }
    }
}

void launchIntegralU8(hipStream_t s, const uint8_t* rgb, const uint16_t* depth, int W, int H, uint32_t* tmp,
        float* col, int2* dep) {
    hipLaunchKernelGGL(k_integral_cols_u8, dim3((W + kColChunk - 1) / kColChunk), dim3(kColChunk * kRowSegs), 0, s,
            rgb, depth, W, H, tmp);
    const int jobs = H * 4;
    hipLaunchKernelGGL(k_integral_rows, dim3((jobs * 64 + 255) / 256), dim3(256), 0, s, tmp, W, H, 0, col, dep);
}

void launchIntegralF32(hipStream_t s, const float* color, const int32_t* depth, int W, int H, uint32_t* tmp,
        double* work, float* col, int2* dep) {
    const size_t lds = static_cast<size_t>(3) * W * sizeof(double);
    if (lds <= 64 * 1024) {
        hipLaunchKernelGGL(k_integral_kahan_f32_lds, dim3(1), dim3(64), lds, s, color, W, H, col);
    } else {
        hipLaunchKernelGGL(k_integral_kahan_f32, dim3(1), dim3(64), 0, s, color, W, H, work, col);
    }
    hipLaunchKernelGGL(k_integral_cols_depth_i32, dim3((W + 255) / 256), dim3(256), 0, s, depth, W, H, tmp);
    const int jobs = H;  // depth job only
    hipLaunchKernelGGL(k_integral_rows, dim3((jobs * 64 + 255) / 256), dim3(256), 0, s, tmp, W, H, 3, col, dep);
}

// An image smaller than its set's slot, integrated at its own size and then
// padded the way loadImages pads integrated images (RGBDImage::resizeImage,
// image.cpp:449-475): colour NaN and depth sums / valid counts 0 outside it.
__global__ void __launch_bounds__(256) k_integral_pad(const float* __restrict__ srcCol,
        const int2* __restrict__ srcDep, int w, int h, int W, int H, float* __restrict__ col, int2* __restrict__ dep) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= W || y >= H) return;
    const bool inside = x < w && y < h;
    const size_t plane = static_cast<size_t>(W) * H, src = static_cast<size_t>(y) * w + x;
    const size_t dst = static_cast<size_t>(y) * W + x;
    const float nan = __int_as_float(0x7fc00000);
#pragma unroll
    for (int c = 0; c < 3; c++) col[c * plane + dst] = inside ? srcCol[c * static_cast<size_t>(w) * h + src] : nan;
    dep[dst] = inside ? srcDep[src] : make_int2(0, 0);
}

void launchIntegralPad(hipStream_t s, const float* srcCol, const int2* srcDep, int w, int h, int W, int H, float* col,
        int2* dep) {
    hipLaunchKernelGGL(k_integral_pad, dim3((W + 255) / 256, H), dim3(256), 0, s, srcCol, srcDep, w, h, W, H, col, dep);
}

}  // namespace cf
