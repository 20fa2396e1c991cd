// Allocation-free batched dense prediction for callers that cannot allocate
// device memory (code under HIP graph capture, tight serving loops).
//
// cf_predict_batch_device stages the batch's frame indices in a device
// buffer that the forest grows on demand.  This variant stages them in
// `scratch`, a file-scope device array of a fixed size, so a call makes no
// device or host allocation: it needs the forest already resident on the
// device (any earlier prediction with the forest uploads it) and refuses
// batches of more than kNoAllocMaxFrames frames.  Every call on one device
// shares that device's copy of `scratch`; it holds the last call's indices
// and is fully rewritten by the next call, so nothing needs clearing.  Calls
// that overlap on one device (from different contexts or host threads
// without a common lock) race on it.
#include <mutex>
#include <string>

#include "cf_host.h"

namespace {

constexpr int kNoAllocMaxFrames = 65536;   // 256 KiB of indices per device, held until the process ends
constexpr int kMaxDevices = 64;

static __device__ __align__(256) int scratch[kNoAllocMaxFrames];

// The device address of `scratch` on each device, resolved once with that device current.  The array's
// name in host code is a host-side stand-in, never a device address: only these pointers go to kernels
// and copies.
std::mutex gScratchMu;
int* gScratch[kMaxDevices] = {};
std::string gScratchError[kMaxDevices];   // why resolving failed, per device

}  // namespace

namespace cf {

void resolvePredictScratch(int device) {
    if (device < 0 || device >= kMaxDevices) throw Error(CF_ERR_UNSUPPORTED, "device index above 63");
    std::lock_guard<std::mutex> g(gScratchMu);
    if (gScratch[device]) return;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
        gScratchError[device] = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
        throw Error(CF_ERR_HIP, gScratchError[device]);
    }
    gScratch[device] = static_cast<int*>(p);
}

}  // namespace cf

using namespace cf;

extern "C" int cf_predict_batch_device_noalloc(cf_ctx* ctx, cf_forest* fo, cf_images* im, const int* indices,
        int count, float* d_probs, uint8_t* d_labels) {
    if (!ctx) return CF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    try {
        if (!fo || !im || !indices) throw Error(CF_ERR_ARG, "null argument");
        if (count <= 0) throw Error(CF_ERR_ARG, "no frames");
        if (count > kNoAllocMaxFrames) {
            throw Error(CF_ERR_ARG, "batch of more than 65536 frames (the allocation-free prediction's limit)");
        }
        for (int i = 0; i < count; i++) {
            if (indices[i] < 0 || indices[i] >= im->count) throw Error(CF_ERR_ARG, "frame index out of range");
        }
        if (fo->ctx && fo->ctx != ctx) throw Error(CF_ERR_ARG, "forest belongs to another context");
        if (!fo->uploaded || fo->ctx != ctx) {
            throw Error(CF_ERR_ARG, "forest not resident on the device (predict with it once first)");
        }
        int* frames = nullptr;
        if (ctx->device >= 0 && ctx->device < kMaxDevices) {
            std::lock_guard<std::mutex> g(gScratchMu);
            frames = gScratch[ctx->device];
            if (!frames) {
                throw Error(CF_ERR_HIP, "the allocation-free prediction's device array is not available on this "
                        "device: " + gScratchError[ctx->device]);
            }
        } else {
            throw Error(CF_ERR_UNSUPPORTED, "device index above 63");
        }
        CF_HIP(hipMemcpyAsync(frames, indices, static_cast<size_t>(count) * sizeof(int), hipMemcpyHostToDevice,
                ctx->stream));
        const DevForest df = fo->devForest();
        launchClassify(ctx->stream, im->desc.get(), frames, count, im->W, im->H, df, d_probs, d_labels);
        CF_HIP(hipStreamSynchronize(ctx->stream));
        ctx->error.clear();
        return CF_OK;
    } catch (const Error& e) {
        ctx->error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->error = e.what();
        return CF_ERR_RUNTIME;
    }
}
