// C ABI of the library (include/curfil_hip.h): handles, argument checks,
// host<->device copies, and the forest wrappers around the training driver.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>

#include "cf_host.h"
#include "forest_pack.h"

using namespace cf;

namespace {

template<class Fn>
int guarded(cf_ctx* ctx, Fn&& fn) {
    if (!ctx) return CF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    try {
        fn();
        ctx->error.clear();
        return CF_OK;
    } catch (const Error& e) {
        ctx->error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        ctx->error = "host allocation failed";
        return CF_ERR_NOMEM;
    } catch (const std::exception& e) {
        ctx->error = e.what();
        return CF_ERR_RUNTIME;
    }
}

void require(bool ok, const char* msg) {
    if (!ok) throw Error(CF_ERR_ARG, msg);
}


// Calls that need no device accept a null context; their errors go to hostError.
template<class Fn>
int guardedHost(cf_ctx* ctx, Fn&& fn) {
    if (ctx) return guarded(ctx, std::forward<Fn>(fn));
    std::string& hostError = hostErrorSlot();
    try {
        fn();
        hostError.clear();
        return CF_OK;
    } catch (const Error& e) {
        hostError = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        hostError = "host allocation failed";
        return CF_ERR_NOMEM;
    } catch (const std::exception& e) {
        hostError = e.what();
        return CF_ERR_RUNTIME;
    }
}

void checkSamples(const cf_images* im, const cf_sample* s, int n) {
    for (int i = 0; i < n; i++) {
        if (s[i].image >= static_cast<uint32_t>(im->count)) throw Error(CF_ERR_ARG, "sample image index out of range");
        if (s[i].x >= im->W || s[i].y >= im->H) throw Error(CF_ERR_ARG, "sample position outside the image");
        if (s[i].flip > 2) throw Error(CF_ERR_ARG, "invalid flip setting");
    }
}

// Feature parameters come from the caller: the type must be depth or colour
// and colour channels must name one of the 3 colour planes, or a response
// would read outside the integral images.
void checkFeature(int type, int c1, int c2) {
    if (type != T_DEPTH && type != T_COLOR) throw Error(CF_ERR_ARG, "invalid feature type");
    if (type == T_COLOR && (c1 < 0 || c1 > 2 || c2 < 0 || c2 > 2)) throw Error(CF_ERR_ARG, "invalid colour channel");
}

// Depth in millimetres above 50 m is refused, as the reference's depth loader
// refuses it (image.cpp:430); the reciprocal coordinate normalisation is proven
// equal to the reference's division for depths up to 50000 mm.
void checkDepthMm(const uint16_t* depth, size_t n) {
    for (size_t p = 0; p < n; p++) {
        if (depth[p] > 50000) throw Error(CF_ERR_ARG, "illegal depth value (> 50000 mm)");
    }
}

void checkFeaturesSoA(const int8_t* soa, int F) {
    for (int f = 0; f < F; f++) {
        checkFeature(soa[static_cast<size_t>(P_TYPE) * F + f], static_cast<uint8_t>(soa[static_cast<size_t>(P_C1) * F + f]),
                static_cast<uint8_t>(soa[static_cast<size_t>(P_C2) * F + f]));
    }
}

void uploadFeaturesSoA(cf_ctx* ctx, FeatTables& tabs, const int8_t* soa, const float* thrTF, int F, int T) {
    tabs.soa.resize(static_cast<size_t>(F) * kFeatParams);
    tabs.thrTF.resize(static_cast<size_t>(F) * T);
    checkFeaturesSoA(soa, F);
    // the reference's thresholds are responses of sampled pixels that are never NaN
    // (random_tree_image.cpp:679-690); a NaN threshold has no rank among the others
    if (thrTF) {
        for (size_t k = 0; k < static_cast<size_t>(F) * T; k++) {
            if (std::isnan(thrTF[k])) throw Error(CF_ERR_ARG, "NaN threshold");
        }
    }
    CF_HIP(hipMemcpyAsync(tabs.soa.get(), soa, static_cast<size_t>(F) * kFeatParams, hipMemcpyHostToDevice, ctx->stream));
    if (thrTF) {
        CF_HIP(hipMemcpyAsync(tabs.thrTF.get(), thrTF, static_cast<size_t>(F) * T * sizeof(float), hipMemcpyHostToDevice,
                ctx->stream));
    } else {
        CF_HIP(hipMemsetAsync(tabs.thrTF.get(), 0, static_cast<size_t>(F) * T * sizeof(float), ctx->stream));
    }
    launchPrepareTables(ctx->stream, 1, F, T, tabs);
}

// Interval histograms of one node (slot 0) into dG.
void singleNodeBins(cf_ctx* ctx, cf_images* im, const cf_sample* samples, int n, const FeatTables& tabs, int C,
        DevBuf<cf_sample>& dS, DevBuf<Chunk>& dC, DevBuf<uint32_t>& dG) {
    dS.resize(std::max(n, 1));
    CF_HIP(hipMemcpyAsync(dS.get(), samples, static_cast<size_t>(n) * sizeof(cf_sample), hipMemcpyHostToDevice, ctx->stream));
    std::vector<Chunk> chunks;
    for (int b = 0; b < n; b += 4096) chunks.push_back(Chunk{0, 0, b, std::min(4096, n - b)});
    dC.resize(std::max<size_t>(chunks.size(), 1));
    if (!chunks.empty()) {
        CF_HIP(hipMemcpyAsync(dC.get(), chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice, ctx->stream));
    }
    const size_t perSlot = static_cast<size_t>(tabs.F) * (tabs.T + 1) * C;
    dG.resize(perSlot);
    CF_HIP(hipMemsetAsync(dG.get(), 0, perSlot * sizeof(uint32_t), ctx->stream));
    launchEvalHistograms(ctx->stream, im->set(), dS.get(), dC.get(), static_cast<int>(chunks.size()), tabs, C,
            dG.get());
}

}  // namespace

namespace cf {

std::string& hostErrorSlot() {
    thread_local std::string message;
    return message;
}

void bindForest(cf_forest* fo, cf_ctx* ctx) {
    if (!fo->ctx) fo->ctx = ctx;
    if (fo->ctx != ctx) throw Error(CF_ERR_ARG, "forest belongs to another context");
}

// the packer's feature columns are the feature table's parameter rows
constexpr bool sameColumn(FeatColumn c, FeatParam p) { return static_cast<int>(c) == static_cast<int>(p); }
static_assert(sameColumn(FC_TYPE, P_TYPE) && sameColumn(FC_O1X, P_O1X) && sameColumn(FC_O2Y, P_O2Y) &&
                sameColumn(FC_R1X, P_R1X) && sameColumn(FC_R2Y, P_R2Y) && sameColumn(FC_C1, P_C1) &&
                sameColumn(FC_C2, P_C2),
        "feature columns");
static_assert(sizeof(NodeBlock) == kBlockWords * sizeof(uint32_t), "node block size");
static_assert(postStride(5) == rowStride(5) && postStride(8) == rowStride(8), "posterior row stride");

void uploadForest(cf_forest* fo) {
    if (fo->uploaded) return;
    if (!fo->ctx) throw Error(CF_ERR_ARG, "forest has no context");
    if (fo->post.size() != fo->trees.size()) throw Error(CF_ERR_RUNTIME, "histograms not normalized");
    std::vector<const cf_tree_node*> trees;
    std::vector<int> sizes;
    std::vector<const double*> post;
    for (size_t t = 0; t < fo->trees.size(); t++) {
        trees.push_back(fo->trees[t].nodes.data());
        sizes.push_back(static_cast<int>(fo->trees[t].nodes.size()));
        post.push_back(fo->post[t].data());
    }
    PackedForest pf;
    std::string msg;
    const int rc = packForest(trees, sizes, post, fo->C, pf, msg);
    if (rc != CF_OK) throw Error(rc, msg);
    if (pf.blocks.empty()) pf.blocks.resize(kBlockWords, 0u);
    if (pf.post.empty()) pf.post.resize(postStride(fo->C));
    const size_t numBlocks = pf.blocks.size() / kBlockWords;
    fo->dBlocks.resize(numBlocks);
    fo->dPost.resize(pf.post.size());
    fo->dRoots.resize(pf.roots.size());
    hipStream_t s = fo->ctx->stream;
    CF_HIP(hipMemcpyAsync(fo->dBlocks.get(), pf.blocks.data(), numBlocks * sizeof(NodeBlock), hipMemcpyHostToDevice, s));
    CF_HIP(hipMemcpyAsync(fo->dPost.get(), pf.post.data(), pf.post.size() * sizeof(double), hipMemcpyHostToDevice, s));
    CF_HIP(hipMemcpyAsync(fo->dRoots.get(), pf.roots.data(), pf.roots.size() * sizeof(int32_t), hipMemcpyHostToDevice,
            s));
    CF_HIP(hipStreamSynchronize(s));
    fo->uploaded = true;
}

}  // namespace cf

cf::ImageSet cf_images::set() const { return cf::ImageSet{col.get(), dep.get(), W, H}; }

namespace {
// true for page-locked host memory (a copy from the device into it runs as one DMA, asynchronously)
bool hostPinned(const void* p) {
    hipPointerAttribute_t a;
    if (hipPointerGetAttributes(&a, p) != hipSuccess) {
        (void)hipGetLastError();
        return false;
    }
    return a.type == hipMemoryTypeHost;
}

// memcpy split over up to 8 host threads (at least 4 MB each)
void parallelCopy(void* dst, const void* src, size_t bytes) {
    constexpr size_t kMinPart = 4u << 20;
    const size_t parts = std::max<size_t>(1, std::min<size_t>(8, bytes / kMinPart));
    const size_t step = (bytes + parts - 1) / parts;
    std::vector<std::thread> th;
    for (size_t i = 1; i < parts; i++) {
        const size_t b = i * step, e = std::min(bytes, b + step);
        if (b < e) th.emplace_back([=] { std::memcpy(static_cast<char*>(dst) + b, static_cast<const char*>(src) + b, e - b); });
    }
    std::memcpy(dst, src, std::min(bytes, step));
    for (auto& t : th) t.join();
}

void ensureEvents(std::vector<hipEvent_t>& v, int n) {
    while (static_cast<int>(v.size()) < n) {
        hipEvent_t e;
        CF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        v.push_back(e);
    }
}

// a stream destroyed at scope exit, after its work is done
struct StreamGuard {
    hipStream_t s = nullptr;
    ~StreamGuard() {
        if (s) {
            (void)hipStreamSynchronize(s);
            (void)hipStreamDestroy(s);
        }
    }
};
}  // namespace

extern "C" {

int cf_ctx_create(int device, cf_ctx** out) {
    if (!out) return CF_ERR_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return CF_ERR_HIP;
    if (device < 0 || device >= n) return CF_ERR_ARG;
    if (hipSetDevice(device) != hipSuccess) return CF_ERR_HIP;
    // the allocation-free prediction's device array: its address on this device, resolved at set-up; a
    // failure is kept and reported by that entry point (every other call works without the array)
    try {
        resolvePredictScratch(device);
    } catch (const Error&) {
    }
    auto* c = new cf_ctx();
    c->device = device;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return CF_ERR_HIP;
    }
    *out = c;
    return CF_OK;
}

int cf_ctx_destroy(cf_ctx* ctx) {
    if (!ctx) return CF_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->scratch0.release();
    ctx->scratch1.release();
    ctx->scratch2.release();
    for (hipEvent_t e : ctx->launchDone) (void)hipEventDestroy(e);
    for (hipEvent_t e : ctx->copyDone) (void)hipEventDestroy(e);
    if (ctx->hostStage) (void)hipHostFree(ctx->hostStage);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CF_OK;
}

const char* cf_last_error(const cf_ctx* ctx) { return ctx ? ctx->error.c_str() : hostErrorSlot().c_str(); }

int cf_ctx_synchronize(cf_ctx* ctx) {
    return guarded(ctx, [&] { CF_HIP(hipStreamSynchronize(ctx->stream)); });
}

int cf_images_create(cf_ctx* ctx, int count, int width, int height, cf_images** out) {
    return guarded(ctx, [&] {
        require(out != nullptr, "null output");
        require(count > 0 && width > 0 && height > 0, "image set dimensions must be positive");
        require(width <= 32767 && height <= 32767, "image too large");
        require(static_cast<uint64_t>(width) * height * sizeof(int2) < (1ull << 32),
                "image too large (an integral plane must stay below 4 GiB)");
        CF_HIP(hipSetDevice(ctx->device));
        auto im = std::make_unique<cf_images>();
        im->ctx = ctx;
        im->count = count;
        im->W = width;
        im->H = height;
        const size_t px = static_cast<size_t>(width) * height;
        im->col.resize(static_cast<size_t>(count) * 3 * px);
        im->dep.resize(static_cast<size_t>(count) * px);
        im->desc.resize(count);
        std::vector<DevImage> d(count);
        for (int i = 0; i < count; i++) {
            d[i].col = im->col.get() + static_cast<size_t>(i) * 3 * px;
            d[i].dep = im->dep.get() + static_cast<size_t>(i) * px;
            d[i].W = width;
            d[i].H = height;
        }
        CF_HIP(hipMemcpyAsync(im->desc.get(), d.data(), count * sizeof(DevImage), hipMemcpyHostToDevice, ctx->stream));
        CF_HIP(hipMemsetAsync(im->col.get(), 0, im->col.bytes(), ctx->stream));
        CF_HIP(hipMemsetAsync(im->dep.get(), 0, im->dep.bytes(), ctx->stream));
        CF_HIP(hipStreamSynchronize(ctx->stream));
        im->labels.assign(static_cast<size_t>(count) * px, 0);
        im->labelsSet.assign(count, 0);
        im->depthHost.assign(static_cast<size_t>(count) * px, 0);
        im->depthSet.assign(count, 0);
        *out = im.release();
    });
}

int cf_images_destroy(cf_images* images) {
    if (!images) return CF_ERR_ARG;
    cf_ctx* ctx = images->ctx;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    (void)hipStreamSynchronize(ctx->stream);
    delete images;
    return CF_OK;
}

int cf_images_set_rgbd_u8(cf_images* im, int index, const uint8_t* rgb, const uint16_t* depth) {
    if (!im) return CF_ERR_ARG;
    return guarded(im->ctx, [&] {
        require(index >= 0 && index < im->count, "image index out of range");
        require(rgb && depth, "null input");
        cf_ctx* ctx = im->ctx;
        const size_t px = static_cast<size_t>(im->W) * im->H;
        checkDepthMm(depth, px);
        ctx->scratch0.resize(px * 3 + px * 2 + 64);
        ctx->scratch1.resize(px * 5 * sizeof(uint32_t));
        uint8_t* dRgb = ctx->scratch0.get();
        uint16_t* dDepth = reinterpret_cast<uint16_t*>(ctx->scratch0.get() + ((px * 3 + 15) / 16) * 16);
        CF_HIP(hipMemcpyAsync(dRgb, rgb, px * 3, hipMemcpyHostToDevice, ctx->stream));
        CF_HIP(hipMemcpyAsync(dDepth, depth, px * 2, hipMemcpyHostToDevice, ctx->stream));
        launchIntegralU8(ctx->stream, dRgb, dDepth, im->W, im->H, reinterpret_cast<uint32_t*>(ctx->scratch1.get()),
                im->col.get() + static_cast<size_t>(index) * 3 * px, im->dep.get() + static_cast<size_t>(index) * px);
        int32_t* hd = im->depthHost.data() + static_cast<size_t>(index) * px;
        for (size_t p = 0; p < px; p++) hd[p] = depth[p];
        im->depthSet[index] = 1;
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_images_set_rgbd_u8_padded(cf_images* im, int index, const uint8_t* rgb, const uint16_t* depth, int width,
        int height) {
    if (!im) return CF_ERR_ARG;
    if (width == im->W && height == im->H) return cf_images_set_rgbd_u8(im, index, rgb, depth);
    return guarded(im->ctx, [&] {
        require(index >= 0 && index < im->count, "image index out of range");
        require(rgb && depth, "null input");
        require(width > 0 && height > 0 && width <= im->W && height <= im->H, "image larger than the set's size");
        cf_ctx* ctx = im->ctx;
        const size_t px = static_cast<size_t>(width) * height;
        const size_t slot = static_cast<size_t>(im->W) * im->H;
        checkDepthMm(depth, px);
        ctx->scratch0.resize(px * 3 + px * 2 + 64);
        ctx->scratch1.resize(px * 5 * sizeof(uint32_t));
        ctx->scratch2.resize(px * 3 * sizeof(float) + px * sizeof(int2) + 256);
        uint8_t* dRgb = ctx->scratch0.get();
        uint16_t* dDepth = reinterpret_cast<uint16_t*>(ctx->scratch0.get() + ((px * 3 + 15) / 16) * 16);
        float* tCol = reinterpret_cast<float*>(ctx->scratch2.get());
        int2* tDep = reinterpret_cast<int2*>(ctx->scratch2.get() + ((px * 3 * sizeof(float) + 255) / 256) * 256);
        CF_HIP(hipMemcpyAsync(dRgb, rgb, px * 3, hipMemcpyHostToDevice, ctx->stream));
        CF_HIP(hipMemcpyAsync(dDepth, depth, px * 2, hipMemcpyHostToDevice, ctx->stream));
        launchIntegralU8(ctx->stream, dRgb, dDepth, width, height, reinterpret_cast<uint32_t*>(ctx->scratch1.get()),
                tCol, tDep);
        launchIntegralPad(ctx->stream, tCol, tDep, width, height, im->W, im->H,
                im->col.get() + static_cast<size_t>(index) * 3 * slot, im->dep.get() + static_cast<size_t>(index) * slot);
        int32_t* hd = im->depthHost.data() + static_cast<size_t>(index) * slot;
        std::fill(hd, hd + slot, 0);
        for (int y = 0; y < height; y++) {
            for (int x = 0; x < width; x++) hd[static_cast<size_t>(y) * im->W + x] = depth[static_cast<size_t>(y) * width + x];
        }
        im->depthSet[index] = 1;
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_images_set_rgbd_f32(cf_images* im, int index, const float* color, const int32_t* depth) {
    if (!im) return CF_ERR_ARG;
    return guarded(im->ctx, [&] {
        require(index >= 0 && index < im->count, "image index out of range");
        require(color && depth, "null input");
        cf_ctx* ctx = im->ctx;
        const size_t px = static_cast<size_t>(im->W) * im->H;
        for (size_t p = 0; p < px; p++) {
            if (depth[p] > 50000) throw Error(CF_ERR_ARG, "illegal depth value (> 50000 mm)");
        }
        ctx->scratch0.resize(px * 3 * sizeof(float) + px * sizeof(int32_t));
        ctx->scratch1.resize(px * 5 * sizeof(uint32_t));
        ctx->scratch2.resize(px * 3 * sizeof(double));
        float* dCol = reinterpret_cast<float*>(ctx->scratch0.get());
        int32_t* dDep = reinterpret_cast<int32_t*>(ctx->scratch0.get() + px * 3 * sizeof(float));
        CF_HIP(hipMemcpyAsync(dCol, color, px * 3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        CF_HIP(hipMemcpyAsync(dDep, depth, px * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        launchIntegralF32(ctx->stream, dCol, dDep, im->W, im->H, reinterpret_cast<uint32_t*>(ctx->scratch1.get()),
                reinterpret_cast<double*>(ctx->scratch2.get()), im->col.get() + static_cast<size_t>(index) * 3 * px,
                im->dep.get() + static_cast<size_t>(index) * px);
        int32_t* hd = im->depthHost.data() + static_cast<size_t>(index) * px;
        for (size_t p = 0; p < px; p++) hd[p] = depth[p] > 0 ? depth[p] : 0;
        im->depthSet[index] = 1;
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_images_set_rgbd_f32_padded(cf_images* im, int index, const float* color, const int32_t* depth, int width,
        int height) {
    if (!im) return CF_ERR_ARG;
    if (width == im->W && height == im->H) return cf_images_set_rgbd_f32(im, index, color, depth);
    return guarded(im->ctx, [&] {
        require(index >= 0 && index < im->count, "image index out of range");
        require(color && depth, "null input");
        require(width > 0 && height > 0 && width <= im->W && height <= im->H, "image larger than the set's size");
        cf_ctx* ctx = im->ctx;
        const size_t px = static_cast<size_t>(width) * height;
        const size_t slot = static_cast<size_t>(im->W) * im->H;
        for (size_t p = 0; p < px; p++) {
            if (depth[p] > 50000) throw Error(CF_ERR_ARG, "illegal depth value (> 50000 mm)");
        }
        auto align = [](size_t b) { return (b + 255) / 256 * 256; };
        ctx->scratch0.resize(align(px * 3 * sizeof(float)) + px * sizeof(int32_t));
        ctx->scratch1.resize(px * 5 * sizeof(uint32_t));
        ctx->scratch2.resize(align(px * 3 * sizeof(double)) + align(px * 3 * sizeof(float)) + px * sizeof(int2));
        float* dCol = reinterpret_cast<float*>(ctx->scratch0.get());
        int32_t* dDep = reinterpret_cast<int32_t*>(ctx->scratch0.get() + align(px * 3 * sizeof(float)));
        double* work = reinterpret_cast<double*>(ctx->scratch2.get());
        float* tCol = reinterpret_cast<float*>(ctx->scratch2.get() + align(px * 3 * sizeof(double)));
        int2* tDep = reinterpret_cast<int2*>(ctx->scratch2.get() + align(px * 3 * sizeof(double)) +
                align(px * 3 * sizeof(float)));
        CF_HIP(hipMemcpyAsync(dCol, color, px * 3 * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        CF_HIP(hipMemcpyAsync(dDep, depth, px * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream));
        launchIntegralF32(ctx->stream, dCol, dDep, width, height, reinterpret_cast<uint32_t*>(ctx->scratch1.get()), work,
                tCol, tDep);
        launchIntegralPad(ctx->stream, tCol, tDep, width, height, im->W, im->H,
                im->col.get() + static_cast<size_t>(index) * 3 * slot, im->dep.get() + static_cast<size_t>(index) * slot);
        int32_t* hd = im->depthHost.data() + static_cast<size_t>(index) * slot;
        std::fill(hd, hd + slot, 0);
        for (int y = 0; y < height; y++) {
            for (int x = 0; x < width; x++) {
                const int32_t d = depth[static_cast<size_t>(y) * width + x];
                hd[static_cast<size_t>(y) * im->W + x] = d > 0 ? d : 0;
            }
        }
        im->depthSet[index] = 1;
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_images_set_labels(cf_images* im, int index, const uint8_t* labels) {
    if (!im) return CF_ERR_ARG;
    return guarded(im->ctx, [&] {
        require(index >= 0 && index < im->count, "image index out of range");
        require(labels != nullptr, "null input");
        const size_t px = static_cast<size_t>(im->W) * im->H;
        std::memcpy(im->labels.data() + static_cast<size_t>(index) * px, labels, px);
        im->labelsSet[index] = 1;
    });
}

int cf_images_get_integrals(cf_images* im, int index, float* color, int32_t* depth, int32_t* valid) {
    if (!im) return CF_ERR_ARG;
    return guarded(im->ctx, [&] {
        require(index >= 0 && index < im->count, "image index out of range");
        cf_ctx* ctx = im->ctx;
        const size_t px = static_cast<size_t>(im->W) * im->H;
        if (color) {
            CF_HIP(hipMemcpyAsync(color, im->col.get() + static_cast<size_t>(index) * 3 * px, px * 3 * sizeof(float),
                    hipMemcpyDeviceToHost, ctx->stream));
        }
        std::vector<int2> d(px);
        CF_HIP(hipMemcpyAsync(d.data(), im->dep.get() + static_cast<size_t>(index) * px, px * sizeof(int2),
                hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipStreamSynchronize(ctx->stream));
        for (size_t p = 0; p < px; p++) {
            if (depth) depth[p] = d[p].x;
            if (valid) valid[p] = d[p].y;
        }
    });
}

int cf_generate_features(cf_ctx* ctx, cf_images* im, const cf_sample* samples, int n, int seed, int F, int T,
        int boxRadius, int regionSize, int useDepth, int8_t* featuresOut, float* thresholdsOut) {
    return guarded(ctx, [&] {
        require(im && samples && featuresOut && thresholdsOut, "null argument");
        require(n > 0 && F > 0 && T > 0, "sizes must be positive");
        require(boxRadius >= 0 && regionSize >= 1, "invalid boxRadius / regionSize");
        checkSamples(im, samples, n);
        DevBuf<cf_sample> dS(n);
        CF_HIP(hipMemcpyAsync(dS.get(), samples, static_cast<size_t>(n) * sizeof(cf_sample), hipMemcpyHostToDevice,
                ctx->stream));
        FeatGenJob j{dS.get(), n, seed, 0, 0};
        DevBuf<FeatGenJob> dJ(1);
        CF_HIP(hipMemcpyAsync(dJ.get(), &j, sizeof(j), hipMemcpyHostToDevice, ctx->stream));
        FeatTables tabs;
        launchFeatureGeneration(ctx->stream, im->set(), dJ.get(), 1, F, T, boxRadius, regionSize, useDepth, &seed, tabs);
        CF_HIP(hipMemcpyAsync(featuresOut, tabs.soa.get(), static_cast<size_t>(F) * kFeatParams, hipMemcpyDeviceToHost,
                ctx->stream));
        CF_HIP(hipMemcpyAsync(thresholdsOut, tabs.thrTF.get(), static_cast<size_t>(F) * T * sizeof(float),
                hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_feature_responses(cf_ctx* ctx, cf_images* im, const cf_sample* samples, int n, const int8_t* features, int F,
        double* out) {
    return guarded(ctx, [&] {
        require(im && samples && features && out, "null argument");
        require(n > 0 && F > 0, "sizes must be positive");
        require(static_cast<size_t>(n) * F < (1u << 31), "n x F too large for one call");
        checkSamples(im, samples, n);
        checkFeaturesSoA(features, F);
        std::vector<FeatRec> fr(F);
        for (int f = 0; f < F; f++) {
            std::memset(&fr[f], 0, sizeof(FeatRec));
            for (int p = 0; p < kFeatParams; p++) fr[f].p[p] = features[p * F + f];
        }
        DevBuf<cf_sample> dS(n);
        DevBuf<FeatRec> dF(F);
        DevBuf<double> dO(static_cast<size_t>(n) * F);
        CF_HIP(hipMemcpyAsync(dS.get(), samples, static_cast<size_t>(n) * sizeof(cf_sample), hipMemcpyHostToDevice, ctx->stream));
        CF_HIP(hipMemcpyAsync(dF.get(), fr.data(), F * sizeof(FeatRec), hipMemcpyHostToDevice, ctx->stream));
        launchResponses(ctx->stream, im->desc.get(), dS.get(), n, dF.get(), F, dO.get());
        CF_HIP(hipMemcpyAsync(out, dO.get(), static_cast<size_t>(n) * F * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_feature_responses_paired(cf_ctx* ctx, cf_images* im, const cf_sample* samples, int n, const int8_t* features,
        double* out) {
    return guarded(ctx, [&] {
        require(im && samples && features && out, "null argument");
        require(n > 0, "sizes must be positive");
        checkSamples(im, samples, n);
        checkFeaturesSoA(features, n);
        std::vector<FeatRec> fr(n);
        for (int f = 0; f < n; f++) {
            std::memset(&fr[f], 0, sizeof(FeatRec));
            for (int p = 0; p < kFeatParams; p++) fr[f].p[p] = features[static_cast<size_t>(p) * n + f];
        }
        DevBuf<cf_sample> dS(n);
        DevBuf<FeatRec> dF(n);
        DevBuf<double> dO(n);
        CF_HIP(hipMemcpyAsync(dS.get(), samples, static_cast<size_t>(n) * sizeof(cf_sample), hipMemcpyHostToDevice, ctx->stream));
        CF_HIP(hipMemcpyAsync(dF.get(), fr.data(), static_cast<size_t>(n) * sizeof(FeatRec), hipMemcpyHostToDevice, ctx->stream));
        launchResponsesPaired(ctx->stream, im->desc.get(), dS.get(), n, dF.get(), dO.get());
        CF_HIP(hipMemcpyAsync(out, dO.get(), static_cast<size_t>(n) * sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_node_counters(cf_ctx* ctx, cf_images* im, const cf_sample* samples, int n, const int8_t* features,
        const float* thresholds, int F, int T, int C, uint32_t* counters) {
    return guarded(ctx, [&] {
        require(im && samples && features && thresholds && counters, "null argument");
        require(n > 0 && F > 0 && T > 0 && C > 0 && C <= kMaxClasses, "invalid sizes");
        checkSamples(im, samples, n);
        for (int i = 0; i < n; i++) require(samples[i].label < C, "sample label >= number of classes");
        FeatTables tabs;
        uploadFeaturesSoA(ctx, tabs, features, thresholds, F, T);
        DevBuf<cf_sample> dS;
        DevBuf<Chunk> dC;
        DevBuf<uint32_t> dG;
        singleNodeBins(ctx, im, samples, n, tabs, C, dS, dC, dG);
        DevBuf<uint32_t> dCnt(static_cast<size_t>(F) * T * C * 2);
        launchCountersFromBins(ctx->stream, dG.get(), tabs, C, dCnt.get());
        CF_HIP(hipMemcpyAsync(counters, dCnt.get(), dCnt.bytes(), hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_node_best_split(cf_ctx* ctx, cf_images* im, const cf_sample* samples, int n, const int8_t* features,
        const float* thresholds, int F, int T, int C, const uint32_t* histogram, int* bestFeature, int* bestThreshold,
        double* bestScore) {
    return guarded(ctx, [&] {
        require(im && samples && features && thresholds && histogram && bestFeature && bestThreshold && bestScore,
                "null argument");
        require(n > 0 && F > 0 && T > 0 && C > 0 && C <= kMaxClasses, "invalid sizes");
        checkSamples(im, samples, n);
        for (int i = 0; i < n; i++) require(samples[i].label < C, "sample label >= number of classes");
        FeatTables tabs;
        uploadFeaturesSoA(ctx, tabs, features, thresholds, F, T);
        DevBuf<cf_sample> dS;
        DevBuf<Chunk> dC;
        DevBuf<uint32_t> dG;
        singleNodeBins(ctx, im, samples, n, tabs, C, dS, dC, dG);
        DevBuf<uint32_t> dH(C);
        DevBuf<int> dT(1);
        DevBuf<FeatSummary> dSumm(F);
        DevBuf<NodeBest> dBest(1);
        const int zero = 0;
        CF_HIP(hipMemcpyAsync(dH.get(), histogram, C * sizeof(uint32_t), hipMemcpyHostToDevice, ctx->stream));
        CF_HIP(hipMemcpyAsync(dT.get(), &zero, sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        launchScores(ctx->stream, dG.get(), dH.get(), dT.get(), 1, tabs, C, dSumm.get(), nullptr, -1);
        launchReduce(ctx->stream, dSumm.get(), 1, F, dBest.get());
        NodeBest b;
        CF_HIP(hipMemcpyAsync(&b, dBest.get(), sizeof(b), hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipStreamSynchronize(ctx->stream));
        if (b.flag) {
            DevBuf<double> dAll(static_cast<size_t>(F) * T);
            launchScores(ctx->stream, dG.get(), dH.get(), dT.get(), 1, tabs, C, dSumm.get(), dAll.get(), 0);
            launchSequentialArgmax(ctx->stream, dAll.get(), T, F, dBest.get());
            CF_HIP(hipMemcpyAsync(&b, dBest.get(), sizeof(b), hipMemcpyDeviceToHost, ctx->stream));
            CF_HIP(hipStreamSynchronize(ctx->stream));
        }
        *bestFeature = b.feat;
        *bestThreshold = b.thresh;
        *bestScore = b.score;
    });
}

int cf_scores(cf_ctx* ctx, const uint32_t* counters, const uint32_t* histogram, int F, int T, int C, double* scores) {
    return guarded(ctx, [&] {
        require(counters && histogram && scores, "null argument");
        require(F > 0 && T > 0 && C > 0 && C <= kMaxClasses, "invalid sizes");
        DevBuf<uint32_t> dCnt(static_cast<size_t>(F) * T * C * 2), dH(C);
        DevBuf<double> dAll(static_cast<size_t>(F) * T);
        CF_HIP(hipMemcpyAsync(dCnt.get(), counters, dCnt.bytes(), hipMemcpyHostToDevice, ctx->stream));
        CF_HIP(hipMemcpyAsync(dH.get(), histogram, C * 4, hipMemcpyHostToDevice, ctx->stream));
        launchScoresFromCounters(ctx->stream, dCnt.get(), dH.get(), F, T, C, dAll.get());
        CF_HIP(hipMemcpyAsync(scores, dAll.get(), dAll.bytes(), hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_argmax_scores(cf_ctx* ctx, const double* scores, int F, int T, int* bestFeature, int* bestThreshold,
        double* bestScore) {
    return guarded(ctx, [&] {
        require(scores && bestFeature && bestThreshold && bestScore, "null argument");
        require(F > 0 && T > 0, "sizes must be positive");
        DevBuf<double> dS(static_cast<size_t>(F) * T);
        DevBuf<NodeBest> dB(1);
        CF_HIP(hipMemcpyAsync(dS.get(), scores, dS.bytes(), hipMemcpyHostToDevice, ctx->stream));
        launchSequentialArgmax(ctx->stream, dS.get(), T, F, dB.get());
        NodeBest nb;
        CF_HIP(hipMemcpyAsync(&nb, dB.get(), sizeof(NodeBest), hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipStreamSynchronize(ctx->stream));
        *bestFeature = nb.feat;
        *bestThreshold = nb.thresh;
        *bestScore = nb.score;
    });
}

int cf_forest_sample(cf_ctx* ctx, cf_images* im, const cf_train_config* cfg, int numTrees, int numLabels,
        cf_forest** out) {
    return guarded(ctx, [&] {
        require(im && cfg && out, "null argument");
        require(numTrees > 0, "cannot construct empty forest");
        auto fo = std::make_unique<cf_forest>();
        fo->ctx = ctx;
        fo->prior = labelPrior(im, numLabels);
        fo->C = static_cast<int>(fo->prior.size());
        // RandomForestImage::train: SEED from the configured seed, tree t uses
        // RandomSource(SEED + t) (random_forest_image.cpp:150-173): first the tree's
        // images (maxImages), then its label prior over them and its pixels.  Trees
        // sample independently, so they sample concurrently.
        HostRandomSource master(cfg->random_seed);
        const int SEED = master.uniformSampler(0xFFFF).next();
        fo->roots.resize(numTrees);
        fo->rsSeeds.resize(numTrees);
        fo->treeIds.resize(numTrees);
        std::vector<std::vector<uint32_t>> priors(numTrees);
        std::vector<std::string> errors(numTrees);
        std::vector<std::thread> threads;
        for (int t = 0; t < numTrees; t++) {
            threads.emplace_back([&, t] {
                try {
                    HostRandomSource rs(SEED + t);
                    const std::vector<int> ids = sampleImages(im->count, cfg->max_images, rs);
                    priors[t] = static_cast<int>(ids.size()) == im->count ? fo->prior : labelPrior(im, ids, numLabels);
                    const size_t perImage = cfg->samples_per_image / numTrees;
                    fo->roots[t] = cfg->subsampling_pixel_uniform ? samplePixelUniform(im, ids, *cfg, rs, perImage)
                                                                  : sampleClassUniform(im, ids, priors[t], *cfg, rs,
                                                                            perImage);
                    fo->rsSeeds[t] = rs.seed();
                    fo->treeIds[t] = t;
                } catch (const std::exception& e) {
                    errors[t] = e.what();
                }
            });
        }
        for (auto& th : threads) th.join();
        for (const auto& e : errors) {
            if (!e.empty()) throw Error(CF_ERR_RUNTIME, e);
        }
        // trees that saw different images have different priors (RandomTreeImage keeps its own)
        bool same = true;
        for (const auto& p : priors) {
            if (p.size() != fo->prior.size()) throw Error(CF_ERR_RUNTIME, "different number of classes in trees");
            same = same && p == fo->prior;
        }
        if (!same) fo->treePriors = priors;
        *out = fo.release();
    });
}

int cf_forest_grow(cf_ctx* ctx, cf_images* im, cf_forest* fo, const cf_train_config* cfg, int mode) {
    return guarded(ctx, [&] {
        require(im && fo && cfg, "null argument");
        require(!fo->roots.empty(), "forest has no sampled roots");
        std::vector<TreeJob> jobs;
        for (size_t t = 0; t < fo->roots.size(); t++) {
            jobs.push_back(TreeJob{fo->treeIds[t], HostRandomSource(fo->rsSeeds[t]), &fo->roots[t]});
        }
        fo->stats = TrainStats();
        fo->trees.clear();
        fo->post.clear();
        fo->uploaded = false;
        trainTrees(ctx, im, *cfg, fo->C, jobs, mode, fo->trees, fo->stats);
    });
}

int cf_forest_train(cf_ctx* ctx, cf_images* im, const cf_train_config* cfg, int numTrees, int numLabels, int mode,
        cf_forest** out) {
    if (!out) return CF_ERR_ARG;
    cf_forest* fo = nullptr;
    int st = cf_forest_sample(ctx, im, cfg, numTrees, numLabels, &fo);
    if (st != CF_OK) return st;
    st = cf_forest_grow(ctx, im, fo, cfg, mode);
    if (st != CF_OK) {
        cf_forest_destroy(fo);
        return st;
    }
    *out = fo;
    return CF_OK;
}

int cf_ctx_set_profiling(cf_ctx* ctx, int enable) {
    return guarded(ctx, [&] { ctx->profiling = enable != 0; });
}

int cf_forest_profile(const cf_forest* fo, double* out, int n) {
    if (!fo || !out) return CF_ERR_ARG;
    const TrainStats& st = fo->stats;
    const double v[] = {st.evaluations, st.seconds, st.evalKernelMs, st.evalLaunches, st.responses, st.featGenMs,
            st.scoreMs, st.partitionMs, static_cast<double>(st.levels), static_cast<double>(st.replays), st.nodes,
            st.histMs, static_cast<double>(st.groupBatchesMax), static_cast<double>(st.slotBatchesMax),
            static_cast<double>(std::max_element(st.shapeLaunches, st.shapeLaunches + 16) - st.shapeLaunches),
            static_cast<double>(st.levelGraphs)};
    const int k = static_cast<int>(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < k; i++) out[i] = v[i];
    return k;
}

int cf_tree_train_from_samples(cf_ctx* ctx, cf_images* im, const cf_train_config* cfg, int treeId, int C, int rsSeed,
        const cf_sample* roots, int n, int mode, cf_forest** out) {
    return guarded(ctx, [&] {
        require(im && cfg && roots && out, "null argument");
        require(n > 0, "no samples");
        checkSamples(im, roots, n);
        auto fo = std::make_unique<cf_forest>();
        fo->ctx = ctx;
        fo->C = C;
        fo->prior.assign(C, 1u);
        std::vector<TreeJob> jobs;
        fo->roots.emplace_back(roots, roots + n);
        fo->treeIds.push_back(treeId);
        fo->rsSeeds.push_back(rsSeed);
        jobs.push_back(TreeJob{treeId, HostRandomSource(rsSeed), &fo->roots.back()});
        trainTrees(ctx, im, *cfg, C, jobs, mode, fo->trees, fo->stats);
        *out = fo.release();
    });
}

int cf_forest_create(cf_ctx* ctx, int numTrees, const int* numNodes, const cf_tree_node* nodes,
        const uint32_t* histograms, int C, const uint32_t* prior, cf_forest** out) {
    return guardedHost(ctx, [&] {
        require(numNodes && nodes && histograms && out, "null argument");
        require(numTrees > 0 && C > 0 && C <= kMaxClasses, "invalid sizes");
        auto fo = std::make_unique<cf_forest>();
        fo->ctx = ctx;
        fo->C = C;
        fo->prior.assign(C, 1u);
        if (prior) fo->prior.assign(prior, prior + C);
        size_t off = 0;
        for (int t = 0; t < numTrees; t++) {
            require(numNodes[t] > 0, "empty tree");
            TreeBuild tb;
            tb.nodes.assign(nodes + off, nodes + off + numNodes[t]);
            tb.hist.assign(histograms + off * C, histograms + (off + numNodes[t]) * C);
            for (int i = 0; i < numNodes[t]; i++) {
                const auto& n = tb.nodes[i];
                if (!n.is_leaf) {
                    // children after their parent (guarantees traversal terminates), stored as a pair
                    require(n.left > i && n.left + 1 < numNodes[t] && n.right == n.left + 1, "invalid child indices");
                    checkFeature(n.feature[P_TYPE], static_cast<uint8_t>(n.feature[P_C1]),
                            static_cast<uint8_t>(n.feature[P_C2]));
                }
            }
            fo->trees.push_back(std::move(tb));
            off += numNodes[t];
        }
        *out = fo.release();
    });
}

int cf_forest_destroy(cf_forest* fo) {
    if (!fo) return CF_ERR_ARG;
    if (!fo->ctx) {
        delete fo;
        return CF_OK;
    }
    std::lock_guard<std::recursive_mutex> lock(fo->ctx->mu);
    delete fo;
    return CF_OK;
}

int cf_forest_num_trees(const cf_forest* fo) { return fo ? static_cast<int>(fo->trees.size()) : CF_ERR_ARG; }
int cf_forest_num_classes(const cf_forest* fo) { return fo ? fo->C : CF_ERR_ARG; }
int cf_forest_tree_size(const cf_forest* fo, int t) {
    if (!fo || t < 0 || t >= static_cast<int>(fo->trees.size())) return CF_ERR_ARG;
    return static_cast<int>(fo->trees[t].nodes.size());
}

int cf_forest_get_tree(const cf_forest* fo, int t, cf_tree_node* nodes, uint32_t* hist) {
    if (!fo || t < 0 || t >= static_cast<int>(fo->trees.size())) return CF_ERR_ARG;
    const auto& tb = fo->trees[t];
    if (nodes) std::memcpy(nodes, tb.nodes.data(), tb.nodes.size() * sizeof(cf_tree_node));
    if (hist) std::memcpy(hist, tb.hist.data(), tb.hist.size() * sizeof(uint32_t));
    return CF_OK;
}

int cf_forest_get_root_samples(const cf_forest* fo, int t, cf_sample* out, int capacity) {
    if (!fo || t < 0 || t >= static_cast<int>(fo->roots.size())) return CF_ERR_ARG;
    const auto& r = fo->roots[t];
    if (out) {
        if (capacity < static_cast<int>(r.size())) return CF_ERR_ARG;
        std::memcpy(out, r.data(), r.size() * sizeof(cf_sample));
    }
    return static_cast<int>(r.size());
}

int cf_forest_get_prior(const cf_forest* fo, uint32_t* prior) {
    if (!fo || !prior) return CF_ERR_ARG;
    std::memcpy(prior, fo->prior.data(), fo->prior.size() * sizeof(uint32_t));
    return static_cast<int>(fo->prior.size());
}

int cf_forest_get_tree_prior(const cf_forest* fo, int t, uint32_t* prior) {
    if (!fo || !prior) return CF_ERR_ARG;
    const size_t numTrees = std::max(fo->trees.size(), fo->roots.size());   // sampled, grown or imported
    if (t < 0 || static_cast<size_t>(t) >= numTrees) return CF_ERR_ARG;
    const std::vector<uint32_t>& p = fo->treePriors.empty() ? fo->prior : fo->treePriors[t];
    std::memcpy(prior, p.data(), p.size() * sizeof(uint32_t));
    return static_cast<int>(p.size());
}

int cf_forest_train_stats(const cf_forest* fo, double* evaluations, double* seconds) {
    if (!fo) return CF_ERR_ARG;
    if (evaluations) *evaluations = fo->stats.evaluations;
    if (seconds) *seconds = fo->stats.seconds;
    return fo->stats.replays;
}

// normalizeHistogram (random_tree.cpp:54-104) for every node.
int cf_forest_normalize(cf_forest* fo, double bias, int usePrior) {
    if (!fo) return CF_ERR_ARG;
    return guardedHost(fo->ctx, [&] {
        const int C = fo->C;
        std::vector<double> w(C);
        fo->post.assign(fo->trees.size(), {});
        for (size_t t = 0; t < fo->trees.size(); t++) {
            // each tree weighs by its own label prior (RandomTree::normalizeHistograms, random_tree.h:822-847)
            const std::vector<uint32_t>& prior = fo->treePriors.empty() ? fo->prior : fo->treePriors[t];
            double psum = 0;
            for (int i = 0; i < C; i++) psum += prior[i];
            for (int i = 0; i < C; i++) w[i] = prior[i] / psum;
            const TreeBuild& tb = fo->trees[t];
            std::vector<double>& out = fo->post[t];
            out.assign(tb.nodes.size() * C, 0.0);
            for (size_t n = 0; n < tb.nodes.size(); n++) {
                const uint32_t* h = &tb.hist[n * C];
                double* o = &out[n * C];
                double sum = 0;
                for (int i = 0; i < C; i++) sum += h[i];
                for (int i = 0; i < C; i++) {
                    o[i] = h[i] - bias * sum;
                    if (o[i] < 0.0) o[i] = 0.0;
                }
                sum = 0;
                for (int i = 0; i < C; i++) sum += o[i];
                if (sum > 0) {
                    for (int i = 0; i < C; i++) o[i] = usePrior ? o[i] / sum * w[i] : o[i] / sum;
                }
            }
        }
        fo->uploaded = false;
    });
}

int cf_forest_export_tree(cf_forest* fo, int tree, const char* path, const cf_train_config* config,
        const uint8_t* palette, const char* folder_training) {
    if (!fo) return CF_ERR_ARG;
    return guardedHost(fo->ctx, [&] {
        require(path && config, "null argument");
        writeTreeFile(fo, tree, path, *config, palette, folder_training);
    });
}

int cf_forest_import(cf_ctx* ctx, const char* const* paths, int count, cf_train_config* config_out, uint8_t* palette,
        cf_forest** out) {
    return guardedHost(ctx, [&] {
        require(out != nullptr, "null output");
        *out = readTreeFiles(ctx, paths, count, config_out, palette).release();
    });
}

int cf_forest_get_posteriors(const cf_forest* fo, int t, double* out) {
    if (!fo || !out || t < 0 || t >= static_cast<int>(fo->post.size())) return CF_ERR_ARG;
    std::memcpy(out, fo->post[t].data(), fo->post[t].size() * sizeof(double));
    return CF_OK;
}

int cf_predict_batch(cf_ctx* ctx, cf_forest* fo, cf_images* im, const int* indices, int count, float* probs,
        uint8_t* labels) {
    return guarded(ctx, [&] {
        require(fo && im && indices, "null argument");
        require(count > 0, "no frames");
        for (int i = 0; i < count; i++) require(indices[i] >= 0 && indices[i] < im->count, "frame index out of range");
        bindForest(fo, ctx);
        uploadForest(fo);
        const size_t plane = static_cast<size_t>(im->W) * im->H;
        // device staging lives in the forest and only grows: no allocation per call at a steady size
        fo->dFrames.resize(count);
        if (probs) fo->dProbs.resize(static_cast<size_t>(count) * fo->C * plane);
        if (labels) fo->dLabels.resize(static_cast<size_t>(count) * plane);
        CF_HIP(hipMemcpyAsync(fo->dFrames.get(), indices, count * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        const DevForest df = fo->devForest();
        // with posteriors: launches of at most kPredictChunk frames, each one's results copied to the host on
        // the copy stream while the next launch runs (64 frames of 640x480: 18.8 ms instead of 25.4 ms, the
        // 8 ms copy no longer after the 16.4 ms launch).  Labels alone (1 byte per pixel) are copied after one
        // launch: in 8-frame launches they took 17.8 ms instead of 16.9 ms.
        constexpr int kPredictChunk = 8;
        const int launches = probs ? (count + kPredictChunk - 1) / kPredictChunk : 1;
        const int per = (count + launches - 1) / launches;
        // The copy stream lives for this call only: a stream the process keeps holds one of its hardware queues
        // (four by default), and a serving loop created afterwards then shared queues between its own streams
        // (its 640x480 p99 at 2000 frames/s rose from 2.9 to 7.6 ms, and it served 2270 instead of 2990
        // frames/s at 3000 offered).
        StreamGuard copy;
        hipStream_t cs = ctx->stream;
        if (launches > 1) {
            CF_HIP(hipStreamCreateWithFlags(&copy.s, hipStreamNonBlocking));
            ensureEvents(ctx->launchDone, launches);
            cs = copy.s;
        }
        // Outputs in pageable memory: the runtime copies into such memory through its own staging, one
        // thread at 7-8 GB/s, and the call blocks.  Here each launch's results go by DMA into one half of a
        // page-locked staging buffer, and host threads copy the other half out meanwhile.
        const bool stageP = launches > 1 && !hostPinned(probs);
        const bool stageL = stageP && labels && !hostPinned(labels);
        if (stageP) {
            const size_t chunkP = static_cast<size_t>(per) * fo->C * plane * sizeof(float);
            const size_t chunkL = stageL ? static_cast<size_t>(per) * plane : 0;
            const size_t half = (chunkP + chunkL + 255) & ~static_cast<size_t>(255);
            if (ctx->hostStageBytes < 2 * half) {
                if (ctx->hostStage) CF_HIP(hipHostFree(ctx->hostStage));
                ctx->hostStage = nullptr;
                ctx->hostStageBytes = 0;
                CF_HIP(hipHostMalloc(&ctx->hostStage, 2 * half, hipHostMallocDefault));
                ctx->hostStageBytes = 2 * half;
            }
            ensureEvents(ctx->copyDone, launches);
            for (int k = 0; k < launches; k++) {
                const int f0 = k * per;
                const int n = std::min(per, count - f0);
                const size_t pOff = static_cast<size_t>(f0) * fo->C * plane, lOff = static_cast<size_t>(f0) * plane;
                launchClassify(ctx->stream, im->desc.get(), fo->dFrames.get() + f0, n, im->W, im->H, df,
                        fo->dProbs.get() + pOff, labels ? fo->dLabels.get() + lOff : nullptr);
                CF_HIP(hipEventRecord(ctx->launchDone[k], ctx->stream));
            }
            auto stageOf = [&](int k) { return static_cast<char*>(ctx->hostStage) + (k & 1) * half; };
            auto drain = [&](int k) {   // launch k's results from the staging into the caller's arrays
                const int f0 = k * per;
                const int n = std::min(per, count - f0);
                CF_HIP(hipEventSynchronize(ctx->copyDone[k]));
                parallelCopy(probs + static_cast<size_t>(f0) * fo->C * plane, stageOf(k),
                        static_cast<size_t>(n) * fo->C * plane * sizeof(float));
                if (stageL) parallelCopy(labels + static_cast<size_t>(f0) * plane, stageOf(k) + chunkP, n * plane);
            };
            for (int k = 0; k < launches; k++) {
                const int f0 = k * per;
                const int n = std::min(per, count - f0);
                const size_t pOff = static_cast<size_t>(f0) * fo->C * plane, lOff = static_cast<size_t>(f0) * plane;
                // the half of launch k held launch k - 2's results, drained in the previous round
                CF_HIP(hipStreamWaitEvent(cs, ctx->launchDone[k], 0));
                CF_HIP(hipMemcpyAsync(stageOf(k), fo->dProbs.get() + pOff,
                        static_cast<size_t>(n) * fo->C * plane * sizeof(float), hipMemcpyDeviceToHost, cs));
                if (stageL) {
                    CF_HIP(hipMemcpyAsync(stageOf(k) + chunkP, fo->dLabels.get() + lOff, n * plane,
                            hipMemcpyDeviceToHost, cs));
                }
                CF_HIP(hipEventRecord(ctx->copyDone[k], cs));
                if (labels && !stageL) {
                    CF_HIP(hipMemcpyAsync(labels + lOff, fo->dLabels.get() + lOff, n * plane, hipMemcpyDeviceToHost,
                            cs));
                }
                if (k > 0) drain(k - 1);
            }
            drain(launches - 1);
            CF_HIP(hipStreamSynchronize(cs));
            CF_HIP(hipStreamSynchronize(ctx->stream));
            return;
        }
        for (int k = 0; k < launches; k++) {
            const int f0 = k * per;
            const int n = std::min(per, count - f0);
            if (n <= 0) break;
            const size_t pOff = static_cast<size_t>(f0) * fo->C * plane, lOff = static_cast<size_t>(f0) * plane;
            launchClassify(ctx->stream, im->desc.get(), fo->dFrames.get() + f0, n, im->W, im->H, df,
                    probs ? fo->dProbs.get() + pOff : nullptr, labels ? fo->dLabels.get() + lOff : nullptr);
            if (cs != ctx->stream) {
                CF_HIP(hipEventRecord(ctx->launchDone[k], ctx->stream));
                CF_HIP(hipStreamWaitEvent(cs, ctx->launchDone[k], 0));
            }
            if (probs) {
                CF_HIP(hipMemcpyAsync(probs + pOff, fo->dProbs.get() + pOff,
                        static_cast<size_t>(n) * fo->C * plane * sizeof(float), hipMemcpyDeviceToHost, cs));
            }
            if (labels) {
                CF_HIP(hipMemcpyAsync(labels + lOff, fo->dLabels.get() + lOff, static_cast<size_t>(n) * plane,
                        hipMemcpyDeviceToHost, cs));
            }
        }
        CF_HIP(hipStreamSynchronize(cs));
        if (cs != ctx->stream) CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_predict_batch_device(cf_ctx* ctx, cf_forest* fo, cf_images* im, const int* indices, int count, float* d_probs,
        uint8_t* d_labels) {
    return guarded(ctx, [&] {
        require(fo && im && indices, "null argument");
        require(count > 0, "no frames");
        for (int i = 0; i < count; i++) require(indices[i] >= 0 && indices[i] < im->count, "frame index out of range");
        bindForest(fo, ctx);
        uploadForest(fo);
        fo->dFrames.resize(count);
        CF_HIP(hipMemcpyAsync(fo->dFrames.get(), indices, count * sizeof(int), hipMemcpyHostToDevice, ctx->stream));
        const DevForest df = fo->devForest();
        launchClassify(ctx->stream, im->desc.get(), fo->dFrames.get(), count, im->W, im->H, df, d_probs, d_labels);
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_device_alloc(cf_ctx* ctx, size_t bytes, void** ptr) {
    return guarded(ctx, [&] {
        require(ptr != nullptr && bytes > 0, "invalid allocation");
        CF_HIP(hipSetDevice(ctx->device));
        CF_HIP(hipMalloc(ptr, bytes));
    });
}

int cf_host_alloc(cf_ctx* ctx, size_t bytes, void** ptr) {
    return guarded(ctx, [&] {
        require(ptr != nullptr && bytes > 0, "invalid allocation");
        CF_HIP(hipSetDevice(ctx->device));
        CF_HIP(hipHostMalloc(ptr, bytes, hipHostMallocDefault));
    });
}

int cf_host_free(cf_ctx* ctx, void* ptr) {
    return guarded(ctx, [&] {
        CF_HIP(hipStreamSynchronize(ctx->stream));
        if (ptr) CF_HIP(hipHostFree(ptr));
    });
}

int cf_device_free(cf_ctx* ctx, void* ptr) {
    return guarded(ctx, [&] {
        CF_HIP(hipStreamSynchronize(ctx->stream));
        if (ptr) CF_HIP(hipFree(ptr));
    });
}

int cf_device_to_host(cf_ctx* ctx, void* host, const void* device, size_t bytes) {
    return guarded(ctx, [&] {
        require(host && device, "null argument");
        CF_HIP(hipMemcpyAsync(host, device, bytes, hipMemcpyDeviceToHost, ctx->stream));
        CF_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int cf_predict(cf_ctx* ctx, cf_forest* fo, cf_images* im, int index, float* probs, uint8_t* labels) {
    return cf_predict_batch(ctx, fo, im, &index, 1, probs, labels);
}

}  // extern "C"
