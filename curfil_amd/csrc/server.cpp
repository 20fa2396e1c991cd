// Serving loop for dense prediction: RGB-D frames of any resolution arrive
// one by one (cf_server_submit), are admitted in arrival order while device
// memory of the budget is free, have their integral images built on an
// upload stream, are classified in launches of bounded size (frames of all
// sizes together, large frames in row bands) on a compute stream, and return
// their posteriors and labels through a download stream.
// One worker thread keeps the three stages busy: frames that arrive while a
// batch runs are admitted and integrated meanwhile and go into the next
// batch (continuous batching), and a frame's device slot is reused as soon
// as its results are on the host.  The forest stays resident.
//
// RandomForestImage::predict (random_forest_image.cpp:258-352) per frame: the
// same kernel body as cf_predict, so a frame's results do not depend on what
// else is in flight.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <climits>
#include <cstdlib>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "cf_host.h"

namespace cf {
namespace {

using Clock = std::chrono::steady_clock;

// pinned host buffers, reused by size
class PinnedPool {
public:
    ~PinnedPool() {
        for (auto& kv : free_) {
            for (void* p : kv.second) (void)hipHostFree(p);
        }
    }
    void* take(size_t bytes) {
        {
            std::lock_guard<std::mutex> g(mu_);
            auto& v = free_[bytes];
            if (!v.empty()) {
                void* p = v.back();
                v.pop_back();
                return p;
            }
        }
        void* p = nullptr;
        CF_HIP(hipHostMalloc(&p, bytes, hipHostMallocDefault));
        allocs++;
        return p;
    }
    void give(void* p, size_t bytes) {
        if (!p) return;
        std::lock_guard<std::mutex> g(mu_);
        free_[bytes].push_back(p);
    }
    std::atomic<uint64_t> allocs{0};   // buffers allocated (not taken from the pool)

private:
    std::mutex mu_;
    std::map<size_t, std::vector<void*>> free_;
};

// device memory of one frame of W x H: raw input, integral scratch, integrals, outputs
struct Slot {
    int W = 0, H = 0;
    size_t bytes = 0;
    DevBuf<uint8_t> in;      // rgb [H][W][3], then depth u16 [H][W] at inDepthOffset
    DevBuf<uint32_t> tmp;    // column prefixes of the 5 channels
    DevBuf<float> col;       // [3][H][W]
    DevBuf<int2> dep;        // [H][W]
    DevBuf<float> probs;     // [C][H][W]
    DevBuf<uint8_t> labels;  // [H][W]
};

size_t inputBytes(int W, int H) {
    const size_t px = static_cast<size_t>(W) * H;
    return ((px * 3 + 15) / 16) * 16 + px * 2;
}
size_t depthOffset(int W, int H) { return ((static_cast<size_t>(W) * H * 3 + 15) / 16) * 16; }

// device bytes of one slot: input, integral scratch, integrals, posteriors and labels
size_t slotBytes(int W, int H, int C) {
    const size_t px = static_cast<size_t>(W) * H;
    return inputBytes(W, H) + px * 5 * sizeof(uint32_t) + px * 3 * sizeof(float) + px * sizeof(int2) +
            px * C * sizeof(float) + px;
}

struct Request {
    uint64_t id = 0;
    int W = 0, H = 0;
    void* hostIn = nullptr;      // pinned input
    float* hostProbs = nullptr;  // pinned outputs
    uint8_t* hostLabels = nullptr;
    Slot* slot = nullptr;
    hipEvent_t ready = nullptr, done = nullptr;
    bool finished = false;
    Clock::time_point submitted, admitted, launched, completed;
};

constexpr int kJobRing = 64;   // job tables in flight
// Launch policy, measured with the 10-tree depth-20 forest and Poisson arrivals of 320x240 / 640x480 /
// 1280x960 frames (p99 at 2000 frames/s and frames served at 3000 offered):
// - classify work (pixels) allowed in unfinished launches: below it ready frames launch at once
//   (latency), above it they wait and gather;
// - a launch holds at most kLaunchPixels, so a burst becomes a pipeline of moderate launches instead of
//   one large launch and an idle device while the next gathers (6.9 against 7.7-12.3 ms, 2974 against
//   2490-2540 frames/s; 1.0 Mpix: 6.7 ms but 2894 frames/s; 2.0 Mpix: 8.6 ms);
// - frames above kBandPixels are classified in row bands of at most that many pixels, so that a
//   1280x960 frame does not hold up the frames behind it for a whole launch;
// - frames admitted and not yet launched hold at most kMaxAdmittedPixels; later frames wait in the host
//   queue.  Under overload the admitted frames would otherwise fill the device budget, and every
//   admission of a size without a free slot then gave back other sizes' slots and allocated anew.
// CURFIL_SERVE_PENDING_MPIX, CURFIL_SERVE_LAUNCH_MPIX (0: no limit), CURFIL_SERVE_BAND_MPIX (0: no
// bands) and CURFIL_SERVE_ADMIT_MPIX (0: no limit) override them when a server is created.
constexpr double kMaxPendingPixels = 1.5e6;
constexpr double kLaunchPixels = 1.25e6;
constexpr double kBandPixels = 0.35e6;
constexpr double kMaxAdmittedPixels = 3.0e6;
double envMpix(const char* name, double def) {
    const char* v = getenv(name);
    return v ? atof(v) * 1e6 : def;
}

// a row band of an integrated frame waiting to be classified; the last band's launch completes the frame
struct Band {
    Request* r;
    int rowBegin, rowEnd;
    bool last;
};

}  // namespace
}  // namespace cf

using namespace cf;

struct cf_server {
    cf_ctx* ctx = nullptr;
    cf_forest* fo = nullptr;
    size_t budget = 0;
    int maxBatch = 64;
    double bandPx = 0, launchPx = 0;   // row bands of larger frames (0: none); pixels per launch (0: no limit)
    double pendingMax = kMaxPendingPixels;
    double admitMax = kMaxAdmittedPixels;   // pixels of admitted frames not yet launched (0: no limit)
    double splitPx = 0;                     // launches up to this many pixels split the trees over waves
    // three streams (with the context's own, four: the hardware queues a process gets by default)
    hipStream_t up = nullptr, comp = nullptr, down = nullptr;
    PinnedPool pinned;
    std::mutex mu;
    std::condition_variable cvWork, cvDone;
    std::deque<Request*> queued;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> requests;
    std::map<std::pair<int, int>, std::vector<Slot*>> freeSlots;
    std::vector<std::unique_ptr<Slot>> slots;
    std::vector<hipEvent_t> freeEvents;
    size_t devUsed = 0;
    uint64_t nextId = 1;
    bool stop = false;
    std::string error;
    // job tables: device ring, pinned host staging, and an event per entry marking its kernel done
    DevBuf<FrameJob> dJobs;
    FrameJob* hJobs = nullptr;
    hipEvent_t jobDone[kJobRing] = {};
    int jobNext = 0;
    std::deque<std::pair<int, double>> launchedPx;   // unfinished launches: job entry, pixels
    // statistics; a worker pass is one round of admission, launches and the completion scan, without the
    // waits for work between passes
    uint64_t framesDone = 0, launches = 0;
    double pixelsDone = 0;
    uint64_t passes = 0, slowPasses = 0, slotAllocs = 0, eventAllocs = 0;
    double longestPassMs = 0;
    std::thread worker;

    hipEvent_t takeEvent() {
        if (!freeEvents.empty()) {
            hipEvent_t e = freeEvents.back();
            freeEvents.pop_back();
            return e;
        }
        hipEvent_t e;
        CF_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        eventAllocs++;
        return e;
    }

    // a free slot of W x H, or a new one if the budget allows (caller holds mu)
    Slot* takeSlot(int W, int H) {
        auto& fl = freeSlots[{W, H}];
        if (!fl.empty()) {
            Slot* s = fl.back();
            fl.pop_back();
            return s;
        }
        const size_t px = static_cast<size_t>(W) * H;
        const size_t bytes = slotBytes(W, H, fo->C);
        if (devUsed + bytes > budget) {
            // give back idle slots of other sizes before refusing
            for (auto& kv : freeSlots) {
                while (!kv.second.empty() && devUsed + bytes > budget) {
                    Slot* s = kv.second.back();
                    kv.second.pop_back();
                    devUsed -= s->bytes;
                    for (auto it = slots.begin(); it != slots.end(); ++it) {
                        if (it->get() == s) {
                            slots.erase(it);
                            break;
                        }
                    }
                }
            }
            if (devUsed + bytes > budget) return nullptr;
        }
        auto s = std::make_unique<Slot>();
        s->W = W;
        s->H = H;
        s->bytes = bytes;
        s->in.resize(inputBytes(W, H));
        s->tmp.resize(px * 5);
        s->col.resize(px * 3);
        s->dep.resize(px);
        s->probs.resize(px * fo->C);
        s->labels.resize(px);
        devUsed += bytes;
        slotAllocs++;
        slots.push_back(std::move(s));
        return slots.back().get();
    }

    void run();
};

void cf_server::run() {
    try {
        CF_HIP(hipSetDevice(ctx->device));
        std::vector<Band> ready;         // bands of frames integrating or integrated, not yet classified
        std::deque<Request*> inflight;   // classified, results on their way to the host, in copy order
        const DevForest df = fo->devForest();
        double passMs = -1;        // the previous pass
        double admittedPx = 0;     // pixels of admitted frames whose last band has not been launched
        for (;;) {
            std::vector<Request*> admit;
            Clock::time_point passStart;
            {
                std::unique_lock<std::mutex> lk(mu);
                if (passMs >= 0) {
                    passes++;
                    if (passMs > 1.0) slowPasses++;
                    longestPassMs = std::max(longestPassMs, passMs);
                }
                if (queued.empty() && ready.empty() && inflight.empty()) {
                    cvWork.wait(lk, [&] { return stop || !queued.empty(); });
                } else if (!inflight.empty()) {
                    cvWork.wait_for(lk, std::chrono::microseconds(20));
                }
                passStart = Clock::now();
                if (stop && queued.empty() && ready.empty() && inflight.empty()) break;
                while (!queued.empty()) {
                    Request* r = queued.front();
                    const double rpx = static_cast<double>(r->W) * r->H;
                    if (admitMax > 0 && admittedPx > 0 && admittedPx + rpx > admitMax) break;
                    Slot* s = takeSlot(r->W, r->H);
                    if (!s) break;   // arrival order: later frames wait for this one
                    queued.pop_front();
                    r->slot = s;
                    r->admitted = Clock::now();
                    r->ready = takeEvent();
                    r->done = takeEvent();
                    admittedPx += rpx;
                    admit.push_back(r);
                }
            }
            // pinned result buffers are taken at admission, so the device budget also bounds them
            for (Request* r : admit) {
                const size_t px = static_cast<size_t>(r->W) * r->H;
                r->hostProbs = static_cast<float*>(pinned.take(px * fo->C * sizeof(float)));
                r->hostLabels = static_cast<uint8_t*>(pinned.take(px));
            }
            // upload and integrate the admitted frames
            for (Request* r : admit) {
                Slot* s = r->slot;
                CF_HIP(hipMemcpyAsync(s->in.get(), r->hostIn, inputBytes(r->W, r->H), hipMemcpyHostToDevice, up));
                launchIntegralU8(up, s->in.get(), reinterpret_cast<const uint16_t*>(s->in.get() + depthOffset(r->W, r->H)),
                        r->W, r->H, s->tmp.get(), s->col.get(), s->dep.get());
                CF_HIP(hipEventRecord(r->ready, up));
                // frames above bandPx pixels in bands of whole 16-row tile rows, so that other frames can
                // be classified between them
                int rows = r->H;
                if (bandPx > 0 && static_cast<double>(r->W) * r->H > bandPx) {
                    rows = std::max(16, static_cast<int>(bandPx / r->W) / 16 * 16);
                }
                for (int y0 = 0; y0 < r->H; y0 += rows) {
                    const int y1 = std::min(r->H, y0 + rows);
                    ready.push_back(Band{r, y0, y1, y1 == r->H});
                }
            }
            // classify what is ready, bands of frames of all sizes together in arrival order (at most maxBatch
            // bands and launchPx pixels per launch), while the unfinished launches hold less than pendingMax
            // pixels of work: under load, ready frames gather into larger launches
            double pendingPx = 0;
            while (!launchedPx.empty()) {
                const hipError_t q = hipEventQuery(jobDone[launchedPx.front().first]);
                if (q == hipErrorNotReady) break;
                CF_HIP(q);
                launchedPx.pop_front();
            }
            for (const auto& lp : launchedPx) pendingPx += lp.second;
            size_t next = 0;
            while (next < ready.size() && pendingPx < pendingMax) {
                // the next launch: bands in arrival order, up to maxBatch of them and launchPx pixels
                size_t end = next;
                double px = 0;
                while (end < ready.size() && end - next < static_cast<size_t>(maxBatch)) {
                    const Band& bd = ready[end];
                    const double bpx = static_cast<double>(bd.r->W) * (bd.rowEnd - bd.rowBegin);
                    if (end > next && launchPx > 0 && px + bpx > launchPx) break;
                    px += bpx;
                    end++;
                }
                const int n = static_cast<int>(end - next);
                const int e = jobNext;
                jobNext = (jobNext + 1) % kJobRing;
                CF_HIP(hipEventSynchronize(jobDone[e]));   // the entry's previous kernel has run
                hipStream_t cs = comp;
                FrameJob* hj = hJobs + static_cast<size_t>(e) * maxBatch;
                for (int i = 0; i < n; i++) {
                    const Band& bd = ready[next + i];
                    Request* r = bd.r;
                    Slot* s = r->slot;
                    hj[i].im = DevImage{s->col.get(), s->dep.get(), r->W, r->H, 0, 0};
                    hj[i].probs = s->probs.get();
                    hj[i].labels = s->labels.get();
                    hj[i].rowBegin = bd.rowBegin;
                    hj[i].rowEnd = bd.rowEnd;
                    hj[i].pad = 0;
                    if (bd.rowBegin == 0) CF_HIP(hipStreamWaitEvent(cs, r->ready, 0));
                }
                const ClassifyPlan plan = planClassifyJobs(hj, n, df, splitPx);
                FrameJob* dj = dJobs.get() + static_cast<size_t>(e) * maxBatch;
                CF_HIP(hipMemcpyAsync(dj, hj, n * sizeof(FrameJob), hipMemcpyHostToDevice, cs));
                launchClassifyJobs(cs, dj, n, plan, df);
                CF_HIP(hipEventRecord(jobDone[e], cs));
                launchedPx.emplace_back(e, px);
                pendingPx += px;
                launches++;
                // results of the frames whose last band this launch holds (the earlier bands ran in earlier
                // launches on the same stream), smallest first: a large frame's copy does not hold up the small
                // ones finished with it
                std::vector<Request*> order;
                for (int i = 0; i < n; i++) {
                    if (ready[next + i].last) {
                        order.push_back(ready[next + i].r);
                        admittedPx -= static_cast<double>(ready[next + i].r->W) * ready[next + i].r->H;
                    }
                }
                if (!order.empty()) CF_HIP(hipStreamWaitEvent(down, jobDone[e], 0));
                std::stable_sort(order.begin(), order.end(), [](const Request* x, const Request* y) {
                    return static_cast<int64_t>(x->W) * x->H < static_cast<int64_t>(y->W) * y->H;
                });
                for (Request* r : order) {
                    Slot* s = r->slot;
                    const size_t rpx = static_cast<size_t>(r->W) * r->H;
                    CF_HIP(hipMemcpyAsync(r->hostProbs, s->probs.get(), rpx * fo->C * sizeof(float),
                            hipMemcpyDeviceToHost, down));
                    CF_HIP(hipMemcpyAsync(r->hostLabels, s->labels.get(), rpx, hipMemcpyDeviceToHost, down));
                    CF_HIP(hipEventRecord(r->done, down));
                    r->launched = Clock::now();
                    inflight.push_back(r);
                }
                next = end;
            }
            ready.erase(ready.begin(), ready.begin() + next);
            // finished frames: results are in pinned host memory; their slots are free again.  The copies run
            // in order on one stream, so the first unfinished frame ends the scan (one query per loop pass
            // however many frames are in flight)
            bool any = false;
            while (!inflight.empty()) {
                Request* r = inflight.front();
                const hipError_t q = hipEventQuery(r->done);
                if (q == hipErrorNotReady) break;
                CF_HIP(q);
                inflight.pop_front();
                std::lock_guard<std::mutex> g(mu);
                r->completed = Clock::now();
                r->finished = true;
                freeSlots[{r->W, r->H}].push_back(r->slot);
                r->slot = nullptr;
                freeEvents.push_back(r->ready);
                freeEvents.push_back(r->done);
                r->ready = r->done = nullptr;
                framesDone++;
                pixelsDone += static_cast<double>(r->W) * r->H;
                any = true;
            }
            if (any) cvDone.notify_all();
            passMs = std::chrono::duration<double, std::milli>(Clock::now() - passStart).count();
        }
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> g(mu);
        error = e.what();
        stop = true;
        cvDone.notify_all();
    }
}

extern "C" {

int cf_server_create(cf_ctx* ctx, cf_forest* fo, double budget_mb, int max_batch, cf_server** out) {
    if (!ctx) return CF_ERR_ARG;
    std::lock_guard<std::recursive_mutex> lock(ctx->mu);
    try {
        if (!fo || !out) throw Error(CF_ERR_ARG, "null argument");
        if (!(budget_mb > 0) || max_batch <= 0) throw Error(CF_ERR_ARG, "budget and batch size must be positive");
        if (fo->post.size() != fo->trees.size()) throw Error(CF_ERR_RUNTIME, "histograms not normalized");
        CF_HIP(hipSetDevice(ctx->device));
        auto sv = std::make_unique<cf_server>();
        sv->ctx = ctx;
        sv->fo = fo;
        sv->budget = static_cast<size_t>(budget_mb * 1024.0 * 1024.0);
        sv->maxBatch = max_batch;
        sv->bandPx = envMpix("CURFIL_SERVE_BAND_MPIX", kBandPixels);
        sv->launchPx = envMpix("CURFIL_SERVE_LAUNCH_MPIX", kLaunchPixels);
        sv->pendingMax = envMpix("CURFIL_SERVE_PENDING_MPIX", kMaxPendingPixels);
        sv->admitMax = envMpix("CURFIL_SERVE_ADMIT_MPIX", kMaxAdmittedPixels);
        sv->splitPx = classifySplitPixels();
        bindForest(fo, ctx);
        uploadForest(fo);
        CF_HIP(hipStreamCreateWithFlags(&sv->up, hipStreamNonBlocking));
        CF_HIP(hipStreamCreateWithFlags(&sv->comp, hipStreamNonBlocking));
        CF_HIP(hipStreamCreateWithFlags(&sv->down, hipStreamNonBlocking));
        sv->dJobs.resize(static_cast<size_t>(kJobRing) * max_batch);
        CF_HIP(hipHostMalloc(reinterpret_cast<void**>(&sv->hJobs), sizeof(FrameJob) * kJobRing * max_batch,
                hipHostMallocDefault));
        for (int i = 0; i < kJobRing; i++) {
            CF_HIP(hipEventCreateWithFlags(&sv->jobDone[i], hipEventDisableTiming));
            CF_HIP(hipEventRecord(sv->jobDone[i], sv->comp));
        }
        cf_server* raw = sv.get();
        sv->worker = std::thread([raw] { raw->run(); });
        *out = sv.release();
        return CF_OK;
    } catch (const Error& e) {
        ctx->error = e.what();
        return e.code;
    } catch (const std::exception& e) {
        ctx->error = e.what();
        return CF_ERR_RUNTIME;
    }
}

int cf_server_submit(cf_server* sv, const uint8_t* rgb, const uint16_t* depth, int width, int height,
        uint64_t* id) {
    // a frame's latency runs from this call, its checks and the copy of its input included
    const Clock::time_point called = Clock::now();
    if (!sv || !rgb || !depth || !id || width <= 0 || height <= 0 || width > 32767 || height > 32767) {
        return CF_ERR_ARG;
    }
    // an integral plane is addressed with 32-bit byte offsets
    if (static_cast<uint64_t>(width) * height * sizeof(int2) >= (1ull << 32)) return CF_ERR_ARG;
    // depth above 50 m is refused, as the reference's depth loader does (image.cpp:430)
    for (size_t p = 0, px = static_cast<size_t>(width) * height; p < px; p++) {
        if (depth[p] > 50000) return CF_ERR_ARG;
    }
    // a frame that does not fit the device budget on its own could never be admitted
    if (slotBytes(width, height, sv->fo->C) > sv->budget) return CF_ERR_NOMEM;
    try {
        auto r = std::make_unique<Request>();
        r->W = width;
        r->H = height;
        const size_t px = static_cast<size_t>(width) * height;
        r->hostIn = sv->pinned.take(inputBytes(width, height));
        std::memcpy(r->hostIn, rgb, px * 3);
        std::memcpy(static_cast<uint8_t*>(r->hostIn) + depthOffset(width, height), depth, px * 2);
        std::lock_guard<std::mutex> g(sv->mu);
        if (sv->stop) {
            sv->pinned.give(r->hostIn, inputBytes(width, height));
            return CF_ERR_RUNTIME;
        }
        r->submitted = called;
        r->id = sv->nextId++;
        *id = r->id;
        sv->queued.push_back(r.get());
        sv->requests[r->id] = std::move(r);
        sv->cvWork.notify_one();
        return CF_OK;
    } catch (const Error& e) {
        return e.code;
    } catch (const std::exception&) {
        return CF_ERR_RUNTIME;
    }
}

int cf_server_wait(cf_server* sv, uint64_t id, float* probs, uint8_t* labels, double* latency_ms) {
    if (!sv) return CF_ERR_ARG;
    std::unique_ptr<Request> r;
    {
        std::unique_lock<std::mutex> lk(sv->mu);
        auto it = sv->requests.find(id);
        if (it == sv->requests.end()) return CF_ERR_ARG;
        Request* q = it->second.get();
        sv->cvDone.wait(lk, [&] { return q->finished || !sv->error.empty(); });
        if (!q->finished) return CF_ERR_RUNTIME;
        r = std::move(it->second);
        sv->requests.erase(it);
    }
    const size_t px = static_cast<size_t>(r->W) * r->H;
    if (probs) std::memcpy(probs, r->hostProbs, px * sv->fo->C * sizeof(float));
    if (labels) std::memcpy(labels, r->hostLabels, px);
    if (latency_ms) {
        using ms = std::chrono::duration<double, std::milli>;
        latency_ms[0] = ms(r->completed - r->submitted).count();
        latency_ms[1] = ms(r->admitted - r->submitted).count();
        latency_ms[2] = ms(r->launched - r->admitted).count();
        latency_ms[3] = ms(r->completed - r->launched).count();
    }
    sv->pinned.give(r->hostIn, inputBytes(r->W, r->H));
    sv->pinned.give(r->hostProbs, px * sv->fo->C * sizeof(float));
    sv->pinned.give(r->hostLabels, px);
    return CF_OK;
}

int cf_server_stats(cf_server* sv, double* out, int n) {
    if (!sv || !out) return CF_ERR_ARG;
    std::lock_guard<std::mutex> g(sv->mu);
    const double v[] = {static_cast<double>(sv->framesDone), sv->pixelsDone, static_cast<double>(sv->launches),
            static_cast<double>(sv->devUsed) / (1024.0 * 1024.0), static_cast<double>(sv->queued.size()),
            static_cast<double>(sv->passes), sv->longestPassMs, static_cast<double>(sv->slowPasses),
            static_cast<double>(sv->slotAllocs), static_cast<double>(sv->pinned.allocs.load()),
            static_cast<double>(sv->eventAllocs)};
    const int k = static_cast<int>(sizeof(v) / sizeof(v[0]));
    for (int i = 0; i < n && i < k; i++) out[i] = v[i];
    return k;
}

const char* cf_server_error(cf_server* sv) {
    if (!sv) return "null server";
    std::lock_guard<std::mutex> g(sv->mu);
    return sv->error.c_str();
}

int cf_server_destroy(cf_server* sv) {
    if (!sv) return CF_ERR_ARG;
    {
        std::lock_guard<std::mutex> g(sv->mu);
        sv->stop = true;
        sv->cvWork.notify_all();
    }
    if (sv->worker.joinable()) sv->worker.join();
    (void)hipSetDevice(sv->ctx->device);
    (void)hipStreamSynchronize(sv->up);
    (void)hipStreamSynchronize(sv->comp);
    (void)hipStreamSynchronize(sv->down);
    for (auto& kv : sv->requests) {
        Request* r = kv.second.get();
        const size_t px = static_cast<size_t>(r->W) * r->H;
        sv->pinned.give(r->hostIn, inputBytes(r->W, r->H));
        sv->pinned.give(r->hostProbs, px * sv->fo->C * sizeof(float));
        sv->pinned.give(r->hostLabels, px);
        if (r->ready) (void)hipEventDestroy(r->ready);
        if (r->done) (void)hipEventDestroy(r->done);
    }
    for (hipEvent_t e : sv->freeEvents) (void)hipEventDestroy(e);
    for (int i = 0; i < kJobRing; i++) {
        if (sv->jobDone[i]) (void)hipEventDestroy(sv->jobDone[i]);
    }
    if (sv->hJobs) (void)hipHostFree(sv->hJobs);
    (void)hipStreamDestroy(sv->up);
    (void)hipStreamDestroy(sv->comp);
    (void)hipStreamDestroy(sv->down);
    delete sv;
    return CF_OK;
}

}  // extern "C"
