// Level-wise training driver (host side of the hot path).
//
// Semantics follow RandomTreeTrain::train (random_tree.h:1249-1415) and
// ImageFeatureEvaluation::evaluateBestSplits (random_tree_image.cpp:239-502):
//   * the root is level 1; a level equal to maxDepth is not evaluated;
//   * one seed per (tree, level) from the tree's RandomSource, one feature
//     table per (tree, level) generated from all of the level's samples in
//     frontier order;
//   * children ids are ++idNode (left, right), rolled back by 2 when a side is
//     empty (the node stays a leaf);
//   * a child keeps growing iff its histogram is not pure and it holds at
//     least minSampleCount samples;
//   * with horizontal flipping, samples set to Both count twice in the
//     histograms and are evaluated in both orientations; a Both sample that the
//     split separates from its flipped self continues as NoFlip plus a Flip
//     copy on the other side (random_tree.h:1295-1360).
// The device keeps the level's samples contiguous per node; evaluation runs
// either node by node (CF_MODE_PLAIN, the reference GPU path's shape: one
// launch sequence and one host sync per node) or for every node of every
// tree of the level together (CF_MODE_BATCHED).
#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>

#include "cf_host.h"

namespace cf {

namespace {

constexpr int kChunk = 4096;

struct ActiveTree {
    int job;                  // index into jobs / out
    int idNode;               // last assigned node id
    std::vector<int> nodes;   // frontier: node indices into out[job].nodes
    std::vector<int> counts;  // samples per frontier node
    int begin = 0, n = 0;     // segment of the level buffer
};

// The pass-1 blocks of one block size at the current level, and the work list of a batch size.
struct BlockSet {
    int level = -1;        // level the blocks were built for
    DevBuf<BinBlock> blocks;
    DevBuf<int> segInfo;   // per image group: first block, block count, XCD prefix; then the XCD totals
    int numSegs = 0, maxXcdBlocks = 0;
    DevBuf<int2> work;
    int workGroups = -1;   // feature groups the work list was built for
    int numWork = 0;
};

// An event pair for timing launches the tuner measures.
class EventPair {
public:
    EventPair() {
        CF_HIP(hipEventCreate(&a));
        CF_HIP(hipEventCreate(&b));
    }
    ~EventPair() {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
    }
    EventPair(const EventPair&) = delete;
    EventPair& operator=(const EventPair&) = delete;
    hipEvent_t a = nullptr, b = nullptr;
};

// A level's evaluation captured from the stream into a graph, instantiated and launched once.  The
// capture holds kernels only (buffers are cleared by a fill kernel, not a memset node; tables are
// uploaded before it), so nothing in it depends on host memory.
class LevelGraph {
public:
    explicit LevelGraph(hipStream_t s) : s_(s) {}
    ~LevelGraph() {
        if (capturing_) {
            hipGraph_t g = nullptr;
            (void)hipStreamEndCapture(s_, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        if (launched_) (void)hipStreamSynchronize(s_);   // normally done already: the level waits for its argmax
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
    }
    LevelGraph(const LevelGraph&) = delete;
    LevelGraph& operator=(const LevelGraph&) = delete;
    void beginCapture() {
        CF_HIP(hipStreamBeginCapture(s_, hipStreamCaptureModeThreadLocal));
        capturing_ = true;
    }
    void launch() {
        capturing_ = false;
        CF_HIP(hipStreamEndCapture(s_, &graph_));
        CF_HIP(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
        CF_HIP(hipGraphLaunch(exec_, s_));
        launched_ = true;
    }
    bool launched() const { return launched_; }

private:
    hipStream_t s_;
    bool capturing_ = false, launched_ = false;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t exec_ = nullptr;
};

struct LevelNode {
    int active;  // index into the active list (feature table)
    int node;    // node index in its tree
    int begin, count;
    int chunkBegin, chunkEnd;
};

// Device time of the work enqueued between begin() and end() on the stream.  The events are read by
// flush(), called where the host waits for the stream anyway, so timing adds no host waits between
// the kernels of a level.
class PhaseTimer {
public:
    PhaseTimer(hipStream_t s, bool on) : s_(s), on_(on) {}
    ~PhaseTimer() {
        for (hipEvent_t e : pool_) (void)hipEventDestroy(e);
        for (const Pending& p : pending_) {
            (void)hipEventDestroy(p.a);
            (void)hipEventDestroy(p.b);
        }
        if (cur_) (void)hipEventDestroy(cur_);
    }
    void begin() {
        if (!on_) return;
        if (!cur_) cur_ = take();
        CF_HIP(hipEventRecord(cur_, s_));
    }
    // the phase's elapsed milliseconds are added to `acc` at the next flush()
    void end(double& acc) {
        if (!on_) return;
        hipEvent_t b = take();
        CF_HIP(hipEventRecord(b, s_));
        pending_.push_back(Pending{cur_, b, &acc});
        cur_ = nullptr;
    }
    // waits for the last ended phase and accumulates every ended phase
    void flush() {
        if (!on_ || pending_.empty()) return;
        CF_HIP(hipEventSynchronize(pending_.back().b));
        for (const Pending& p : pending_) {
            float ms = 0;
            CF_HIP(hipEventElapsedTime(&ms, p.a, p.b));
            *p.acc += ms;
            pool_.push_back(p.a);
            pool_.push_back(p.b);
        }
        pending_.clear();
    }

private:
    struct Pending {
        hipEvent_t a, b;
        double* acc;
    };
    hipEvent_t take() {
        if (!pool_.empty()) {
            hipEvent_t e = pool_.back();
            pool_.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        CF_HIP(hipEventCreate(&e));
        return e;
    }
    hipStream_t s_;
    bool on_;
    hipEvent_t cur_ = nullptr;
    std::vector<hipEvent_t> pool_;
    std::vector<Pending> pending_;
};

// Device memory budgets of a batched level.  CURFIL_*_BUDGET_MB sets them in MiB
// (at least 16); CURFIL_*_BUDGET_KB in KiB without that floor, so that small
// workloads can be made to run the evaluation in several feature-group and
// node-slot batches, as the full-size workloads do.
size_t budgetBytes(const char* mbVar, const char* kbVar, size_t defaultMb) {
    if (const char* kb = std::getenv(kbVar)) {
        const long long v = std::atoll(kb);
        return static_cast<size_t>(std::max(1LL, v)) << 10;
    }
    const char* e = std::getenv(mbVar);
    size_t mb = e ? static_cast<size_t>(std::atoll(e)) : defaultMb;
    if (mb < 16) mb = 16;
    return mb << 20;
}

size_t evalBudgetBytes() { return budgetBytes("CURFIL_EVAL_BUDGET_MB", "CURFIL_EVAL_BUDGET_KB", 2048); }

size_t binsBudgetBytes() { return budgetBytes("CURFIL_BINS_BUDGET_MB", "CURFIL_BINS_BUDGET_KB", 4096); }

bool isPure(const uint32_t* h, int C) {
    int nz = 0;
    for (int c = 0; c < C; c++) {
        if (h[c] == 0) continue;
        if (++nz > 1) return false;
    }
    return nz == 1;
}

}  // namespace

void trainTrees(cf_ctx* ctx, cf_images* images, const cf_train_config& cfg, int C, std::vector<TreeJob>& jobs,
        int mode, std::vector<TreeBuild>& out, TrainStats& stats) {
    if (cfg.max_depth <= 0) throw Error(CF_ERR_ARG, "maxDepth must be positive");
    if (C < 1 || C > kMaxClasses) throw Error(CF_ERR_ARG, "number of classes must be in [1, 32]");
    if (cfg.feature_count == 0 || cfg.thresholds == 0) throw Error(CF_ERR_ARG, "featureCount and thresholds must be positive");
    if (mode != CF_MODE_PLAIN && mode != CF_MODE_BATCHED) throw Error(CF_ERR_ARG, "unknown training mode");
    const int F = static_cast<int>(cfg.feature_count);
    const int T = cfg.thresholds;
    hipStream_t s = ctx->stream;
    const auto t0 = std::chrono::steady_clock::now();

    out.assign(jobs.size(), TreeBuild());
    size_t total = 0;
    bool anyBoth = false;
    for (auto& j : jobs) {
        total += j.roots->size();
        for (const cf_sample& x : *j.roots) anyBoth = anyBoth || x.flip == 2;
    }
    // a Both sample splits at most once into a NoFlip sample and a Flip copy, which
    // never split again: a level holds at most twice the root samples
    const size_t capacity = std::max<size_t>(anyBoth ? 2 * total : total, 1);
    DevBuf<cf_sample> level[2];
    level[0].resize(capacity);
    level[1].resize(capacity);
    DevBuf<uint8_t> dir(capacity);

    std::vector<ActiveTree> active;
    {
        size_t offset = 0;   // each tree's roots are copied straight from the forest into level 0
        for (size_t j = 0; j < jobs.size(); j++) {
            TreeBuild& tb = out[j];
            cf_tree_node root{};
            root.id = jobs[j].treeId;
            root.level = 1;
            root.parent = -1;
            root.left = root.right = -1;
            root.num_samples = static_cast<int32_t>(jobs[j].roots->size());
            root.is_leaf = 1;
            tb.nodes.push_back(root);
            tb.hist.assign(C, 0u);
            for (const cf_sample& x : *jobs[j].roots) {
                if (x.label >= C) throw Error(CF_ERR_ARG, "sample label >= number of classes");
                if (x.image >= static_cast<uint32_t>(images->count)) throw Error(CF_ERR_ARG, "sample image out of range");
                if (x.x >= images->W || x.y >= images->H) throw Error(CF_ERR_ARG, "sample position outside the image");
                tb.hist[x.label] += 1;
                if (x.flip == 2) tb.hist[x.label] += 1;
            }
            if (jobs[j].roots->empty()) continue;
            ActiveTree a;
            a.job = static_cast<int>(j);
            a.idNode = jobs[j].treeId;
            a.nodes.push_back(0);
            a.counts.push_back(static_cast<int>(jobs[j].roots->size()));
            a.begin = static_cast<int>(offset);
            a.n = static_cast<int>(jobs[j].roots->size());
            CF_HIP(hipMemcpyAsync(level[0].get() + offset, jobs[j].roots->data(), jobs[j].roots->size() * sizeof(cf_sample),
                    hipMemcpyHostToDevice, s));
            offset += jobs[j].roots->size();
            active.push_back(std::move(a));
        }
    }
    int cur = 0;

    // Batched levels use the two-pass evaluation (image-major bins, then
    // node histograms; samples set to Both get a second, flipped bin row)
    // unless T exceeds the 8-bit bins or CURFIL_EVAL_DIRECT is set.
    const bool twoPass = mode == CF_MODE_BATCHED && T <= kMaxBinThresholds &&
            std::getenv("CURFIL_EVAL_DIRECT") == nullptr;

    FeatTables tabs;
    DevBuf<FeatGenJob> dJobs;
    DevBuf<Chunk> dChunks;
    DevBuf<int2> dGraphWork[2];   // work lists of a captured level: full and last feature-group batch
    DevBuf<uint32_t> dNodeHist;
    DevBuf<int> dSlotTree;
    DevBuf<uint32_t> dG;
    DevBuf<FeatSummary> dSumm;
    DevBuf<NodeBest> dBest;
    DevBuf<double> dAllScores;
    DevBuf<SplitRec> dSplits;
    DevBuf<NodeBest> dBestAll;
    DevBuf<int> dNodeTree;
    DevBuf<int2> dChunkCounts;
    DevBuf<uint32_t> dLeftHist;
    DevBuf<int2> dDst;
    DevBuf<uint32_t> dKeys, dVals, dPerm, dSortCount, dSortCursor, dSortBlocks;
    DevBuf<uint8_t> dLab8;
    std::map<int, BlockSet> blockSets;   // by pass-1 block size
    EventPair tuneEvents;
    const hipEvent_t tuneStart = tuneEvents.a, tuneEnd = tuneEvents.b;
    DevBuf<uint32_t> dRepHist;
    DevBuf<int> dRepTree;
    DevBuf<FeatSummary> dRepSumm;
    DevBuf<NodeBest> dRepBest;
    DevBuf<FeatRec> dGFeat;
    DevBuf<float> dGThr;
    DevBuf<int> dStarts;
    DevBuf<uint16_t> dEvalOrder;
    DevBuf<uint4> dBins;
    const size_t binsBudget = binsBudgetBytes();
    // CURFIL_LEVEL_GRAPH=1: levels whose pass-1 shape is settled run as one captured graph.  Off by default:
    // once a graph has been instantiated the HIP runtime holds one of the process's hardware queues
    // (GPU_MAX_HW_QUEUES, 4 by default), so streams created later share queues (the serving loop's three
    // lost a fifth of their capacity), and a level launched kernel by kernel is as fast: its host reads
    // wait for nothing between kernels.  Not while the context profiles either: the phase timers' events
    // have to be recorded on the stream (events recorded inside a captured graph give no elapsed times).
    const char* lg = std::getenv("CURFIL_LEVEL_GRAPH");
    const bool levelGraphs = lg && std::strcmp(lg, "1") == 0 && !ctx->profiling;
    PhaseTimer timer(s, ctx->profiling);
    const size_t perSlot = static_cast<size_t>(F) * (T + 1) * C;
    const size_t budget = evalBudgetBytes();
    size_t maxSlots = std::max<size_t>(1, budget / (perSlot * sizeof(uint32_t)));

    // per-level evaluation log on stderr (with profiling on), like the reference's per-level INFO lines
    const bool levelLog = ctx->profiling && std::getenv("CURFIL_LEVEL_LOG") != nullptr;
    for (int levelNr = 1; !active.empty(); levelNr++) {
        if (levelNr == cfg.max_depth) break;
        stats.levels++;
        const int numActive = static_cast<int>(active.size());
        // ---- feature generation, one job per tree
        std::vector<int> seeds(numActive);
        std::vector<FeatGenJob> fj(numActive);
        for (int a = 0; a < numActive; a++) {
            seeds[a] = jobs[active[a].job].rs.uniformSampler(0xFFFFFF).next();
            fj[a].samples = level[cur].get() + active[a].begin;
            fj[a].n = active[a].n;
            fj[a].seed = seeds[a];
            fj[a].outOffset = a * F;
            fj[a].pad = 0;
        }
        dJobs.resize(numActive);
        CF_HIP(hipMemcpyAsync(dJobs.get(), fj.data(), numActive * sizeof(FeatGenJob), hipMemcpyHostToDevice, s));
        timer.begin();
        launchFeatureGeneration(s, images->set(), dJobs.get(), numActive, F, T, cfg.box_radius, cfg.region_size,
                cfg.use_depth_images, seeds.data(), tabs);
        timer.end(stats.featGenMs);

        // ---- nodes and chunks of the level (slot = position in `nodes`)
        std::vector<LevelNode> nodes;
        std::vector<Chunk> chunks;
        std::vector<int> chunkSlotTree;
        for (int a = 0; a < numActive; a++) {
            int off = active[a].begin;
            for (size_t k = 0; k < active[a].nodes.size(); k++) {
                LevelNode ln;
                ln.active = a;
                ln.node = active[a].nodes[k];
                ln.begin = off;
                ln.count = active[a].counts[k];
                ln.chunkBegin = static_cast<int>(chunks.size());
                for (int b = 0; b < ln.count; b += kChunk) {
                    Chunk c;
                    c.slot = static_cast<int>(nodes.size());
                    c.tree = a;
                    c.begin = off + b;
                    c.count = std::min(kChunk, ln.count - b);
                    chunks.push_back(c);
                }
                ln.chunkEnd = static_cast<int>(chunks.size());
                off += ln.count;
                nodes.push_back(ln);
                stats.evaluations += static_cast<double>(ln.count) * F * T;
                stats.responses += static_cast<double>(ln.count) * F;
                stats.nodes += 1;
            }
        }
        const int numNodes = static_cast<int>(nodes.size());
        std::vector<NodeBest> best(numNodes);
        const double evalMsBefore = stats.evalKernelMs;

        // ---- evaluation
        auto evalSlots = [&](int s0, int s1) {
            const int ns = s1 - s0;
            const int c0 = nodes[s0].chunkBegin, c1 = nodes[s1 - 1].chunkEnd;
            std::vector<Chunk> bc(chunks.begin() + c0, chunks.begin() + c1);
            for (auto& c : bc) c.slot -= s0;
            std::vector<uint32_t> nh(static_cast<size_t>(ns) * C);
            std::vector<int> st(ns);
            for (int i = 0; i < ns; i++) {
                const LevelNode& ln = nodes[s0 + i];
                std::memcpy(&nh[static_cast<size_t>(i) * C], &out[active[ln.active].job].hist[static_cast<size_t>(ln.node) * C],
                        C * sizeof(uint32_t));
                st[i] = ln.active;
            }
            dChunks.resize(bc.size());
            dNodeHist.resize(nh.size());
            dSlotTree.resize(ns);
            dG.resize(static_cast<size_t>(ns) * perSlot);
            dSumm.resize(static_cast<size_t>(ns) * F);
            dBest.resize(ns);
            CF_HIP(hipMemcpyAsync(dChunks.get(), bc.data(), bc.size() * sizeof(Chunk), hipMemcpyHostToDevice, s));
            CF_HIP(hipMemcpyAsync(dNodeHist.get(), nh.data(), nh.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
            CF_HIP(hipMemcpyAsync(dSlotTree.get(), st.data(), ns * sizeof(int), hipMemcpyHostToDevice, s));
            CF_HIP(hipMemsetAsync(dG.get(), 0, static_cast<size_t>(ns) * perSlot * sizeof(uint32_t), s));
            timer.begin();
            launchEvalHistograms(s, images->set(), level[cur].get(), dChunks.get(), static_cast<int>(bc.size()),
                    tabs, C, dG.get());
            timer.end(stats.evalKernelMs);
            stats.evalLaunches += 1;
            timer.begin();
            launchScores(s, dG.get(), dNodeHist.get(), dSlotTree.get(), ns, tabs, C, dSumm.get(), nullptr, -1);
            launchReduce(s, dSumm.get(), ns, F, dBest.get());
            timer.end(stats.scoreMs);
            CF_HIP(hipMemcpyAsync(&best[s0], dBest.get(), ns * sizeof(NodeBest), hipMemcpyDeviceToHost, s));
            CF_HIP(hipStreamSynchronize(s));
            timer.flush();
            for (int i = 0; i < ns; i++) {
                if (!best[s0 + i].flag) continue;
                // near-tie band: replay the reference's sequential scan for this node
                dAllScores.resize(static_cast<size_t>(T) * F);
                launchScores(s, dG.get(), dNodeHist.get(), dSlotTree.get(), ns, tabs, C, dSumm.get(), dAllScores.get(), i);
                launchSequentialArgmax(s, dAllScores.get(), T, F, dBest.get() + i);
                CF_HIP(hipMemcpyAsync(&best[s0 + i], dBest.get() + i, sizeof(NodeBest), hipMemcpyDeviceToHost, s));
                CF_HIP(hipStreamSynchronize(s));
                timer.flush();
                stats.replays++;
            }
        };
        // the literal sequential argmax of flagged nodes from their direct-path histograms, as many nodes
        // per launch as the histogram budget holds
        auto replayNodes = [&](const std::vector<int>& flagged) {
            for (size_t a = 0; a < flagged.size();) {
                const int ns = static_cast<int>(std::min<size_t>(flagged.size() - a, maxSlots));
                std::vector<Chunk> bc;
                std::vector<uint32_t> nh(static_cast<size_t>(ns) * C);
                std::vector<int> st(ns);
                for (int k = 0; k < ns; k++) {
                    const LevelNode& ln = nodes[flagged[a + k]];
                    for (int c = ln.chunkBegin; c < ln.chunkEnd; c++) {
                        Chunk ch = chunks[c];
                        ch.slot = k;
                        bc.push_back(ch);
                    }
                    std::memcpy(&nh[static_cast<size_t>(k) * C],
                            &out[active[ln.active].job].hist[static_cast<size_t>(ln.node) * C], C * sizeof(uint32_t));
                    st[k] = ln.active;
                }
                dChunks.resize(bc.size());
                dG.resize(static_cast<size_t>(ns) * perSlot);
                dAllScores.resize(static_cast<size_t>(ns) * T * F);
                dRepHist.resize(nh.size());
                dRepTree.resize(ns);
                dRepSumm.resize(static_cast<size_t>(ns) * F);
                dRepBest.resize(ns);
                CF_HIP(hipMemcpyAsync(dChunks.get(), bc.data(), bc.size() * sizeof(Chunk), hipMemcpyHostToDevice, s));
                CF_HIP(hipMemcpyAsync(dRepHist.get(), nh.data(), nh.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
                CF_HIP(hipMemcpyAsync(dRepTree.get(), st.data(), ns * sizeof(int), hipMemcpyHostToDevice, s));
                CF_HIP(hipMemsetAsync(dG.get(), 0, static_cast<size_t>(ns) * perSlot * sizeof(uint32_t), s));
                launchEvalHistograms(s, images->set(), level[cur].get(), dChunks.get(), static_cast<int>(bc.size()),
                        tabs, C, dG.get());
                launchScores(s, dG.get(), dRepHist.get(), dRepTree.get(), ns, tabs, C, dRepSumm.get(), dAllScores.get(),
                        -1);
                launchSequentialArgmax(s, dAllScores.get(), T, F, dRepBest.get(), ns);
                std::vector<NodeBest> rb(ns);
                CF_HIP(hipMemcpyAsync(rb.data(), dRepBest.get(), ns * sizeof(NodeBest), hipMemcpyDeviceToHost, s));
                CF_HIP(hipStreamSynchronize(s));
                timer.flush();
                for (int k = 0; k < ns; k++) best[flagged[a + k]] = rb[k];
                stats.replays += ns;
                a += ns;
            }
        };
        auto evalTwoPass = [&]() {
            int N = 0;
            for (int a = 0; a < numActive; a++) N = std::max(N, active[a].begin + active[a].n);
            // image-major order of the level: (tree, image, row) keys, counting sort of level indices.
            // With Both samples a second entry per level index evaluates it flipped; the entries of samples
            // that are not Both get the key one past the last (tree, image) and sort to the end.
            const int M = anyBoth ? 2 * N : N;
            dKeys.resize(M);
            dVals.resize(M);
            dPerm.resize(M);
            dLab8.resize(N);
            const uint64_t keyRange = static_cast<uint64_t>(numActive) * static_cast<uint64_t>(images->count);
            if (keyRange >= (1ull << 27)) throw Error(CF_ERR_UNSUPPORTED, "too many (tree, image) segments in one level");
            // within a segment, samples ordered by row while the buckets stay below 2^26 (CURFIL_ROW_ORDER=0: off)
            int yBits = 0;
            while ((1 << yBits) < images->H) yBits++;
            const char* rowOrder = std::getenv("CURFIL_ROW_ORDER");
            if (((keyRange + 2) << yBits) > (1ull << 26) || (rowOrder && std::strcmp(rowOrder, "0") == 0)) yBits = 0;
            // segments (image, tree) in image-major order: the trees' segments of one image are adjacent
            for (int a = 0; a < numActive; a++) {
                launchLevelKeys(s, level[cur].get(), active[a].begin, active[a].n, static_cast<uint32_t>(a),
                        static_cast<uint32_t>(numActive), dKeys.get(), dVals.get(), anyBoth ? N : 0,
                        static_cast<uint32_t>(keyRange), yBits);
            }
            const int numKeys = static_cast<int>(keyRange);
            dSortCount.resize(levelSortBuckets(numKeys, yBits));
            dSortCursor.resize(levelSortBuckets(numKeys, yBits));
            dSortBlocks.resize(levelSortScanBlocks(numKeys, yBits));
            dStarts.resize(numKeys + 1);
            levelCountingSort(s, dKeys.get(), dVals.get(), M, numKeys, yBits, dSortCount.get(), dSortCursor.get(),
                    dSortBlocks.get(), dPerm.get(), dStarts.get());
            launchLevelLabels(s, level[cur].get(), N, dLab8.get());
            // (image, tree) segments of the sorted order -> blocks of <= kBinBlockSamples samples within a segment;
            // the segments of one image form a group of consecutive blocks
            std::vector<int> starts(numKeys + 1);
            CF_HIP(hipMemcpyAsync(starts.data(), dStarts.get(), (numKeys + 1) * sizeof(int), hipMemcpyDeviceToHost, s));
            CF_HIP(hipStreamSynchronize(s));
            timer.flush();
            if (starts[numKeys] < 0) starts[numKeys] = M;   // the end of the last segment
            // blocks of one pass-1 block size: consecutive runs of <= blockSamples entries within a (image, tree)
            // segment; the segments of one image form a group of consecutive blocks.  Built on the host from the
            // segment starts for each block size the tuner asks for; the work list of a batch of ngb feature
            // groups is then built on the device: image group i goes to XCD i % 8; each XCD takes its image
            // groups in order, all feature groups of an image group (feature group major, then tree, then block)
            // before the next, so the trees' workgroups of one feature group read the same planes of one image
            // from the XCD's L2; workgroup w runs on XCD w % 8, so the lists are interleaved
            for (auto& kv : blockSets) kv.second.level = -1;
            auto blocksFor = [&](int blockSamples) -> BlockSet& {
                BlockSet& bs = blockSets[blockSamples];
                if (bs.level == levelNr) return bs;
                bs.level = levelNr;
                bs.workGroups = -1;
                std::vector<BinBlock> blocks;
                std::vector<std::pair<int, int>> segBlocks;   // per image group: first block, block count
                int k = 0;
                while (k < numKeys && starts[k] < 0) k++;
                while (k < numKeys) {
                    int k2 = k + 1;
                    while (k2 < numKeys && starts[k2] < 0) k2++;
                    const int b = starts[k], e = starts[k2];
                    const int first = static_cast<int>(blocks.size());
                    for (int p = b; p < e; p += blockSamples) {
                        blocks.push_back(BinBlock{k % numActive, p, std::min(blockSamples, e - p), k / numActive});
                    }
                    const int image = k / numActive;
                    if (!segBlocks.empty() && blocks[segBlocks.back().first].image == image) {
                        segBlocks.back().second += static_cast<int>(blocks.size()) - first;
                    } else {
                        segBlocks.emplace_back(first, static_cast<int>(blocks.size()) - first);
                    }
                    k = k2;
                }
                bs.blocks.resize(std::max<size_t>(blocks.size(), 1));
                if (!blocks.empty()) {
                    CF_HIP(hipMemcpyAsync(bs.blocks.get(), blocks.data(), blocks.size() * sizeof(BinBlock),
                            hipMemcpyHostToDevice, s));
                }
                constexpr int kXcds = 8;
                bs.numSegs = static_cast<int>(segBlocks.size());
                // first, count, XCD prefix per image group; XCD totals
                std::vector<int> segInfo(3 * static_cast<size_t>(bs.numSegs) + kXcds, 0);
                int* xcdBlocks = segInfo.data() + 3 * static_cast<size_t>(bs.numSegs);
                for (int i = 0; i < bs.numSegs; i++) {
                    segInfo[i] = segBlocks[i].first;
                    segInfo[bs.numSegs + i] = segBlocks[i].second;
                    segInfo[2 * static_cast<size_t>(bs.numSegs) + i] = xcdBlocks[i % kXcds];
                    xcdBlocks[i % kXcds] += segBlocks[i].second;
                }
                bs.maxXcdBlocks = *std::max_element(xcdBlocks, xcdBlocks + kXcds);
                bs.segInfo.resize(segInfo.size());
                CF_HIP(hipMemcpyAsync(bs.segInfo.get(), segInfo.data(), segInfo.size() * sizeof(int),
                        hipMemcpyHostToDevice, s));
                return bs;
            };
            // the work list of a batch of ngb feature groups into `work`; returns its length
            auto buildWork = [&](BlockSet& bs, int ngb, DevBuf<int2>& work) {
                const int maxLen = bs.maxXcdBlocks * ngb;
                const int numWork = maxLen * kNumXcds;
                work.resize(std::max<size_t>(static_cast<size_t>(numWork), 1));
                const int ns = bs.numSegs;
                launchBuildWork(s, bs.segInfo.get(), bs.segInfo.get() + ns, bs.segInfo.get() + 2 * ns,
                        bs.segInfo.get() + 3 * ns, ns, ngb, maxLen, work.get());
                return numWork;
            };
            auto workFor = [&](BlockSet& bs, int ngb) {
                if (bs.workGroups == ngb) return;
                bs.workGroups = ngb;
                bs.numWork = buildWork(bs, ngb, bs.work);
            };
            dEvalOrder.resize(static_cast<size_t>(numActive) * F);
            launchFeatureEvalOrder(s, tabs, numActive, dEvalOrder.get());
            dGFeat.resize(static_cast<size_t>(numActive) * F);
            dGThr.resize(static_cast<size_t>(numActive) * F * binThresholdWidth(T));
            launchBinTables(s, tabs, numActive, dEvalOrder.get(), dGFeat.get(), dGThr.get());
            // node histograms and trees of the whole level
            std::vector<uint32_t> nh(static_cast<size_t>(numNodes) * C);
            std::vector<int> st(numNodes);
            for (int i = 0; i < numNodes; i++) {
                const LevelNode& ln = nodes[i];
                std::memcpy(&nh[static_cast<size_t>(i) * C],
                        &out[active[ln.active].job].hist[static_cast<size_t>(ln.node) * C], C * sizeof(uint32_t));
                st[i] = ln.active;
            }
            dNodeHist.resize(nh.size());
            dSlotTree.resize(numNodes);
            dSumm.resize(static_cast<size_t>(numNodes) * F);
            dBest.resize(numNodes);
            CF_HIP(hipMemcpyAsync(dNodeHist.get(), nh.data(), nh.size() * sizeof(uint32_t), hipMemcpyHostToDevice, s));
            CF_HIP(hipMemcpyAsync(dSlotTree.get(), st.data(), numNodes * sizeof(int), hipMemcpyHostToDevice, s));
            // feature groups in batches that fit the bin-row budget; node slots in batches that fit the histogram budget
            const int groupsTotal = (F + kBinGroup - 1) / kBinGroup;
            // a group's rows: one per level index, and a second (flipped) one with Both samples
            const size_t rowBytes = static_cast<size_t>(N) * sizeof(uint4) * (anyBoth ? 2 : 1);
            const int ng = static_cast<int>(std::max<size_t>(1, std::min<size_t>(groupsTotal, binsBudget / rowBytes)));
            const size_t flipOffset = anyBoth ? static_cast<size_t>(ng) * N : 0;
            dBins.resize(static_cast<size_t>(ng) * N * (anyBoth ? 2 : 1));
            stats.groupBatchesMax = std::max(stats.groupBatchesMax, (groupsTotal + ng - 1) / ng);
            const int TM = binThresholdWidth(T);
            bool tuning = false;
            const int levelShape = evalTunerChoose(TM, static_cast<size_t>(M), &tuning);
            // node-slot batches of a batch of ngb feature groups: the slots whose histograms fit the budget
            auto slotBatches = [&](int ngb) {
                const size_t perSlot = static_cast<size_t>(ngb) * kBinGroup * (T + 1) * C;
                const int maxSlots = static_cast<int>(std::max<size_t>(1, budget / (perSlot * sizeof(uint32_t))));
                std::vector<std::pair<int, int>> out;
                for (int s0 = 0; s0 < numNodes; s0 += maxSlots) out.emplace_back(s0, std::min(numNodes, s0 + maxSlots));
                return out;
            };
            LevelGraph graph(s);
            if (!tuning && levelGraphs) {
                // The shape is settled: the level's evaluation (every batch's pass 1; per node-slot batch the
                // histogram clear, pass 2 and the scores; the argmax) is captured into one graph and launched
                // once, so the device runs it back to back with no host work between its kernels.  Work
                // lists, chunk tables and buffers are set up before the capture.
                const EvalShape& es = evalShape(levelShape);
                BlockSet& bs = blocksFor(es.lanes * es.perLane);
                struct Batch {
                    int g0, ngb;
                    const DevBuf<int2>* work;
                    int numWork;
                    std::vector<std::pair<int, int>> slots;
                };
                const int lastNgb = groupsTotal - ((groupsTotal - 1) / ng) * ng;
                const int numWorkFull = buildWork(bs, ng, dGraphWork[0]);
                const int numWorkLast = lastNgb == ng ? numWorkFull : buildWork(bs, lastNgb, dGraphWork[1]);
                bs.workGroups = -1;
                std::vector<Batch> batches;
                std::vector<Chunk> allChunks;
                std::vector<size_t> chunkOff;   // per (batch, slot batch), in order
                size_t maxG = 1;
                for (int g0 = 0; g0 < groupsTotal; g0 += ng) {
                    const int ngb = std::min(ng, groupsTotal - g0);
                    Batch b{g0, ngb, ngb == ng ? &dGraphWork[0] : &dGraphWork[1], ngb == ng ? numWorkFull : numWorkLast,
                            slotBatches(ngb)};
                    const size_t perSlotB = static_cast<size_t>(ngb) * kBinGroup * (T + 1) * C;
                    for (const auto& sb : b.slots) {
                        chunkOff.push_back(allChunks.size());
                        for (int c = nodes[sb.first].chunkBegin; c < nodes[sb.second - 1].chunkEnd; c++) {
                            Chunk ch = chunks[c];
                            ch.slot -= sb.first;
                            allChunks.push_back(ch);
                        }
                        maxG = std::max(maxG, static_cast<size_t>(sb.second - sb.first) * perSlotB);
                    }
                    batches.push_back(std::move(b));
                }
                dChunks.resize(std::max<size_t>(allChunks.size(), 1));
                CF_HIP(hipMemcpyAsync(dChunks.get(), allChunks.data(), allChunks.size() * sizeof(Chunk),
                        hipMemcpyHostToDevice, s));
                dG.resize(maxG);
                CF_HIP(hipStreamSynchronize(s));   // allChunks is read by the copy; nothing host-side in the capture
                timer.flush();
                graph.beginCapture();
                size_t k = 0;
                for (const Batch& b : batches) {
                    timer.begin();
                    launchEvalBins(s, images->set(), level[cur].get(), dPerm.get(), bs.blocks.get(), b.work->get(),
                            b.numWork, b.g0, dGFeat.get(), dGThr.get(), tabs, N, dBins.get(), flipOffset, levelShape);
                    timer.end(stats.evalKernelMs);
                    stats.evalLaunches += 1;
                    stats.shapeLaunches[levelShape] += 1;
                    stats.slotBatchesMax = std::max(stats.slotBatchesMax, static_cast<int>(b.slots.size()));
                    const size_t perSlotB = static_cast<size_t>(b.ngb) * kBinGroup * (T + 1) * C;
                    for (const auto& sb : b.slots) {
                        const int s0 = sb.first, ns = sb.second - sb.first;
                        const int nc = nodes[sb.second - 1].chunkEnd - nodes[s0].chunkBegin;
                        launchFillU32(s, dG.get(), static_cast<size_t>(ns) * perSlotB, 0u);
                        timer.begin();
                        launchHistBins(s, dBins.get(), dLab8.get(), dChunks.get() + chunkOff[k], nc, b.g0, b.ngb, tabs,
                                C, N, dG.get(), flipOffset);
                        timer.end(stats.histMs);
                        timer.begin();
                        launchScores(s, dG.get(), dNodeHist.get() + static_cast<size_t>(s0) * C, dSlotTree.get() + s0,
                                ns, tabs, C, dSumm.get() + static_cast<size_t>(s0) * F, nullptr, -1, b.g0 * kBinGroup,
                                b.ngb * kBinGroup, dEvalOrder.get());
                        timer.end(stats.scoreMs);
                        k++;
                    }
                }
                timer.begin();
                launchReduce(s, dSumm.get(), numNodes, F, dBest.get());
                timer.end(stats.scoreMs);
                graph.launch();
                stats.levelGraphs += 1;
            }
            for (int g0 = 0; g0 < (graph.launched() ? 0 : groupsTotal); g0 += ng) {
                const int ngb = std::min(ng, groupsTotal - g0);
                bool measure = false;
                const int shape = evalTunerChoose(TM, static_cast<size_t>(M), &measure);
                const EvalShape& es = evalShape(shape);
                BlockSet& bs = blocksFor(es.lanes * es.perLane);
                workFor(bs, ngb);
                if (measure) CF_HIP(hipEventRecord(tuneStart, s));
                timer.begin();
                launchEvalBins(s, images->set(), level[cur].get(), dPerm.get(), bs.blocks.get(), bs.work.get(),
                        bs.numWork, g0, dGFeat.get(), dGThr.get(), tabs, N, dBins.get(), flipOffset, shape);
                timer.end(stats.evalKernelMs);
                if (measure) {
                    CF_HIP(hipEventRecord(tuneEnd, s));
                    CF_HIP(hipEventSynchronize(tuneEnd));
                    float ms = 0;
                    CF_HIP(hipEventElapsedTime(&ms, tuneStart, tuneEnd));
                    // responses of the launch: every entry of the permutation for each feature of the batch
                    const int feats = std::min(F - g0 * kBinGroup, ngb * kBinGroup);
                    evalTunerRecord(TM, static_cast<size_t>(M), shape, ms, static_cast<double>(M) * feats);
                }
                stats.evalLaunches += 1;
                stats.shapeLaunches[shape] += 1;
                const size_t perSlotB = static_cast<size_t>(ngb) * kBinGroup * (T + 1) * C;
                const int maxSlotsB = static_cast<int>(std::max<size_t>(1, budget / (perSlotB * sizeof(uint32_t))));
                stats.slotBatchesMax = std::max(stats.slotBatchesMax, (numNodes + maxSlotsB - 1) / maxSlotsB);
                for (int s0 = 0; s0 < numNodes;) {
                    const int s1 = std::min(numNodes, s0 + maxSlotsB);
                    const int ns = s1 - s0;
                    std::vector<Chunk> bc(chunks.begin() + nodes[s0].chunkBegin, chunks.begin() + nodes[s1 - 1].chunkEnd);
                    for (auto& c : bc) c.slot -= s0;
                    dChunks.resize(bc.size());
                    dG.resize(static_cast<size_t>(ns) * perSlotB);
                    CF_HIP(hipMemcpyAsync(dChunks.get(), bc.data(), bc.size() * sizeof(Chunk), hipMemcpyHostToDevice, s));
                    CF_HIP(hipMemsetAsync(dG.get(), 0, static_cast<size_t>(ns) * perSlotB * sizeof(uint32_t), s));
                    timer.begin();
                    launchHistBins(s, dBins.get(), dLab8.get(), dChunks.get(), static_cast<int>(bc.size()), g0, ngb, tabs,
                            C, N, dG.get(), flipOffset);
                    timer.end(stats.histMs);
                    timer.begin();
                    launchScores(s, dG.get(), dNodeHist.get() + static_cast<size_t>(s0) * C, dSlotTree.get() + s0, ns,
                            tabs, C, dSumm.get() + static_cast<size_t>(s0) * F, nullptr, -1, g0 * kBinGroup,
                            ngb * kBinGroup, dEvalOrder.get());
                    timer.end(stats.scoreMs);
                    s0 = s1;
                }
            }
            if (!graph.launched()) {
                timer.begin();
                launchReduce(s, dSumm.get(), numNodes, F, dBest.get());
                timer.end(stats.scoreMs);
            }
            CF_HIP(hipMemcpyAsync(best.data(), dBest.get(), numNodes * sizeof(NodeBest), hipMemcpyDeviceToHost, s));
            CF_HIP(hipStreamSynchronize(s));
            timer.flush();
            std::vector<int> flagged;
            for (int i = 0; i < numNodes; i++) {
                if (best[i].flag) flagged.push_back(i);
            }
            replayNodes(flagged);
        };
        if (mode == CF_MODE_PLAIN) {
            for (int i = 0; i < numNodes; i++) evalSlots(i, i + 1);
        } else if (twoPass) {
            evalTwoPass();
        } else {
            for (int s0 = 0; s0 < numNodes;) {
                const int s1 = static_cast<int>(std::min<size_t>(numNodes, s0 + maxSlots));
                evalSlots(s0, s1);
                s0 = s1;
            }
        }

        if (levelLog) {
            double levelSamples = 0;
            for (const auto& ln : nodes) levelSamples += ln.count;
            std::fprintf(stderr, "[curfil] level %d: trees %d nodes %d samples %.0f chunks %zu eval %.2f ms (%.3g responses/s)\n",
                    levelNr, numActive, numNodes, levelSamples, chunks.size(), stats.evalKernelMs - evalMsBefore,
                    levelSamples * F / ((stats.evalKernelMs - evalMsBefore) * 1e-3));
        }

        // ---- split records, built on the device from the argmax and the level's feature tables
        std::vector<int> nodeTree(numNodes);
        for (int i = 0; i < numNodes; i++) nodeTree[i] = nodes[i].active;
        dSplits.resize(numNodes);
        dBestAll.resize(numNodes);
        dNodeTree.resize(numNodes);
        dChunkCounts.resize(chunks.size());
        dLeftHist.resize(static_cast<size_t>(numNodes) * C);
        dChunks.resize(chunks.size());
        CF_HIP(hipMemcpyAsync(dBestAll.get(), best.data(), numNodes * sizeof(NodeBest), hipMemcpyHostToDevice, s));
        CF_HIP(hipMemcpyAsync(dNodeTree.get(), nodeTree.data(), numNodes * sizeof(int), hipMemcpyHostToDevice, s));
        launchSplitRecords(s, dBestAll.get(), dNodeTree.get(), numNodes, tabs, dSplits.get());
        CF_HIP(hipMemcpyAsync(dChunks.get(), chunks.data(), chunks.size() * sizeof(Chunk), hipMemcpyHostToDevice, s));
        CF_HIP(hipMemsetAsync(dLeftHist.get(), 0, static_cast<size_t>(numNodes) * C * sizeof(uint32_t), s));
        timer.begin();
        launchPartitionDirections(s, images->desc.get(), level[cur].get(), dChunks.get(), static_cast<int>(chunks.size()),
                dSplits.get(), C, dir.get(), dChunkCounts.get(), dLeftHist.get());
        timer.end(stats.partitionMs);
        std::vector<int2> chunkCounts(chunks.size());   // left and right entries per chunk
        std::vector<uint32_t> leftHist(static_cast<size_t>(numNodes) * C);
        std::vector<SplitRec> splits(numNodes);
        CF_HIP(hipMemcpyAsync(chunkCounts.data(), dChunkCounts.get(), chunks.size() * sizeof(int2), hipMemcpyDeviceToHost,
                s));
        CF_HIP(hipMemcpyAsync(leftHist.data(), dLeftHist.get(), leftHist.size() * sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        CF_HIP(hipMemcpyAsync(splits.data(), dSplits.get(), numNodes * sizeof(SplitRec), hipMemcpyDeviceToHost, s));
        CF_HIP(hipStreamSynchronize(s));
        timer.flush();

        // ---- tree growth on the host, next frontier and sample destinations
        std::vector<int2> dst(chunks.size(), make_int2(-1, -1));
        std::vector<ActiveTree> next;
        int nextOff = 0;
        int slot = 0;
        for (int a = 0; a < numActive; a++) {
            ActiveTree& at = active[a];
            TreeBuild& tb = out[at.job];
            ActiveTree na;
            na.job = at.job;
            na.begin = nextOff;
            for (size_t k = 0; k < at.nodes.size(); k++, slot++) {
                const LevelNode& ln = nodes[slot];
                int nL = 0, nR = 0;
                for (int c = ln.chunkBegin; c < ln.chunkEnd; c++) {
                    nL += chunkCounts[c].x;
                    nR += chunkCounts[c].y;
                }
                const int leftId = ++at.idNode;
                const int rightId = ++at.idNode;
                if (nL == 0 || nR == 0) {
                    at.idNode -= 2;
                    continue;
                }
                const int pi = ln.node;
                const int li = static_cast<int>(tb.nodes.size());
                for (int side = 0; side < 2; side++) {
                    cf_tree_node cn{};
                    cn.id = side == 0 ? leftId : rightId;
                    cn.level = tb.nodes[pi].level + 1;
                    cn.parent = pi;
                    cn.left = cn.right = -1;
                    cn.num_samples = side == 0 ? nL : nR;
                    cn.is_leaf = 1;
                    tb.nodes.push_back(cn);
                    for (int c = 0; c < C; c++) {
                        const uint32_t l = leftHist[static_cast<size_t>(slot) * C + c];
                        tb.hist.push_back(side == 0 ? l : tb.hist[static_cast<size_t>(pi) * C + c] - l);
                    }
                }
                cf_tree_node& p = tb.nodes[pi];
                p.is_leaf = 0;
                p.left = li;
                p.right = li + 1;
                p.feature_id = best[slot].feat;
                p.threshold = splits[slot].threshold;
                p.score = best[slot].score;
                for (int q = 0; q < kFeatParams; q++) p.feature[q] = splits[slot].feat.p[q];
                int childBegin[2] = {-1, -1};
                for (int side = 0; side < 2; side++) {
                    const int ci = li + side;
                    const int cnt = side == 0 ? nL : nR;
                    if (!isPure(&tb.hist[static_cast<size_t>(ci) * C], C) &&
                            cnt >= static_cast<int>(cfg.min_sample_count)) {
                        childBegin[side] = nextOff;
                        na.nodes.push_back(ci);
                        na.counts.push_back(cnt);
                        nextOff += cnt;
                    }
                }
                int doneL = 0, doneR = 0;
                for (int c = ln.chunkBegin; c < ln.chunkEnd; c++) {
                    const int cl = chunkCounts[c].x;
                    const int cr = chunkCounts[c].y;
                    dst[c].x = childBegin[0] >= 0 ? childBegin[0] + doneL : -1;
                    dst[c].y = childBegin[1] >= 0 ? childBegin[1] + doneR : -1;
                    doneL += cl;
                    doneR += cr;
                }
            }
            na.n = nextOff - na.begin;
            na.idNode = at.idNode;
            if (!na.nodes.empty()) next.push_back(std::move(na));
        }
        if (nextOff > 0) {
            dDst.resize(dst.size());
            CF_HIP(hipMemcpyAsync(dDst.get(), dst.data(), dst.size() * sizeof(int2), hipMemcpyHostToDevice, s));
            timer.begin();
            launchPartitionScatter(s, level[cur].get(), dChunks.get(), static_cast<int>(chunks.size()), dir.get(),
                    dDst.get(), level[1 - cur].get());
            timer.end(stats.partitionMs);
            cur = 1 - cur;
        }
        // the host vectors handed to the copies above live until here
        CF_HIP(hipStreamSynchronize(s));
        timer.flush();
        active.swap(next);
    }
    CF_HIP(hipStreamSynchronize(s));
    timer.flush();
    stats.seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace cf
