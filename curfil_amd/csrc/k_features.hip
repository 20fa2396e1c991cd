// Per-(tree, level) random feature generation on the device, with the exact
// semantics of the reference CPU generator:
//   ImageFeatureEvaluation::generateRandomFeatures  random_tree_image.cpp:577-646
//   ImageFeatureEvaluation::sampleFeature           random_tree_image.cpp:648-700
//   sortFeatures (stable sort by getSortKey)        random_tree_image.cpp:504-566,
//                                                   random_tree_image.h:444-452
//
// RandomSource(seed) hands seed to the sample-index sampler and seeds
// seed+1+4k .. seed+4+4k to candidate k's four samplers (type, offset,
// region, channel).  Candidates are therefore independent and are drawn in
// parallel (k_fg_candidates).  Thresholds consume one shared draw stream in
// order, and how many draws a candidate consumes depends on which responses
// are NaN; k_fg_walk keeps that order exactly: one workgroup per (tree, level)
// owns the sample-index generator (MT19937 state twisted in parallel in LDS),
// evaluates the next 16 candidates, each in a band of 128 of the next 512
// queued draws, in parallel, and lets one wave walk them with the reference's
// retry rules.
#include "cf_internal.h"
#include "cf_rng.h"

namespace cf {

constexpr int kPrefixDraws = 24;   // engine outputs available to one feature sampler
constexpr int kDrawQueue = 2048;   // ring of pending sample-index draws

// g++ on x86-64 evaluates the two getNext() calls of `Offset(rgen.getNext(),
// rgen.getNext())` right to left, so the second coordinate is drawn first
// (pinned by the reference's golden feature hash).
__global__ void __launch_bounds__(256) k_fg_candidates(const int* __restrict__ seeds, int numCand, int boxRadius,
        int regionSize, int useDepth, FeatRec* __restrict__ cand, int* __restrict__ err) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int job = blockIdx.y;
    if (k >= numCand) return;
    const int base = seeds[job] + 1 + 4 * k;
    FeatRec f;
#pragma unroll
    for (int i = 0; i < 16; i++) f.p[i] = 0;
    bool ok = true;
    int a, b;
    {
        DevPrefixSampler<kPrefixDraws> xy;
        xy.init(base + 1, -boxRadius, boxRadius);
        int o1x = 0, o1y = 0, o2x = 0, o2y = 0;
        ok &= xy.next(b);
        ok &= xy.next(a);
        o1x = a;
        o1y = b;
        do {
            ok &= xy.next(b);
            ok &= xy.next(a);
            o2x = a;
            o2y = b;
        } while (ok && o1x == o2x && o1y == o2y);
        f.p[P_O1X] = static_cast<int8_t>(o1x);
        f.p[P_O1Y] = static_cast<int8_t>(o1y);
        f.p[P_O2X] = static_cast<int8_t>(o2x);
        f.p[P_O2Y] = static_cast<int8_t>(o2y);
    }
    {
        DevPrefixSampler<kPrefixDraws> rg;
        rg.init(base + 2, 1, regionSize);
        ok &= rg.next(b);
        ok &= rg.next(a);
        f.p[P_R1X] = static_cast<int8_t>(a);
        f.p[P_R1Y] = static_cast<int8_t>(b);
        ok &= rg.next(b);
        ok &= rg.next(a);
        f.p[P_R2X] = static_cast<int8_t>(a);
        f.p[P_R2Y] = static_cast<int8_t>(b);
    }
    int type = T_COLOR;
    if (useDepth) {
        DevPrefixSampler<kPrefixDraws> ft;
        ft.init(base, 0, 1);
        int v = 0;
        ok &= ft.next(v);
        type = (v == 0) ? T_DEPTH : T_COLOR;
    }
    f.p[P_TYPE] = static_cast<int8_t>(type);
    if (type == T_COLOR) {
        DevPrefixSampler<kPrefixDraws> cc;
        cc.init(base + 3, 0, 2);
        ok &= cc.next(b);
        ok &= cc.next(a);
        f.p[P_C1] = static_cast<int8_t>(a);
        f.p[P_C2] = static_cast<int8_t>(b);
    }
    if (!ok) atomicExch(err, 2);
    cand[static_cast<size_t>(job) * numCand + k] = f;
}

// getSortKey, random_tree_image.h:444-452
__device__ __forceinline__ int featureSortKey(const FeatRec& f) {
    int k = 0;
    k |= static_cast<int>(static_cast<uint8_t>(f.p[P_TYPE] & 0x03)) << 30;
    k |= static_cast<int>(static_cast<uint8_t>(f.p[P_C1] & 0x0F)) << 26;
    k |= static_cast<int>(static_cast<uint8_t>(f.p[P_C2] & 0x0F)) << 22;
    k |= static_cast<int>(static_cast<uint8_t>((f.p[P_O1Y] + 127) & 0xFF)) << 14;
    k |= static_cast<int>(static_cast<uint8_t>((f.p[P_O1X] + 127) & 0xFF)) << 6;
    return k;
}

constexpr int kWalkCand = 16;   // candidates evaluated per round
constexpr int kWalkDraws = 512; // queued draws a round can reach
constexpr int kWalkBand = 128;  // draws evaluated per candidate per round
constexpr int kWalkWords = kWalkDraws / 64;
static_assert(kDrawQueue >= kWalkDraws + 624, "the draw ring holds a round's window and one refill");

// What the walk did with candidate j of a round: window positions
// [begin, end) were consumed with st = 0 (still open at its band's end),
// 1 (accepted into output slot `slot`) or 2 (rejected, or not reached); t0
// thresholds had been found before the round.
struct WalkStep {
    int st, begin, end, t0, slot;
};

struct WalkShared {
    uint32_t mt[624];
    int dq[kDrawQueue];
    float resp[kWalkCand][kWalkDraws];
    unsigned long long finite[kWalkCand][kWalkWords];   // bit: response is not NaN (set inside the band only)
    FeatRec cand[kWalkCand];
    WalkStep step[kWalkCand];
    int waveCounts[4];
    // walk state after a round: next draw, next candidate, features generated,
    // thresholds found and tries left for a candidate that ran past the round;
    // band spacing of the next round and the draws / candidates finished so far
    int head, k, generated, t, tries, err;
    int spacing, drawsDone, candsDone;
};

// One MT19937 twist of the LDS state, in three dependency-free phases.
__device__ void mtTwistShared(uint32_t* mt) {
    const int tid = threadIdx.x;
    uint32_t v = 0;
    if (tid < 227) v = mtMix(mt[tid], mt[tid + 1], mt[tid + 397]);
    __syncthreads();
    if (tid < 227) mt[tid] = v;
    __syncthreads();
    if (tid < 227) v = mtMix(mt[227 + tid], mt[228 + tid], mt[tid]);
    __syncthreads();
    if (tid < 227) mt[227 + tid] = v;
    __syncthreads();
    if (tid < 169) v = mtMix(mt[454 + tid], mt[455 + tid], mt[227 + tid]);
    __syncthreads();
    if (tid < 169) mt[454 + tid] = v;
    __syncthreads();
    if (tid == 0) mt[623] = mtMix(mt[623], mt[0], mt[396]);
    __syncthreads();
}

// Appends the accepted uniform_int draws of one engine block to the queue.
__device__ int refillDraws(WalkShared& sh, int tail, const UniformIntParams& up) {
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6;
    if (up.range == 0) {
        for (int i = tid; i < 624; i += 256) sh.dq[(tail + i) % kDrawQueue] = up.lo;
        __syncthreads();
        return tail + 624;
    }
    mtTwistShared(sh.mt);
    for (int r = 0; r < 3; r++) {
        const int i = r * 256 + tid;
        bool acc = false;
        uint32_t v = 0;
        if (i < 624) {
            v = mtTemper(sh.mt[i]) / up.bucket;
            acc = v <= up.range;
        }
        const unsigned long long mask = __ballot(acc);
        const int lanePrefix = __popcll(mask & ((1ull << lane) - 1ull));
        if (lane == 0) sh.waveCounts[wave] = __popcll(mask);
        __syncthreads();
        int wavePrefix = 0, total = 0;
        for (int w = 0; w < 4; w++) {
            if (w < wave) wavePrefix += sh.waveCounts[w];
            total += sh.waveCounts[w];
        }
        if (acc) sh.dq[(tail + wavePrefix + lanePrefix) % kDrawQueue] = static_cast<int>(static_cast<uint32_t>(up.lo) + v);
        tail += total;
        __syncthreads();
    }
    return tail;
}

// Per round: candidate j of the next kWalkCand is evaluated at the draws
// [j S, j S + kWalkBand) of the next kWalkDraws queued draws (S: about 7/8 of
// the draws a finished candidate has used, at least min(T, 10), the fewest a
// candidate can use), 2 candidates x 2 draws per lane and wave; then wave 0
// walks the candidates in order with the reference's rules
// (random_tree_image.cpp:648-700: a NaN response costs one of 10 tries, a
// finite one is a threshold and resets the tries; T thresholds accept the
// candidate, 10 NaN in a row reject it).  A candidate's first draw is the one
// after its predecessor's last.  The round ends at the first candidate whose
// first draw lies outside its band or which runs to its band's end; that
// candidate carries its state into the next round, which starts with it.
// The bands only decide how far a round gets, never which draws a candidate
// uses.
__global__ void __launch_bounds__(256) k_fg_walk(const FeatGenJob* __restrict__ jobs, ImageSet imgs,
        const FeatRec* __restrict__ cand, int numCand, int F, int T, FeatRec* __restrict__ outFeat,
        float* __restrict__ outThr, int* __restrict__ outKey, int* __restrict__ err) {
    __shared__ WalkShared sh;
    const FeatGenJob job = jobs[blockIdx.x];
    const FeatRec* myCand = cand + static_cast<size_t>(blockIdx.x) * numCand;
    const int tid = threadIdx.x;
    const int wave = tid >> 6, lane = tid & 63;
    const UniformIntParams up = uniformIntParams(0, job.n - 1);
    const int minUse = min(T, 10);
    if (tid == 0) {
        uint32_t v = static_cast<uint32_t>(job.seed);
        sh.mt[0] = v;
        for (uint32_t i = 1; i < 624; i++) {
            v = mtInitStep(v, i);
            sh.mt[i] = v;
        }
        sh.head = 0;
        sh.k = 0;
        sh.generated = 0;
        sh.t = 0;
        sh.tries = 10;
        sh.err = 0;
        sh.spacing = minUse;
        sh.drawsDone = 0;
        sh.candsDone = 0;
    }
    __syncthreads();
    constexpr int kPerWave = kWalkCand / 4;          // candidates per wave: wave, wave + 4, ...
    constexpr int kReps = kWalkBand / 64;            // draws per lane and candidate
    constexpr int kSlots = kPerWave * kReps;
    int tail = 0;
    for (;;) {
        const int head = sh.head, k0 = sh.k, S = sh.spacing;
        const bool done = sh.generated >= F || sh.err != 0;
        __syncthreads();
        if (done) break;
        while (tail - head < kWalkDraws) tail = refillDraws(sh, tail, up);
        for (int i = tid; i < kWalkCand * kWalkWords; i += 256) (&sh.finite[0][0])[i] = 0ull;
        if (tid < kWalkCand && k0 + tid < numCand) sh.cand[tid] = myCand[k0 + tid];
        __syncthreads();
        {
            // two rounds of loads for all slots together: first every slot's sample record, then every
            // slot's corner gathers (planes addressed from the image index, no descriptor load); a
            // response is formed only after all of them are in flight
            bool ok[kSlots], isDepth[kSlots], active[kSlots];
            int2 cv[kSlots][8];   // depth corners, or colour corners' bits in .x
            cf_sample smp[kSlots];
#pragma unroll
            for (int q = 0; q < kPerWave; q++) {
                const int c = wave + 4 * q;
                const bool candOk = k0 + c < numCand;
#pragma unroll
                for (int r = 0; r < kReps; r++) {
                    const int i = q * kReps + r;
                    const int dpos = c * S + 64 * r + lane;
                    active[i] = candOk && dpos < kWalkDraws;
                    // inactive slots read the level's first sample (in range) and issue no gathers
                    const int d = active[i] ? sh.dq[(head + dpos) % kDrawQueue] : 0;
                    smp[i] = job.samples[d];
                }
            }
            const size_t plane = static_cast<size_t>(imgs.W) * imgs.H;
#pragma unroll
            for (int q = 0; q < kPerWave; q++) {
                const int c = wave + 4 * q;
                const bool candOk = k0 + c < numCand;
                Feat ff{};
                if (candOk) ff = loadFeat(&sh.cand[c]);
#pragma unroll
                for (int r = 0; r < kReps; r++) {
                    const int i = q * kReps + r;
                    isDepth[i] = ff.type == T_DEPTH;
                    const bool valid = smp[i].depth_mm > 0;
                    const double rinv = valid ? depthReciprocal(depthToFloat(smp[i].depth_mm)) : 0.0;
                    int idx[8];
                    ok[i] = active[i] && valid && responseCorners(ff, smp[i].x, smp[i].y, imgs.W, imgs.H, rinv, false,
                            idx);
                    if (ok[i]) {
                        if (isDepth[i]) {
                            const int2* dep = imgs.dep + static_cast<size_t>(smp[i].image) * plane;
#pragma unroll
                            for (int e = 0; e < 8; e++) cv[i][e] = planeAt(dep, idx[e]);
                        } else {
                            const float* col = imgs.col + static_cast<size_t>(smp[i].image) * 3 * plane;
                            const float* colA = col + static_cast<size_t>(ff.c1) * plane;
                            const float* colB = col + static_cast<size_t>(ff.c2) * plane;
#pragma unroll
                            for (int e = 0; e < 4; e++) cv[i][e].x = __float_as_int(planeAt(colA, idx[e]));
#pragma unroll
                            for (int e = 4; e < 8; e++) cv[i][e].x = __float_as_int(planeAt(colB, idx[e]));
                        }
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < kPerWave; q++) {
                const int c = wave + 4 * q;
#pragma unroll
                for (int r = 0; r < kReps; r++) {
                    const int i = q * kReps + r;
                    const int start = c * S + 64 * r;   // draw of lane 0
                    const int dpos = start + lane;
                    float v = __builtin_nanf("");
                    if (ok[i]) {
                        if (isDepth[i]) {
                            v = static_cast<float>(depthResponseFromCorners(cv[i]));
                        } else {
                            float c8[8];
#pragma unroll
                            for (int e = 0; e < 8; e++) c8[e] = __int_as_float(cv[i][e].x);
                            v = static_cast<float>(colorResponseFromCorners(c8));
                        }
                    }
                    if (active[i]) sh.resp[c][dpos] = v;
                    // the wave's 64 draws start at `start`, which need not be word aligned
                    const unsigned long long fm = __ballot(active[i] && !__builtin_isnan(v));
                    if (lane == 0 && fm != 0ull) {
                        const int w = start >> 6, off = start & 63;
                        atomicOr(&sh.finite[c][w], fm << off);
                        if (off != 0 && w + 1 < kWalkWords) atomicOr(&sh.finite[c][w + 1], fm >> (64 - off));
                    }
                }
            }
        }
        __syncthreads();
        if (tid < 64) {
            // wave 0 walks the finite-response bits a 64-draw word at a time: lane q
            // stands for draw q of the word and tests whether the walk stops there
            // (the T-th threshold, or the NaN that uses up the tries: a NaN run
            // costs one try per draw, a finite response is the next threshold and
            // resets the tries to 10); the first such lane ends the candidate
            const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));   // bits [0, lane)
            int p = 0;
            int k = k0;
            int gen = __builtin_amdgcn_readfirstlane(sh.generated);
            int t = __builtin_amdgcn_readfirstlane(sh.t);
            int tries = __builtin_amdgcn_readfirstlane(sh.tries);
            if (tid < kWalkCand) sh.step[tid].st = 2;   // not walked: nothing to copy
            while (k < k0 + kWalkCand && gen < F) {
                const int j = k - k0;
                const int lo = j * S, hi = min(lo + kWalkBand, kWalkDraws);
                if (p < lo || p >= hi) break;   // its first draw was not evaluated: the next round starts here
                // `--maxLoops <= 0` with maxLoops = 3F, checked right after sampling candidate k
                if (k + 1 >= 3 * F) {
                    if (tid == 0) {
                        sh.err = 1;
                        atomicExch(err, 1);
                    }
                    break;
                }
                const int begin = p, t0 = t, slot = job.outOffset + gen;
                int st = 0;
                while (st == 0 && p < hi) {
                    const int w = p >> 6, o = p & 63;
                    const int e = min(64, hi - (w << 6));   // lanes [o, e) lie in the band
                    const unsigned long long word = sh.finite[j][w];
                    const uint32_t lo32 = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(word));
                    const uint32_t hi32 = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(word >> 32));
                    const unsigned long long inBand = (~0ull << o) & (e == 64 ? ~0ull : ((1ull << e) - 1ull));
                    const unsigned long long m = ((static_cast<unsigned long long>(hi32) << 32) | lo32) & inBand;
                    // finite draws in [p, q) and the NaN run that ends at q: it starts after the last of
                    // them and is held to 10 tries, or, with none, starts at p with the tries carried in
                    const unsigned long long prevF = m & below;
                    const bool fin = (m >> lane) & 1ull;
                    const int run = prevF ? lane - (63 - __clzll(prevF)) : lane - o + 1;
                    const bool accept = fin && t + __popcll(prevF) + 1 == T;
                    const bool reject = lane >= o && lane < e && !fin && run == (prevF ? 10 : tries);
                    const unsigned long long stop = __ballot(accept || reject);
                    if (stop) {
                        const int q = __builtin_ctzll(stop);
                        st = ((m >> q) & 1ull) ? 1 : 2;
                        p = (w << 6) + q + 1;
                    } else {
                        // no stop before e: its thresholds count, and the NaN draws after the last
                        // finite one (fewer than the tries left) carry on
                        t += __popcll(m);
                        tries = m ? 10 - (e - 1 - (63 - __clzll(m))) : tries - (e - o);
                        p = (w << 6) + e;
                    }
                }
                if (tid == 0) {
                    WalkStep ws;
                    ws.st = st;
                    ws.begin = begin;
                    ws.end = p;
                    ws.t0 = t0;
                    ws.slot = slot;
                    sh.step[j] = ws;
                    if (st == 1) {
                        const FeatRec fr = sh.cand[j];
                        outFeat[slot] = fr;
                        outKey[slot] = featureSortKey(fr);
                    }
                }
                if (st == 0) break;   // the band ran out inside candidate k
                if (st == 1) gen++;
                k++;
                t = 0;
                tries = 10;
            }
            if (tid == 0) {
                sh.head = head + p;
                sh.k = k;
                sh.generated = gen;
                sh.t = t;
                sh.tries = tries;
                // draws and candidates finished this round (an open candidate's draws count with it later)
                int used = p;
                if (k < k0 + kWalkCand) {
                    const WalkStep& open = sh.step[k - k0];
                    if (open.st == 0) used = open.begin;
                }
                const int drawsDone = sh.drawsDone + used, candsDone = sh.candsDone + (k - k0);
                sh.drawsDone = drawsDone;
                sh.candsDone = candsDone;
                sh.spacing = candsDone > 0 ? max(minUse, (7 * drawsDone) / (8 * candsDone)) : minUse;
            }
        }
        __syncthreads();
        // thresholds of the accepted and still open candidates: the finite
        // draws of [begin, end) in order, after the t0 found earlier
#pragma unroll
        for (int q = 0; q < kPerWave; q++) {
            const int c = wave + 4 * q;
            const WalkStep ws = sh.step[c];
            if (ws.st == 2) continue;
#pragma unroll
            for (int r = 0; r < kReps; r++) {
                const int dpos = c * S + 64 * r + lane;
                if (dpos < ws.begin || dpos >= ws.end) continue;
                if (!((sh.finite[c][dpos >> 6] >> (dpos & 63)) & 1ull)) continue;
                int before = 0;   // finite draws in [begin, dpos)
                for (int w = ws.begin >> 6; w <= (dpos >> 6); w++) {
                    unsigned long long m = sh.finite[c][w];
                    if (w == (ws.begin >> 6)) m &= ~0ull << (ws.begin & 63);
                    if (w == (dpos >> 6)) m &= (1ull << (dpos & 63)) - 1ull;
                    before += __popcll(m);
                }
                outThr[static_cast<size_t>(ws.slot) * T + ws.t0 + before] = sh.resp[c][dpos];
            }
        }
        __syncthreads();
    }
}

// Stable sort of each job's F features by key (rank = #smaller keys + #equal
// keys earlier), scattering features and thresholds into both layouts.  The
// grid is (F / 256 blocks, jobs): every workgroup holds all of its job's keys
// in LDS and ranks 256 of its features.
__global__ void __launch_bounds__(256) k_fg_sort(const int* __restrict__ keys, const FeatRec* __restrict__ inFeat,
        const float* __restrict__ inThr, int F, int T, FeatRec* __restrict__ feat, float* __restrict__ thrFT,
        float* __restrict__ thrTF, int8_t* __restrict__ soa) {
    extern __shared__ int sk[];
    const int job = blockIdx.y;
    const int* jk = keys + static_cast<size_t>(job) * F;
    for (int i = threadIdx.x; i < F; i += blockDim.x) sk[i] = jk[i];
    __syncthreads();
    {
        const int i = blockIdx.x * blockDim.x + threadIdx.x;
        if (i >= F) return;
        const int ki = sk[i];
        int rank = 0;
        for (int j = 0; j < F; j++) {
            const int kj = sk[j];
            rank += (kj < ki || (kj == ki && j < i)) ? 1 : 0;
        }
        const size_t src = static_cast<size_t>(job) * F + i;
        const size_t dst = static_cast<size_t>(job) * F + rank;
        const FeatRec fr = inFeat[src];
        feat[dst] = fr;
        for (int p = 0; p < kFeatParams; p++) soa[(static_cast<size_t>(job) * kFeatParams + p) * F + rank] = fr.p[p];
        for (int t = 0; t < T; t++) {
            const float v = inThr[src * T + t];
            thrFT[dst * T + t] = v;
            thrTF[(static_cast<size_t>(job) * T + t) * F + rank] = v;
        }
    }
}

// Reference layout in (soa [trees][11][F], thrTF [trees][T][F]) -> feat, thrFT.
__global__ void __launch_bounds__(256) k_tables_from_reference(const int8_t* __restrict__ soa,
        const float* __restrict__ thrTF, int trees, int F, int T, FeatRec* __restrict__ feat,
        float* __restrict__ thrFT) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= trees * F) return;
    const int job = i / F, f = i % F;
    FeatRec fr;
#pragma unroll
    for (int p = 0; p < 16; p++) fr.p[p] = 0;
    for (int p = 0; p < kFeatParams; p++) fr.p[p] = soa[(static_cast<size_t>(job) * kFeatParams + p) * F + f];
    feat[i] = fr;
    for (int t = 0; t < T; t++) thrFT[static_cast<size_t>(i) * T + t] = thrTF[(static_cast<size_t>(job) * T + t) * F + f];
}

// Per feature: rank[t] = #{k : thr[k] < thr[t]} and the thresholds ordered by
// (value, index).  A sample with float key q goes left at threshold t iff
// q <= thr[t] iff #{k : thr[k] < q} <= rank[t].
__global__ void __launch_bounds__(256) k_threshold_ranks(const float* __restrict__ thrFT, int total, int T,
        uint16_t* __restrict__ rank, uint16_t* __restrict__ order) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float* th = thrFT + static_cast<size_t>(i) * T;
    // order: by (value, index) with NaN after every number, so `order` is a permutation
    // whatever the table holds (the callers refuse NaN thresholds)
    for (int t = 0; t < T; t++) {
        const float v = th[t];
        const bool vNan = v != v;
        int r = 0, pos = 0;
        for (int k = 0; k < T; k++) {
            const float u = th[k];
            const bool uNan = u != u;
            const bool below = vNan ? !uNan : u < v;
            const bool same = vNan ? uNan : u == v;
            r += (u < v) ? 1 : 0;
            pos += (below || (same && k < t)) ? 1 : 0;
        }
        rank[static_cast<size_t>(i) * T + t] = static_cast<uint16_t>(r);
        order[static_cast<size_t>(i) * T + pos] = static_cast<uint16_t>(t);
    }
}

void launchFeatureGeneration(hipStream_t s, const ImageSet& imgs, const FeatGenJob* jobsDev, int numJobs, int F,
        int T, int boxRadius, int regionSize, int useDepth, const int* seedsHost, FeatTables& tabs) {
    if (T > kMaxThresholds) throw Error(CF_ERR_ARG, "too many thresholds (max 256)");
    if (F > 8192) throw Error(CF_ERR_ARG, "too many features (max 8192)");
    if (boxRadius > 127 || regionSize > 127) throw Error(CF_ERR_ARG, "boxRadius/regionSize must fit int8");
    const int numCand = 3 * F;
    tabs.trees = numJobs;
    tabs.F = F;
    tabs.T = T;
    tabs.cand.resize(static_cast<size_t>(numJobs) * numCand);
    tabs.unsortedFeat.resize(static_cast<size_t>(numJobs) * F);
    tabs.unsortedThr.resize(static_cast<size_t>(numJobs) * F * T);
    tabs.unsortedKey.resize(static_cast<size_t>(numJobs) * F);
    tabs.feat.resize(static_cast<size_t>(numJobs) * F);
    tabs.thrFT.resize(static_cast<size_t>(numJobs) * F * T);
    tabs.thrTF.resize(static_cast<size_t>(numJobs) * F * T);
    tabs.soa.resize(static_cast<size_t>(numJobs) * F * kFeatParams);
    tabs.rank.resize(static_cast<size_t>(numJobs) * F * T);
    tabs.order.resize(static_cast<size_t>(numJobs) * F * T);
    tabs.err.resize(1);
    CF_HIP(hipMemsetAsync(tabs.err.get(), 0, sizeof(int), s));
    DevBuf<int> seeds(numJobs);
    CF_HIP(hipMemcpyAsync(seeds.get(), seedsHost, numJobs * sizeof(int), hipMemcpyHostToDevice, s));
    hipLaunchKernelGGL(k_fg_candidates, dim3((numCand + 255) / 256, numJobs), dim3(256), 0, s, seeds.get(), numCand,
            boxRadius, regionSize, useDepth, tabs.cand.get(), tabs.err.get());
    hipLaunchKernelGGL(k_fg_walk, dim3(numJobs), dim3(256), 0, s, jobsDev, imgs, tabs.cand.get(), numCand, F, T,
            tabs.unsortedFeat.get(), tabs.unsortedThr.get(), tabs.unsortedKey.get(), tabs.err.get());
    hipLaunchKernelGGL(k_fg_sort, dim3((F + 255) / 256, numJobs), dim3(256), F * sizeof(int), s, tabs.unsortedKey.get(),
            tabs.unsortedFeat.get(), tabs.unsortedThr.get(), F, T, tabs.feat.get(), tabs.thrFT.get(),
            tabs.thrTF.get(), tabs.soa.get());
    const int total = numJobs * F;
    hipLaunchKernelGGL(k_threshold_ranks, dim3((total + 255) / 256), dim3(256), 0, s, tabs.thrFT.get(), total, T,
            tabs.rank.get(), tabs.order.get());
    int errHost = 0;
    CF_HIP(hipMemcpyAsync(&errHost, tabs.err.get(), sizeof(int), hipMemcpyDeviceToHost, s));
    CF_HIP(hipStreamSynchronize(s));
    if (errHost == 1) throw Error(CF_ERR_RUNTIME, "failed to generate random features. max loops exceeded");
    if (errHost == 2) throw Error(CF_ERR_RUNTIME, "feature sampler needed more engine outputs than provisioned");
}

void launchPrepareTables(hipStream_t s, int trees, int F, int T, FeatTables& tabs) {
    if (T > kMaxThresholds) throw Error(CF_ERR_ARG, "too many thresholds (max 256)");
    tabs.trees = trees;
    tabs.F = F;
    tabs.T = T;
    tabs.feat.resize(static_cast<size_t>(trees) * F);
    tabs.thrFT.resize(static_cast<size_t>(trees) * F * T);
    tabs.rank.resize(static_cast<size_t>(trees) * F * T);
    tabs.order.resize(static_cast<size_t>(trees) * F * T);
    const int total = trees * F;
    hipLaunchKernelGGL(k_tables_from_reference, dim3((total + 255) / 256), dim3(256), 0, s, tabs.soa.get(),
            tabs.thrTF.get(), trees, F, T, tabs.feat.get(), tabs.thrFT.get());
    hipLaunchKernelGGL(k_threshold_ranks, dim3((total + 255) / 256), dim3(256), 0, s, tabs.thrFT.get(), total, T,
            tabs.rank.get(), tabs.order.get());
}

}  // namespace cf
