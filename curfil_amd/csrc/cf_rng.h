// Random numbers of the reference CPU path: boost::mt19937 driven through
// boost::uniform_int<int> (Sampler, random_tree.h:1009-1063; getNext,
// random_tree.cpp:107-112), with RandomSource handing out consecutive seeds
// (random_tree.h:1129-1156).
//
// MT19937 is the standard generator (Matsumoto & Nishimura 1998).  The
// integer distribution is Boost's bucket method: with the 32-bit engine range
// brange = 2^32-1 and range r = hi-lo, bucket = brange / (r+1), plus one when
// brange % (r+1) == r; draws v = engine() / bucket until v <= r; the result is
// lo + v.  A degenerate range (lo == hi) returns lo without touching the engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cf {

__host__ __device__ __forceinline__ uint32_t mtTemper(uint32_t y) {
    y ^= (y >> 11);
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= (y >> 18);
    return y;
}

__host__ __device__ __forceinline__ uint32_t mtMix(uint32_t cur, uint32_t nxt, uint32_t far) {
    const uint32_t y = (cur & 0x80000000u) | (nxt & 0x7fffffffu);
    return far ^ (y >> 1) ^ ((y & 1u) ? 0x9908b0dfu : 0u);
}

__host__ __device__ __forceinline__ uint32_t mtInitStep(uint32_t prev, uint32_t i) {
    return 1812433253u * (prev ^ (prev >> 30)) + i;
}

struct UniformIntParams {
    int lo;
    uint32_t range;
    uint32_t bucket;
};

__host__ __device__ __forceinline__ UniformIntParams uniformIntParams(int lo, int hi) {
    UniformIntParams p;
    p.lo = lo;
    p.range = static_cast<uint32_t>(hi) - static_cast<uint32_t>(lo);
    if (p.range == 0) {
        p.bucket = 1;
    } else {
        p.bucket = 0xFFFFFFFFu / (p.range + 1u);
        if (0xFFFFFFFFu % (p.range + 1u) == p.range) p.bucket++;
    }
    return p;
}

// Host generator: full 624-word state.
class HostMT {
public:
    explicit HostMT(uint32_t s) { seed(s); }
    void seed(uint32_t s) {
        mt_[0] = s;
        for (uint32_t i = 1; i < 624; i++) mt_[i] = mtInitStep(mt_[i - 1], i);
        idx_ = 624;
    }
    uint32_t next() {
        if (idx_ >= 624) twist();
        return mtTemper(mt_[idx_++]);
    }

private:
    void twist() {
        for (int i = 0; i < 624; i++) mt_[i] = mtMix(mt_[i], mt_[(i + 1) % 624], mt_[(i + 397) % 624]);
        idx_ = 0;
    }
    uint32_t mt_[624];
    int idx_;
};

// Sampler(seed, lower, upper)
class HostSampler {
public:
    HostSampler(int seed, int lo, int hi) : rng_(static_cast<uint32_t>(seed)), p_(uniformIntParams(lo, hi)) {}
    int next() {
        if (p_.range == 0) return p_.lo;
        for (;;) {
            const uint32_t v = rng_.next() / p_.bucket;
            if (v <= p_.range) return static_cast<int>(static_cast<uint32_t>(p_.lo) + v);
        }
    }

private:
    HostMT rng_;
    UniformIntParams p_;
};

class HostRandomSource {
public:
    explicit HostRandomSource(int seed) : seed_(seed) {}
    HostSampler uniformSampler(int val) { return uniformSampler(0, val - 1); }
    HostSampler uniformSampler(int lo, int hi) { return HostSampler(seed_++, lo, hi); }
    int seed() const { return seed_; }

private:
    int seed_;
};

// Device generator for the first K outputs of MT19937(seed) without the full
// state: output k (k < 227) is temper(mix(s[k], s[k+1], s[k+397])) over the
// initial state s, so only s[0..K] and s[397..397+K) are kept.  Samplers of
// the feature generator draw a handful of values each.
template<int K>
struct DevMTPrefix {
    uint32_t lo[K + 1];
    uint32_t hi[K];
    int pos;

    __device__ void init(uint32_t seed) {
        uint32_t v = seed;
        lo[0] = v;
#pragma unroll 1
        for (uint32_t i = 1; i < 397u + K; i++) {
            v = mtInitStep(v, i);
            if (i <= static_cast<uint32_t>(K)) lo[i] = v;
            if (i >= 397u) hi[i - 397u] = v;
        }
        pos = 0;
    }
    // false when more than K outputs would be needed
    __device__ bool next(uint32_t& out) {
        if (pos >= K) return false;
        out = mtTemper(mtMix(lo[pos], lo[pos + 1], hi[pos]));
        pos++;
        return true;
    }
};

template<int K>
struct DevPrefixSampler {
    DevMTPrefix<K> mt;
    UniformIntParams p;
    __device__ void init(int seed, int lo, int hi) {
        mt.init(static_cast<uint32_t>(seed));
        p = uniformIntParams(lo, hi);
    }
    __device__ bool next(int& out) {
        if (p.range == 0) {
            out = p.lo;
            return true;
        }
        for (;;) {
            uint32_t r;
            if (!mt.next(r)) return false;
            const uint32_t v = r / p.bucket;
            if (v <= p.range) {
                out = static_cast<int>(static_cast<uint32_t>(p.lo) + v);
                return true;
            }
        }
    }
};

}  // namespace cf
