// Training-pixel sampling on the host (the step before the device hot path).
#include <algorithm>
#include <cmath>
#include <map>

#include "cf_host.h"

namespace cf {

// random_tree_image.cpp:885-909: counts of every label over all pixels; the
// prior has max(numLabels, #distinct labels) entries, filled for labels
// 0 .. #distinct-1.
std::vector<uint32_t> labelPrior(const cf_images* images, int numLabels) {
    std::map<int, size_t> m;
    const size_t px = static_cast<size_t>(images->W) * images->H;
    for (int i = 0; i < images->count; i++) {
        if (!images->labelsSet[i]) throw Error(CF_ERR_ARG, "labels of image " + std::to_string(i) + " not set");
        const uint8_t* l = images->labels.data() + i * px;
        size_t counts[256] = {0};
        for (size_t p = 0; p < px; p++) counts[l[p]]++;
        for (int v = 0; v < 256; v++) {
            if (counts[v]) m[v] += counts[v];
        }
    }
    std::vector<uint32_t> prior(std::max<size_t>(static_cast<size_t>(numLabels), m.size()), 0u);
    for (size_t l = 0; l < m.size(); l++) {
        auto it = m.find(static_cast<int>(l));
        prior[l] = it == m.end() ? 0u : static_cast<uint32_t>(it->second);
    }
    return prior;
}

namespace {

struct Reservoir {
    size_t count = 0, capacity = 0;
    std::vector<cf_sample> items;
    explicit Reservoir(size_t n) : capacity(n) { items.reserve(n); }
    // ReservoirSampler::sample, random_tree.h:1089-1109
    void offer(HostSampler& s, const cf_sample& v) {
        if (capacity == 0) throw Error(CF_ERR_RUNTIME, "no samples to sample");
        if (items.size() < capacity) {
            items.push_back(v);
        } else {
            const size_t r = static_cast<size_t>(s.next()) % (count + 1);
            if (r < capacity) items[r] = v;
        }
        count++;
    }
};

}  // namespace

std::vector<cf_sample> sampleClassUniform(const cf_images* images, const std::vector<uint32_t>& prior,
        const cf_train_config& cfg, HostRandomSource& rs, size_t subsampleCount) {
    if (images->count == 0) throw Error(CF_ERR_RUNTIME, "got no label images");
    const size_t numLabelsAll = prior.size();
    bool ignored[256] = {false};
    for (int i = 0; i < cfg.num_ignored; i++) {
        const int l = cfg.ignored_labels[i];
        if (l < 0 || l >= static_cast<int>(numLabelsAll)) throw Error(CF_ERR_RUNTIME, "color to ignore was not found");
        ignored[l] = true;
    }
    const size_t numLabels = numLabelsAll - static_cast<size_t>(cfg.num_ignored);
    const size_t spc = static_cast<size_t>(std::ceil(images->count * static_cast<double>(subsampleCount) /
            static_cast<double>(numLabels)));
    const int randomSeed = rs.uniformSampler(0xFFFFFF).next();
    const uint8_t flip = cfg.horizontal_flipping ? 2 : 0;

    std::vector<Reservoir> chunk;
    chunk.reserve(numLabelsAll);
    for (size_t l = 0; l < numLabelsAll; l++) chunk.emplace_back(spc);
    {
        HostRandomSource crs(randomSeed);
        HostSampler sampler = crs.uniformSampler(0x7FFFFFFF);
        const int W = images->W, H = images->H;
        const size_t px = static_cast<size_t>(W) * H;
        for (int i = 0; i < images->count; i++) {
            if (!images->depthSet[i]) throw Error(CF_ERR_ARG, "depth of image " + std::to_string(i) + " not set");
            const uint8_t* lab = images->labels.data() + i * px;
            const int32_t* dep = images->depthHost.data() + i * px;
            for (int y = 0; y < H; y++) {
                for (int x = 0; x < W; x++) {
                    const size_t p = static_cast<size_t>(y) * W + x;
                    const int label = lab[p];
                    if (ignored[label]) continue;
                    // PixelInstance depth (random_tree_image.h:148-164) recovers the raw
                    // depth; invalid pixels (depth <= 0) are skipped.
                    const int32_t d = dep[p];
                    if (!(d > 0)) continue;
                    if (static_cast<size_t>(label) >= numLabelsAll) {
                        throw Error(CF_ERR_RUNTIME, "label outside the class prior");
                    }
                    cf_sample s{};
                    s.x = static_cast<uint16_t>(x);
                    s.y = static_cast<uint16_t>(y);
                    s.image = static_cast<uint32_t>(i);
                    s.depth_mm = d;
                    s.label = static_cast<uint8_t>(label);
                    s.flip = flip;
                    chunk[label].offer(sampler, s);
                }
            }
        }
    }
    std::vector<cf_sample> all;
    all.reserve(spc * numLabels);
    HostSampler sampler = rs.uniformSampler(0x7FFFFFFF);
    for (size_t l = 0; l < numLabelsAll; l++) {
        Reservoir merged(spc);
        if (!ignored[l]) {
            for (const cf_sample& s : chunk[l].items) merged.offer(sampler, s);
        }
        all.insert(all.begin(), merged.items.begin(), merged.items.end());
    }
    if (all.size() != spc * numLabels) throw Error(CF_ERR_RUNTIME, "failed to sample enough pixels");
    const size_t BLOCK = 200000;
    for (size_t b = 0; b < all.size(); b += BLOCK) {
        const size_t e = std::min(all.size(), b + BLOCK);
        std::sort(all.begin() + b, all.begin() + e, [](const cf_sample& a, const cf_sample& c) {
            if (a.image != c.image) return a.image < c.image;
            if (a.label != c.label) return a.label < c.label;
            if (a.y / 10 != c.y / 10) return a.y / 10 < c.y / 10;
            if (a.x / 10 != c.x / 10) return a.x / 10 < c.x / 10;
            return false;
        });
    }
    return all;
}

}  // namespace cf
