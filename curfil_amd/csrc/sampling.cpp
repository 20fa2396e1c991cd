// Training-pixel sampling on the host (the step before the device hot path).
#include <algorithm>
#include <cmath>
#include <map>

#include "cf_host.h"

namespace cf {

// random_tree_image.cpp:885-909 over the images `ids`: counts of every label.
// The prior has max(numLabels, #distinct labels) entries; the reference fills
// entries 0 .. (largest label) (its loop bound grows as missing labels are
// inserted into the map), which must fit.
std::vector<uint32_t> labelPrior(const cf_images* images, const std::vector<int>& ids, int numLabels) {
    size_t counts[256] = {0};
    const size_t px = static_cast<size_t>(images->W) * images->H;
    for (int i : ids) {
        if (!images->labelsSet[i]) throw Error(CF_ERR_ARG, "labels of image " + std::to_string(i) + " not set");
        const uint8_t* l = images->labels.data() + i * px;
        for (size_t p = 0; p < px; p++) counts[l[p]]++;
    }
    int distinct = 0, maxLabel = -1;
    for (int v = 0; v < 256; v++) {
        if (counts[v]) {
            distinct++;
            maxLabel = v;
        }
    }
    std::vector<uint32_t> prior(std::max(numLabels, distinct), 0u);
    if (maxLabel >= static_cast<int>(prior.size())) {
        throw Error(CF_ERR_ARG, "label " + std::to_string(maxLabel) + " does not fit " + std::to_string(prior.size()) +
                " classes");
    }
    for (int l = 0; l <= maxLabel; l++) prior[l] = static_cast<uint32_t>(counts[l]);
    return prior;
}

std::vector<uint32_t> labelPrior(const cf_images* images, int numLabels) {
    std::vector<int> ids(images->count);
    for (int i = 0; i < images->count; i++) ids[i] = i;
    return labelPrior(images, ids, numLabels);
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

// RandomForestImage::train's image subsampling (random_forest_image.cpp:161-171):
// with maxImages > 0 and more images than that, a reservoir of maxImages
// images drawn with the tree's RandomSource; else every image in order.
std::vector<int> sampleImages(int count, int maxImages, HostRandomSource& rs) {
    std::vector<int> ids;
    if (maxImages > 0 && count > maxImages) {
        HostSampler sampler = rs.uniformSampler(0, 10 * count);
        size_t seen = 0;
        ids.reserve(maxImages);
        for (int i = 0; i < count; i++) {
            // ReservoirSampler::sample (random_tree.h:1089-1109)
            if (ids.size() < static_cast<size_t>(maxImages)) {
                ids.push_back(i);
            } else {
                const size_t r = static_cast<size_t>(sampler.next()) % (seen + 1);
                if (r < static_cast<size_t>(maxImages)) ids[r] = i;
            }
            seen++;
        }
    } else {
        ids.resize(count);
        for (int i = 0; i < count; i++) ids[i] = i;
    }
    return ids;
}

namespace {

bool shouldIgnore(const cf_train_config& cfg, int label) {
    for (int i = 0; i < cfg.num_ignored; i++) {
        if (cfg.ignored_labels[i] == label) return true;
    }
    return false;
}

// the root sort of RandomTreeImage::train (random_tree_image.cpp:776-802): std::sort
// in blocks of 200000 by (image, label, y/10, x/10), image index for the image address
void sortRoots(std::vector<cf_sample>& all) {
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
}

}  // namespace

std::vector<cf_sample> samplePixelUniform(const cf_images* images, const std::vector<int>& ids,
        const cf_train_config& cfg, HostRandomSource& rs, size_t subsampleCount) {
    // subsampleTrainingDataPixelUniform (random_tree_image.cpp:840-883): per sample an image, then
    // fresh x and y samplers drawn until the label is not ignored; depth is not checked
    if (ids.empty()) throw Error(CF_ERR_RUNTIME, "got no label images");
    for (int l = 0; l < 256; l++) {
        if (!shouldIgnore(cfg, l)) break;
        if (l == 255) throw Error(CF_ERR_ARG, "every label is ignored");
    }
    const int W = images->W, H = images->H;
    const size_t px = static_cast<size_t>(W) * H;
    const uint8_t flip = cfg.horizontal_flipping ? 2 : 0;
    // an image whose pixels are all ignored would make the reference draw forever
    std::vector<char> usable(images->count, 0);
    for (int img : ids) {
        const uint8_t* lab = images->labels.data() + img * px;
        for (size_t p = 0; p < px && !usable[img]; p++) usable[img] = !shouldIgnore(cfg, lab[p]);
    }
    HostSampler rImage = rs.uniformSampler(static_cast<int>(ids.size()));
    std::vector<cf_sample> all;
    const size_t n = subsampleCount * ids.size();
    all.reserve(n);
    for (size_t k = 0; k < n; k++) {
        const int img = ids[rImage.next()];
        if (!images->depthSet[img]) throw Error(CF_ERR_ARG, "depth of image " + std::to_string(img) + " not set");
        const uint8_t* lab = images->labels.data() + img * px;
        const int32_t* dep = images->depthHost.data() + img * px;
        HostSampler rx = rs.uniformSampler(W);
        HostSampler ry = rs.uniformSampler(H);
        if (!usable[img]) throw Error(CF_ERR_RUNTIME, "image " + std::to_string(img) + " has no pixel that is not ignored");
        for (;;) {
            const int x = rx.next(), y = ry.next();
            const size_t p = static_cast<size_t>(y) * W + x;
            const int label = lab[p];
            if (shouldIgnore(cfg, label)) continue;
            cf_sample s{};
            s.x = static_cast<uint16_t>(x);
            s.y = static_cast<uint16_t>(y);
            s.image = static_cast<uint32_t>(img);
            s.depth_mm = dep[p] > 0 ? dep[p] : -1;   // PixelInstance depth; Depth::INVALID is -1 (image.cpp:50)
            s.label = static_cast<uint8_t>(label);
            s.flip = flip;
            all.push_back(s);
            break;
        }
    }
    sortRoots(all);
    return all;
}

std::vector<cf_sample> sampleClassUniform(const cf_images* images, const std::vector<int>& ids,
        const std::vector<uint32_t>& prior, const cf_train_config& cfg, HostRandomSource& rs,
        size_t subsampleCount) {
    if (ids.empty()) throw Error(CF_ERR_RUNTIME, "got no label images");
    const size_t numLabelsAll = prior.size();
    bool ignored[256] = {false};
    for (int i = 0; i < cfg.num_ignored; i++) {
        const int l = cfg.ignored_labels[i];
        if (l < 0 || l >= static_cast<int>(numLabelsAll)) throw Error(CF_ERR_RUNTIME, "color to ignore was not found");
        ignored[l] = true;
    }
    const size_t numLabels = numLabelsAll - static_cast<size_t>(cfg.num_ignored);
    const size_t spc = static_cast<size_t>(std::ceil(ids.size() * static_cast<double>(subsampleCount) /
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
        for (int i : ids) {
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
    sortRoots(all);
    return all;
}

}  // namespace cf
