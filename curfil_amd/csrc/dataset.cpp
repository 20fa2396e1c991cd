// Training and test datasets from PNG files: loadImagePair / loadImages
// (image.cpp:818-965), RGBDImage loading (image.cpp:123-184, 396-442), CIELab
// colour (image.cpp:89-96), depth
// filling (image.cpp:279-377), dummy depth (image.cpp:379-394), label images
// and the colour -> label map (image.cpp:660-760), and padding to the largest
// image (image.cpp:449-475, 763-780).
//
// The reference keeps one process-wide colour map that assigns labels in the
// order colours are met; here the map is an object the caller passes along
// (training and test sets share it as they share the global one).  Files are
// decoded in parallel and their colours encoded in file order afterwards, so
// labels do not depend on thread timing (the reference's parallel loader
// encodes in completion order).
#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <thread>

#include "cf_host.h"

struct cf_label_map {
    std::map<std::array<uint8_t, 3>, int> ids;
    std::vector<std::array<uint8_t, 3>> colors;   // by label
    std::string error;
    // LabelImage::encodeColor (image.cpp:680-689)
    int encode(const std::array<uint8_t, 3>& c) {
        auto it = ids.find(c);
        if (it != ids.end()) return it->second;
        const int id = static_cast<int>(colors.size());
        if (id > 255) throw cf::Error(CF_ERR_ARG, "more than 256 label colours");
        ids[c] = id;
        colors.push_back(c);
        return id;
    }
};

namespace cf {

struct LoadedPair {
    std::string colorPath;
    int W = 0, H = 0;
    std::vector<uint8_t> rgb;       // [H][W][3]
    std::vector<uint16_t> depth;    // [H][W] mm
    std::vector<uint8_t> gtRgb;     // ground truth colours [H][W][3]
    std::vector<uint8_t> labels;    // [H][W], after encoding
};

namespace {

std::string replaceSuffix(const std::string& path, const char* with) {
    const size_t pos = path.find("_colors.png");
    if (pos == std::string::npos) throw Error(CF_ERR_ARG, "illegal image filename (no '_colors.png'): " + path);
    return path.substr(0, pos) + with;
}

std::vector<uint8_t> rgb8(const PngImage& im, const std::string& path, const char* what) {
    // loadImage (image.cpp:109-121): three bands are required
    if (im.channels != 3) throw Error(CF_ERR_ARG, std::string("failed to load ") + what + " '" + path +
            "': loading of non-RGB images is not yet supported");
    if (im.bitDepth != 8) throw Error(CF_ERR_UNSUPPORTED, std::string(what) + " '" + path + "': 16-bit colour");
    std::vector<uint8_t> out(im.samples.size());
    for (size_t i = 0; i < out.size(); i++) out[i] = static_cast<uint8_t>(im.samples[i]);
    return out;
}

// CIELab colour planes [3][H][W] from 8-bit RGB: VIGRA's RGB2LabFunctor<double>
// with its default maximum 255 (image.cpp:89-96), i.e. linear RGB -> XYZ
// (D65 matrix), then L = 116 Y^(1/3) - 16 (or kappa * Y below epsilon),
// a = 500 (Xn^(1/3) - Y^(1/3)), b = 200 (Y^(1/3) - Zn^(1/3)) with the white
// point (0.950456, 1, 1.088754); rounded once to float as the reference stores
// its colour planes.
std::vector<float> labPlanes(const std::vector<uint8_t>& rgb, int W, int H) {
    const size_t px = static_cast<size_t>(W) * H;
    std::vector<float> out(3 * px);
    const double third = 1.0 / 3.0, kappa = 24389.0 / 27.0, epsilon = 216.0 / 24389.0;
    for (size_t i = 0; i < px; i++) {
        const double r = rgb[3 * i] / 255.0, g = rgb[3 * i + 1] / 255.0, b = rgb[3 * i + 2] / 255.0;
        const double X = 0.412453 * r + 0.357580 * g + 0.180423 * b;
        const double Y = 0.212671 * r + 0.715160 * g + 0.072169 * b;
        const double Z = 0.019334 * r + 0.119193 * g + 0.950227 * b;
        const double xg = std::pow(X / 0.950456, third);
        const double yg = std::pow(Y, third);
        const double zg = std::pow(Z / 1.088754, third);
        const double L = Y < epsilon ? kappa * Y : 116.0 * yg - 16.0;
        out[i] = static_cast<float>(L);
        out[px + i] = static_cast<float>(500.0 * (xg - yg));
        out[2 * px + i] = static_cast<float>(200.0 * (yg - zg));
    }
    return out;
}

// RGBDImage::fillDepth (image.cpp:279-377): from the right, left, top and bottom in turn
void fillDepth(std::vector<uint16_t>& d, int W, int H) {
    for (int y = 0; y < H; y++) {
        uint16_t* r = &d[static_cast<size_t>(y) * W];
        for (int x = W - 2; x >= 0; x--) if (!r[x]) r[x] = r[x + 1];
        for (int x = 1; x < W; x++) if (!r[x]) r[x] = r[x - 1];
    }
    for (int y = 1; y < H; y++) {
        for (int x = 0; x < W; x++) {
            uint16_t& v = d[static_cast<size_t>(y) * W + x];
            if (!v) v = d[static_cast<size_t>(y - 1) * W + x];
        }
    }
    for (int y = H - 2; y >= 0; y--) {
        for (int x = 0; x < W; x++) {
            uint16_t& v = d[static_cast<size_t>(y) * W + x];
            if (!v) v = d[static_cast<size_t>(y + 1) * W + x];
        }
    }
}

LoadedPair loadPair(const std::string& colorPath, bool useDepthImages, bool useDepthFilling) {
    LoadedPair p;
    p.colorPath = colorPath;
    const PngImage c = readPng(colorPath);
    p.W = c.width;
    p.H = c.height;
    p.rgb = rgb8(c, colorPath, "image");
    const size_t px = static_cast<size_t>(p.W) * p.H;
    if (useDepthImages) {
        const std::string dp = replaceSuffix(colorPath, "_depth.png");
        const PngImage d = readPng(dp);
        // loadDepthImage (image.cpp:396-442)
        if (d.channels != 1 || d.colorType != 0) throw Error(CF_ERR_ARG, "failed to load depth image '" + dp +
                "': invalid depth image format");
        if (d.width != p.W || d.height != p.H) throw Error(CF_ERR_ARG, "failed to load depth image '" + dp +
                "': width or height of color and depth image differ");
        p.depth.resize(px);
        double sum = 0;
        size_t count = 0;
        for (size_t i = 0; i < px; i++) {
            const uint16_t v = d.samples[i];
            if (v > 50000) throw Error(CF_ERR_ARG, "illegal depth value in image " + dp + ": " + std::to_string(v));
            if (v > 0) {
                sum += v / 1000.0;
                count++;
            }
            p.depth[i] = v;
        }
        const double avg = count ? sum / static_cast<double>(count) : 0.0;
        if (avg < 0.01 || avg > 10.0) {
            throw Error(CF_ERR_ARG, "illegal average depth of '" + dp + "': " + std::to_string(avg));
        }
    } else {
        p.depth.assign(px, 1000);   // loadDummyDepthValues (image.cpp:379-394)
    }
    if (useDepthFilling) fillDepth(p.depth, p.W, p.H);
    const std::string gp = replaceSuffix(colorPath, "_ground_truth.png");
    const PngImage g = readPng(gp);
    p.gtRgb = rgb8(g, gp, "label image");
    if (g.width != p.W || g.height != p.H) {
        throw Error(CF_ERR_ARG, "RGB-D image '" + colorPath + "' and label image '" + gp + "' have different sizes");
    }
    return p;
}

// LabelImage(filename) (image.cpp:736-760): colours are encoded column by column
void encodeLabels(LoadedPair& p, cf_label_map* map) {
    p.labels.resize(static_cast<size_t>(p.W) * p.H);
    for (int x = 0; x < p.W; x++) {
        for (int y = 0; y < p.H; y++) {
            const size_t i = static_cast<size_t>(y) * p.W + x;
            p.labels[i] = static_cast<uint8_t>(map->encode({p.gtRgb[3 * i], p.gtRgb[3 * i + 1], p.gtRgb[3 * i + 2]}));
        }
    }
    p.gtRgb.clear();
    p.gtRgb.shrink_to_fit();
}

}  // namespace
}  // namespace cf

struct cf_dataset {
    std::vector<cf::LoadedPair> pairs;
    int W = 0, H = 0;          // the largest width and height (every image is padded to it)
    int paddingLabel = 0;
    bool cielab = false;       // colour planes in CIELab (float) instead of 8-bit RGB
    std::string error;
};

using namespace cf;

namespace {

// calls without a context report through cf_last_error(NULL)
template<class Fn>
int hostCall(std::string* err, Fn&& fn) {
    std::string& slot = hostErrorSlot();
    try {
        fn();
        if (err) err->clear();
        slot.clear();
        return CF_OK;
    } catch (const Error& e) {
        if (err) *err = e.what();
        slot = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        slot = "host allocation failed";
        return CF_ERR_NOMEM;
    } catch (const std::exception& e) {
        if (err) *err = e.what();
        slot = e.what();
        return CF_ERR_RUNTIME;
    }
}

}  // namespace

extern "C" {

int cf_label_map_create(cf_label_map** out) {
    if (!out) return CF_ERR_ARG;
    *out = new cf_label_map();
    return CF_OK;
}

int cf_label_map_destroy(cf_label_map* map) {
    delete map;
    return CF_OK;
}

int cf_label_map_encode(cf_label_map* map, const uint8_t* rgb, int* label) {
    if (!map || !rgb || !label) return CF_ERR_ARG;
    return hostCall(&map->error, [&] { *label = map->encode({rgb[0], rgb[1], rgb[2]}); });
}

int cf_label_map_find(const cf_label_map* map, const uint8_t* rgb) {
    if (!map || !rgb) return CF_ERR_ARG;
    auto it = map->ids.find({rgb[0], rgb[1], rgb[2]});
    return it == map->ids.end() ? -1 : it->second;
}

int cf_label_map_size(const cf_label_map* map) { return map ? static_cast<int>(map->colors.size()) : CF_ERR_ARG; }

int cf_label_map_colors(const cf_label_map* map, uint8_t* rgb) {
    if (!map || !rgb) return CF_ERR_ARG;
    for (size_t i = 0; i < map->colors.size(); i++) std::memcpy(rgb + 3 * i, map->colors[i].data(), 3);
    return static_cast<int>(map->colors.size());
}

int cf_dataset_load(const char* const* colorFiles, int count, int useDepthImages, int useDepthFilling, int useCIELab,
        const uint8_t* paddingColor, cf_label_map* map, int threads, cf_dataset** out) {
    if (!out || !map) return CF_ERR_ARG;
    *out = nullptr;
    auto ds = std::make_unique<cf_dataset>();
    ds->cielab = useCIELab != 0;
    const int st = hostCall(nullptr, [&] {
        if (!colorFiles || count <= 0) throw Error(CF_ERR_ARG, "no images to load");
        for (int i = 0; i < count; i++) {
            if (!colorFiles[i]) throw Error(CF_ERR_ARG, "null file name");
        }
        // getPaddingLabel (image.cpp:859-867): the first ignored colour, else black, encoded first;
        // without a padding colour (single test images, predict.cpp:254) nothing is encoded and
        // images must not need padding
        ds->paddingLabel = paddingColor ? map->encode({paddingColor[0], paddingColor[1], paddingColor[2]}) : -1;
        ds->pairs.resize(count);
        std::vector<std::string> errors(count);
        const int nt = std::max(1, std::min(threads > 0 ? threads : 8, count));
        std::vector<std::thread> pool;
        std::mutex mu;
        int next = 0;
        for (int t = 0; t < nt; t++) {
            pool.emplace_back([&] {
                for (;;) {
                    int i;
                    {
                        std::lock_guard<std::mutex> g(mu);
                        i = next++;
                    }
                    if (i >= count) return;
                    try {
                        ds->pairs[i] = loadPair(colorFiles[i], useDepthImages != 0, useDepthFilling != 0);
                    } catch (const std::exception& e) {
                        errors[i] = e.what();
                    }
                }
            });
        }
        for (auto& th : pool) th.join();
        for (const auto& e : errors) {
            if (!e.empty()) throw Error(CF_ERR_ARG, e);
        }
        for (auto& p : ds->pairs) {
            encodeLabels(p, map);
            ds->W = std::max(ds->W, p.W);
            ds->H = std::max(ds->H, p.H);
        }
    });
    if (st == CF_OK) *out = ds.release();
    return st;
}

int cf_dataset_destroy(cf_dataset* ds) {
    delete ds;
    return CF_OK;
}

int cf_dataset_info(const cf_dataset* ds, int* count, int* width, int* height, int* padding_label) {
    if (!ds) return CF_ERR_ARG;
    if (count) *count = static_cast<int>(ds->pairs.size());
    if (width) *width = ds->W;
    if (height) *height = ds->H;
    if (padding_label) *padding_label = ds->paddingLabel;
    return CF_OK;
}

int cf_dataset_image_size(const cf_dataset* ds, int index, int* width, int* height) {
    if (!ds || index < 0 || index >= static_cast<int>(ds->pairs.size())) return CF_ERR_ARG;
    if (width) *width = ds->pairs[index].W;
    if (height) *height = ds->pairs[index].H;
    return CF_OK;
}

int cf_dataset_get(const cf_dataset* ds, int index, uint8_t* rgb, uint16_t* depth, uint8_t* labels) {
    if (!ds || index < 0 || index >= static_cast<int>(ds->pairs.size())) return CF_ERR_ARG;
    const LoadedPair& p = ds->pairs[index];
    const size_t px = static_cast<size_t>(p.W) * p.H;
    if (rgb) std::memcpy(rgb, p.rgb.data(), px * 3);
    if (depth) std::memcpy(depth, p.depth.data(), px * sizeof(uint16_t));
    if (labels) std::memcpy(labels, p.labels.data(), px);
    return CF_OK;
}

int cf_dataset_get_color(const cf_dataset* ds, int index, float* color) {
    if (!ds || !color || index < 0 || index >= static_cast<int>(ds->pairs.size())) return CF_ERR_ARG;
    const LoadedPair& p = ds->pairs[index];
    const size_t px = static_cast<size_t>(p.W) * p.H;
    if (ds->cielab) {
        const std::vector<float> lab = labPlanes(p.rgb, p.W, p.H);
        std::memcpy(color, lab.data(), lab.size() * sizeof(float));
    } else {
        for (size_t i = 0; i < px; i++) {
            for (int c = 0; c < 3; c++) color[c * px + i] = p.rgb[3 * i + c];
        }
    }
    return CF_OK;
}

int cf_dataset_to_images(const cf_dataset* ds, cf_ctx* ctx, cf_images** out) {
    if (!ds || !ctx || !out) return CF_ERR_ARG;
    *out = nullptr;
    cf_images* im = nullptr;
    int st = cf_images_create(ctx, static_cast<int>(ds->pairs.size()), ds->W, ds->H, &im);
    if (st != CF_OK) return st;
    std::vector<uint8_t> lab(static_cast<size_t>(ds->W) * ds->H);
    for (size_t i = 0; i < ds->pairs.size() && st == CF_OK; i++) {
        const LoadedPair& p = ds->pairs[i];
        if (ds->cielab) {
            const std::vector<float> lab = labPlanes(p.rgb, p.W, p.H);
            const std::vector<int32_t> d32(p.depth.begin(), p.depth.end());
            st = cf_images_set_rgbd_f32_padded(im, static_cast<int>(i), lab.data(), d32.data(), p.W, p.H);
        } else {
            st = cf_images_set_rgbd_u8_padded(im, static_cast<int>(i), p.rgb.data(), p.depth.data(), p.W, p.H);
        }
        if (st != CF_OK) break;
        // LabelImage::resizeImage (image.cpp:763-780): padding label outside the image
        if ((p.W != ds->W || p.H != ds->H) && ds->paddingLabel < 0) {
            st = CF_ERR_ARG;
            ctx->error = "images of different sizes need a padding colour";
            break;
        }
        std::fill(lab.begin(), lab.end(), static_cast<uint8_t>(std::max(ds->paddingLabel, 0)));
        for (int y = 0; y < p.H; y++) {
            std::memcpy(&lab[static_cast<size_t>(y) * ds->W], &p.labels[static_cast<size_t>(y) * p.W], p.W);
        }
        st = cf_images_set_labels(im, static_cast<int>(i), lab.data());
    }
    if (st != CF_OK) {
        const std::string msg = ctx->error;
        cf_images_destroy(im);
        ctx->error = msg;
        return st;
    }
    *out = im;
    return CF_OK;
}

int cf_png_read_info(const char* path, int* width, int* height, int* channels, int* bit_depth) {
    if (!path) return CF_ERR_ARG;
    return hostCall(nullptr, [&] {
        const PngImage im = readPng(path);
        if (width) *width = im.width;
        if (height) *height = im.height;
        if (channels) *channels = im.channels;
        if (bit_depth) *bit_depth = im.bitDepth;
    });
}

int cf_png_read(const char* path, uint16_t* samples, size_t capacity) {
    if (!path || !samples) return CF_ERR_ARG;
    return hostCall(nullptr, [&] {
        const PngImage im = readPng(path);
        if (im.samples.size() > capacity) throw Error(CF_ERR_ARG, "output buffer too small");
        std::memcpy(samples, im.samples.data(), im.samples.size() * sizeof(uint16_t));
    });
}

int cf_png_write(const char* path, int width, int height, int channels, int bit_depth, const void* pixels) {
    if (!path || !pixels) return CF_ERR_ARG;
    return hostCall(nullptr, [&] { writePng(path, width, height, channels, bit_depth, pixels); });
}

}  // extern "C"
