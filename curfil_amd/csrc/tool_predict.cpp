// curfil_predict: predict_main.cpp (predict_main.cpp:59-166) and test()
// (predict.cpp:202-400) over the C ABI.
//
// Same options and positional order as the reference: the forest is read from
// tree files, every `*_colors.png` under folderTesting is classified, per-image
// and average pixel accuracies (with and without void) and the confusion
// matrix are logged, and with folderPrediction the colour, depth, ground truth
// and prediction images are written.  Probability images (writeProbabilityImages)
// are 16-bit PNGs holding round(p * 65535) per label: the reference writes
// p * INT_MAX through a 16-bit depth image, which keeps only the low 16 bits.
// Prediction runs on the GPU; mode 'cpu' is refused.  Images are loaded,
// classified and scored in file order; the PNG files of each image are
// encoded by up to numThreads writers (0: the host's hardware threads, at
// most 16) while the next images are classified, and images are decoded
// 2 x numThreads at a time.
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <filesystem>
#include <future>
#include <sstream>
#include <thread>
#include <vector>

#include "cli.h"
#include "cf_host.h"

namespace {

double average(const std::vector<double>& v) {
    if (v.empty()) return 0.0;
    double s = 0;
    for (double x : v) s += x;
    return s / static_cast<double>(v.size());
}

}  // namespace

extern "C" int cf_tool_predict(int argc, char** argv) {
    cli::Options o;
    o.add("folderPrediction", "folder to output prediction images. leave it empty to suppress writing of prediction images", "");
    o.add("folderTesting", "folder with test images", "", true);
    o.addRepeatable("treeFile", "serialized tree(s) (JSON)", true);
    o.add("histogramBias", "histogram bias", "0");
    o.add("numThreads", "number of threads", "0");
    o.add("mode", "mode: 'cpu' or 'gpu'", "gpu");
    o.add("deviceId", "GPU device id", "0");
    o.addFlag("profile", "profiling", false);
    o.addFlag("useDepthFilling", "whether to do simple depth filling", false);
    o.addFlag("writeProbabilityImages", "whether to write probability PNGs of the prediction", false);
    o.addFlag("useLabelsPrior", "whether to mulitply the trees histograms by labels prior", true);
    o.positional({"folderPrediction", "folderTesting", "treeFile"});
    try {
        if (!o.parse(argc, argv)) return EXIT_FAILURE;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return EXIT_FAILURE;
    }
    cf_ctx* ctx = nullptr;
    try {
        const std::string folderPrediction = o.str("folderPrediction");
        const std::string folderTesting = o.str("folderTesting");
        const double histogramBias = o.real("histogramBias");
        const bool writeProbabilityImages = o.flag("writeProbabilityImages");
        const bool useLabelsPrior = o.flag("useLabelsPrior");
        if (histogramBias < 0.0 || histogramBias >= 1.0) {
            throw std::runtime_error("illegal histogram bias: " + std::to_string(histogramBias));
        }
        if (writeProbabilityImages && folderPrediction.empty()) {
            throw std::runtime_error("specified to write probability images but prediction folder was not set");
        }
        const std::string mode = o.str("mode");
        if (mode != "gpu" && mode != "cpu") throw std::runtime_error("illegal acceleration mode: " + mode);
        if (mode != "gpu") throw std::runtime_error("acceleration mode '" + mode + "' is not available: prediction runs on the GPU");
        cli::info("histogramBias: " + std::to_string(histogramBias));
        cli::info(std::string("writeProbabilityImages: ") + (writeProbabilityImages ? "1" : "0"));
        cli::info(std::string("useLabelsPrior: ") + (useLabelsPrior ? "1" : "0"));
        if (cf_ctx_create(static_cast<int>(o.integer("deviceId")), &ctx) != CF_OK) {
            throw std::runtime_error("cannot open GPU device " + o.str("deviceId"));
        }

        // RandomForestImage(treeFiles, deviceIds, mode, histogramBias, useLabelsPrior)
        const std::vector<std::string> treeFiles = o.all("treeFile");
        std::vector<const char*> tp;
        for (const auto& f : treeFiles) tp.push_back(f.c_str());
        cf_train_config cfg{};
        uint8_t palette[32 * 3];
        cf_forest* forest = nullptr;
        cli::check(ctx, cf_forest_import(ctx, tp.data(), static_cast<int>(tp.size()), &cfg, palette, &forest));
        cli::check(ctx, cf_forest_normalize(forest, histogramBias, useLabelsPrior));
        const int C = cf_forest_num_classes(forest);
        cf_label_map* map = nullptr;
        cli::check(nullptr, cf_label_map_create(&map));
        cli::info("label/color map:");
        for (int l = 0; l < C; l++) {
            int id = -1;
            cli::check(nullptr, cf_label_map_encode(map, palette + 3 * l, &id));
            cli::info("label: " + std::to_string(l) + ", color: RGB(" + std::to_string(palette[3 * l]) + "," +
                    std::to_string(palette[3 * l + 1]) + "," + std::to_string(palette[3 * l + 2]) + ")");
        }
        bool useDepthFilling = cfg.use_depth_filling != 0;
        if (o.given("useDepthFilling")) {
            const bool opt = o.flag("useDepthFilling");
            if (opt != useDepthFilling) {
                cli::warning(useDepthFilling ? "random forest was trained with depth filling. but prediction is performed without!"
                                             : "random forest was NOT trained with depth filling. but prediction is performed with depth filling!");
            }
            useDepthFilling = opt;
        }
        const bool useCIELab = cfg.use_cielab != 0, useDepthImages = cfg.use_depth_images != 0;
        cli::info(std::string("CIELab: ") + (useCIELab ? "1" : "0"));
        cli::info(std::string("DepthFilling: ") + (useDepthFilling ? "1" : "0"));
        cli::info(std::string("useDepthImages: ") + (useDepthImages ? "1" : "0"));
        std::vector<int> ignored(cfg.ignored_labels, cfg.ignored_labels + cfg.num_ignored);

        // test() (predict.cpp:202-400)
        const std::vector<std::string> files = cli::listImageFilenames(folderTesting);
        if (files.empty()) throw std::runtime_error("found no files in " + folderTesting);
        cli::info("got " + std::to_string(files.size()) + " files for prediction");
        const bool writeImages = !folderPrediction.empty();
        if (!writeImages) cli::warning("no prediction folder given. will not write images");
        else std::filesystem::create_directories(folderPrediction);
        std::vector<double> total(static_cast<size_t>(C) * C, 0.0);
        std::vector<double> accuracies, accuraciesWithoutVoid;
        double predictMs = 0;
        const long long threadsOpt = o.integer("numThreads");
        const size_t writers = threadsOpt > 0 ? static_cast<size_t>(threadsOpt)
                                              : std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency()));
        std::deque<std::future<void>> pending;   // PNG writes in flight, oldest first
        // files are decoded a batch at a time on the host's threads (labels encoded in file order, as one
        // at a time would); a batch of equal-sized images goes to the device as one image set
        const size_t batch = 2 * writers;
        for (size_t b0 = 0; b0 < files.size(); b0 += batch) {
            const size_t nb = std::min(batch, files.size() - b0);
            std::vector<const char*> names(nb);
            for (size_t k = 0; k < nb; k++) names[k] = files[b0 + k].c_str();
            cf_dataset* ds = nullptr;
            cli::check(nullptr, cf_dataset_load(names.data(), static_cast<int>(nb), useDepthImages, useDepthFilling,
                    useCIELab, nullptr, map, static_cast<int>(writers), &ds));
            int W0 = 0, H0 = 0;
            cf_dataset_image_size(ds, 0, &W0, &H0);
            bool uniform = true;
            for (size_t k = 1; k < nb; k++) {
                int w = 0, h = 0;
                cf_dataset_image_size(ds, static_cast<int>(k), &w, &h);
                uniform = uniform && w == W0 && h == H0;
            }
            cf_images* batchImages = nullptr;
            if (uniform) cli::check(ctx, cf_dataset_to_images(ds, ctx, &batchImages));
            for (size_t k = 0; k < nb; k++) {
                const size_t i = b0 + k;
                int W = 0, H = 0;
                cf_dataset_image_size(ds, static_cast<int>(k), &W, &H);
                const size_t px = static_cast<size_t>(W) * H;
                std::vector<uint8_t> rgb(px * 3), gt(px), prediction(px);
                std::vector<uint16_t> depth(px);
                cf_dataset_get(ds, static_cast<int>(k), rgb.data(), depth.data(), gt.data());
                for (size_t p = 0; p < px; p++) {
                    if (gt[p] >= C) {
                        std::ostringstream m;
                        m << "illegal label in ground truth image '" << files[i] << "' at pixel (" << p % W << "," << p / W
                          << "): " << static_cast<int>(gt[p]) << " (numClasses: " << C << ")";
                        throw std::runtime_error(m.str());
                    }
                }
                // an image of a batch of mixed sizes goes to the device alone, at its own size
                cf_images* im = batchImages;
                cf_dataset* single = nullptr;
                int index = static_cast<int>(k);
                if (!uniform) {
                    const char* f = files[i].c_str();
                    cli::check(nullptr, cf_dataset_load(&f, 1, useDepthImages, useDepthFilling, useCIELab, nullptr, map, 1,
                            &single));
                    cli::check(ctx, cf_dataset_to_images(single, ctx, &im));
                    index = 0;
                }
                std::vector<float> probs(px * C);
                const auto t0 = std::chrono::steady_clock::now();
                cli::check(ctx, cf_predict(ctx, forest, im, index, probs.data(), prediction.data()));
                predictMs += cli::seconds(t0) * 1e3;
                if (!uniform) {
                    cf_images_destroy(im);
                    cf_dataset_destroy(single);
                }

                const std::string basepath = folderPrediction + "/" + std::filesystem::path(files[i]).stem().string();
                if (writeImages) {
                    if (pending.size() >= writers) {
                        pending.front().get();   // rethrows a failed write
                        pending.pop_front();
                    }
                    pending.push_back(std::async(std::launch::async,
                            [=, rgb = std::move(rgb), depth = std::move(depth), probs = std::move(probs)]() {
                        cli::check(nullptr, cf_png_write((basepath + ".png").c_str(), W, H, 3, 8, rgb.data()));
                        if (useDepthImages) {
                            cli::check(nullptr, cf_png_write((basepath + "_depth.png").c_str(), W, H, 1, 16, depth.data()));
                        }
                        std::vector<uint8_t> gtRgb(px * 3), prRgb(px * 3);
                        for (size_t p = 0; p < px; p++) {
                            std::memcpy(&gtRgb[3 * p], palette + 3 * gt[p], 3);
                            std::memcpy(&prRgb[3 * p], palette + 3 * prediction[p], 3);
                        }
                        cli::check(nullptr, cf_png_write((basepath + "_ground_truth.png").c_str(), W, H, 3, 8, gtRgb.data()));
                        cli::check(nullptr, cf_png_write((basepath + "_prediction.png").c_str(), W, H, 3, 8, prRgb.data()));
                        if (writeProbabilityImages) {
                            std::vector<uint16_t> pi(px);
                            for (int l = 0; l < C; l++) {
                                bool skip = false;
                                for (int g : ignored) skip = skip || g == l;
                                if (skip) continue;
                                for (size_t p = 0; p < px; p++) {
                                    pi[p] = static_cast<uint16_t>(
                                            std::lround(std::min(1.0f, std::max(0.0f, probs[l * px + p])) * 65535.0));
                                }
                                const std::string pf = basepath + "_label_" + std::to_string(l) + ".png";
                                cli::check(nullptr, cf_png_write(pf.c_str(), W, H, 1, 16, pi.data()));
                            }
                        }
                    }));
                }
                double acc = 0, accWithoutVoid = 0;
                std::vector<double> confusion(static_cast<size_t>(C) * C, 0.0);
                cli::check(nullptr, cf_pixel_accuracy(prediction.data(), gt.data(), px, 1, ignored.data(),
                        static_cast<int>(ignored.size()), nullptr, C, &acc));
                cli::check(nullptr, cf_pixel_accuracy(prediction.data(), gt.data(), px, 0, ignored.data(),
                        static_cast<int>(ignored.size()), confusion.data(), C, &accWithoutVoid));
                for (size_t q = 0; q < total.size(); q++) total[q] += confusion[q];
                accuracies.push_back(acc);
                accuraciesWithoutVoid.push_back(accWithoutVoid);
                std::ostringstream m;
                m << "prediction " << (i + 1) << "/" << files.size() << " (" << files[i] << "): pixel accuracy (without void): "
                  << 100 * acc << " (" << 100 * accWithoutVoid << ")";
                cli::info(m.str());
            }
            if (batchImages) cf_images_destroy(batchImages);
            cf_dataset_destroy(ds);
        }
        while (!pending.empty()) {
            pending.front().get();
            pending.pop_front();
        }
        cli::info(std::to_string(predictMs / static_cast<double>(files.size())) + " ms/image");

        // the confusion matrix as the reference prints it (predict.cpp:404-448)
        std::vector<double> normalized = total;
        cf_confusion_normalize(normalized.data(), C);
        std::ostringstream m;
        m << C << "x" << C << " confusion matrix (y: labels, x: predictions):\n                               ";
        for (int l = 0; l < C; l++) {
            char b[16];
            std::snprintf(b, sizeof(b), "cl %-2d  ", l);
            m << b;
        }
        m << "\n";
        for (int l = 0; l < C; l++) {
            char b[64];
            std::snprintf(b, sizeof(b), " class %-2d RGB(%3d,%3d,%3d) : ", l, palette[3 * l], palette[3 * l + 1],
                    palette[3 * l + 2]);
            m << b;
            for (int p = 0; p < C; p++) {
                std::snprintf(b, sizeof(b), "%7.3f", normalized[static_cast<size_t>(l) * C + p]);
                m << b;
            }
            m << "\n";
        }
        double classAcc = 0, classAccNoVoid = 0;
        cf_confusion_class_accuracy(normalized.data(), C, 1, 1, ignored.data(), static_cast<int>(ignored.size()), &classAcc);
        cf_confusion_class_accuracy(normalized.data(), C, 1, 0, ignored.data(), static_cast<int>(ignored.size()),
                &classAccNoVoid);
        m << "\naverage class accuracy (incl void): " << 100 * classAcc << "\naverage class accuracy (no void): "
          << 100 * classAccNoVoid;
        cli::info(m.str());
        cli::info("pixel accuracy: " + std::to_string(100 * average(accuracies)));
        cli::info("pixel accuracy without void: " + std::to_string(100 * average(accuraciesWithoutVoid)));
        cf_forest_destroy(forest);
        cf_label_map_destroy(map);
        cf_ctx_destroy(ctx);
        cli::info("finished");
        return EXIT_SUCCESS;
    } catch (const std::exception& e) {
        cli::log("ERROR", e.what());
        if (ctx) cf_ctx_destroy(ctx);
        return EXIT_FAILURE;
    }
}
