// curfil_train: train_main.cpp (train_main.cpp:40-187) over the C ABI.
//
// Same options, positional order and defaults as the reference.  Training
// runs on the GPU (mode 'gpu'); 'cpu' and 'compare' need the reference's CPU
// evaluation and are refused.  imageCacheSize, numThreads, profile and
// trainTreesInParallel are accepted for compatibility: images stay resident
// in device memory (no cache), trees of a level are trained together, and
// host threads are sized by the library.
#include <chrono>
#include <cstring>
#include <filesystem>
#include <map>
#include <sstream>
#include <vector>

#include "cli.h"
#include "cf_host.h"

extern "C" int cf_tool_train(int argc, char** argv) {
    cli::Options o;
    o.add("folderTraining", "folder with training images", "", true);
    o.add("trees", "number of trees to train", "", true);
    o.add("samplesPerImage", "samples per image", "", true);
    o.add("featureCount", "feature count", "", true);
    o.add("minSampleCount", "min samples count", "", true);
    o.add("maxDepth", "maximum tree depth", "", true);
    o.add("boxRadius", "box radius", "", true);
    o.add("regionSize", "region size", "", true);
    o.add("numThresholds", "number of thresholds to evaluate", "", true);
    o.add("outputFolder", "folder to output predictions and trees", "");
    o.add("numThreads", "number of threads", "0");
    o.addFlag("useCIELab", "convert images to CIElab color space", true);
    o.addFlag("useDepthFilling", "whether to do simple depth filling", false);
    o.add("deviceId", "GPU device id", "0");
    o.add("subsamplingType", "subsampling type: 'pixelUniform' or 'classUniform'", "classUniform");
    o.add("maxImages", "maximum number of images to load for training. set to 0 if all images should be loaded", "0");
    o.add("imageCacheSize", "image cache size on GPU in MB. 0 means automatic adjustment", "0");
    o.add("mode", "mode: 'gpu' (default), 'cpu' or 'compare'", "gpu");
    o.addFlag("profile", "profiling", false);
    o.add("randomSeed", "random seed", "4711");
    o.addRepeatable("ignoreColor", "do not sample pixels of this color. format: R,G,B where 0 <= R,G,B <= 255");
    o.addFlag("verboseTree", "whether to write verbose tree include profiling and debugging information", false);
    o.addFlag("trainTreesInParallel", "whether to train multiple trees sequentially (default) or in parallel", false);
    o.addFlag("useDepthImages", "whether to use depth images", true);
    // the reference gives this option the implicit value false (train_main.cpp:119-120): a bare
    // --horizontalFlipping leaves flipping off; --horizontalFlipping=1 (or 'true') turns it on
    o.addFlag("horizontalFlipping", "whether to horizontally flip features", false, false);
    o.positional({"folderTraining", "trees", "samplesPerImage", "featureCount", "minSampleCount", "maxDepth",
            "boxRadius", "regionSize", "numThresholds", "outputFolder"});
    cf_ctx* ctx = nullptr;
    try {
        if (!o.parse(argc, argv)) return EXIT_FAILURE;
    } catch (const std::exception& e) {
        std::fprintf(stderr, "%s\n", e.what());
        return EXIT_FAILURE;
    }
    try {
        const bool horizontalFlipping = o.flag("horizontalFlipping");
        const std::string mode = o.str("mode");
        if (mode != "gpu" && mode != "cpu" && mode != "compare") throw std::runtime_error("illegal acceleration mode: " + mode);
        if (mode != "gpu") throw std::runtime_error("acceleration mode '" + mode + "' is not available: training runs on the GPU");
        const std::string subsampling = o.str("subsamplingType");
        if (subsampling != "classUniform" && subsampling != "pixelUniform") {
            throw std::runtime_error("unknown subsamplingType: " + subsampling);
        }
        const std::string folderTraining = o.str("folderTraining");
        const std::string outputFolder = o.str("outputFolder");
        const bool useCIELab = o.flag("useCIELab"), useDepthFilling = o.flag("useDepthFilling");
        const bool useDepthImages = o.flag("useDepthImages");
        const std::vector<std::string> ignoredColors = o.all("ignoreColor");
        cli::info("acceleration mode: " + mode);
        cli::info(std::string("CIELab: ") + (useCIELab ? "1" : "0"));
        cli::info(std::string("DepthFilling: ") + (useDepthFilling ? "1" : "0"));

        if (cf_ctx_create(static_cast<int>(o.integer("deviceId")), &ctx) != CF_OK) {
            throw std::runtime_error("cannot open GPU device " + o.str("deviceId"));
        }
        if (o.flag("profile")) cli::check(ctx, cf_ctx_set_profiling(ctx, 1));

        // loadImages (image.cpp:869-965)
        const auto t0 = std::chrono::steady_clock::now();
        const std::vector<std::string> files = cli::listImageFilenames(folderTraining);
        if (files.empty()) throw std::runtime_error("found no files in " + folderTraining);
        cli::info("going to load " + std::to_string(files.size()) + " images from " + folderTraining);
        uint8_t padding[3] = {0, 0, 0};
        if (!ignoredColors.empty()) cli::parseColor(ignoredColors[0], padding);
        cf_label_map* map = nullptr;
        cli::check(nullptr, cf_label_map_create(&map));
        std::vector<const char*> paths;
        for (const auto& f : files) paths.push_back(f.c_str());
        cf_dataset* ds = nullptr;
        cli::check(nullptr, cf_dataset_load(paths.data(), static_cast<int>(paths.size()), useDepthImages,
                useDepthFilling, useCIELab, padding, map, 0, &ds));
        cf_images* images = nullptr;
        cli::check(ctx, cf_dataset_to_images(ds, ctx, &images));
        int count = 0, W = 0, H = 0;
        cf_dataset_info(ds, &count, &W, &H, nullptr);
        const int numLabels = cf_label_map_size(map);
        cli::info("finished loading " + std::to_string(count) + " images (" + std::to_string(W) + "x" +
                std::to_string(H) + ", " + std::to_string(numLabels) + " labels) in " +
                std::to_string(cli::seconds(t0)) + " s");

        cf_train_config cfg{};
        cfg.random_seed = static_cast<int32_t>(o.integer("randomSeed"));
        cfg.samples_per_image = static_cast<uint32_t>(o.integer("samplesPerImage"));
        cfg.feature_count = static_cast<uint32_t>(o.integer("featureCount"));
        cfg.min_sample_count = static_cast<uint32_t>(o.integer("minSampleCount"));
        cfg.max_depth = static_cast<int32_t>(o.integer("maxDepth"));
        cfg.box_radius = static_cast<uint16_t>(o.integer("boxRadius"));
        cfg.region_size = static_cast<uint16_t>(o.integer("regionSize"));
        cfg.thresholds = static_cast<uint16_t>(o.integer("numThresholds"));
        cfg.num_threads = static_cast<int32_t>(o.integer("numThreads"));
        cfg.max_images = static_cast<int32_t>(o.integer("maxImages"));
        cfg.use_depth_images = useDepthImages;
        cfg.horizontal_flipping = horizontalFlipping;
        cfg.subsampling_pixel_uniform = subsampling == "pixelUniform";
        cfg.use_depth_filling = useDepthFilling;
        cfg.use_cielab = useCIELab;
        if (ignoredColors.size() > 8) throw std::runtime_error("at most 8 colours can be ignored");
        for (const auto& c : ignoredColors) {
            uint8_t rgb[3];
            cli::parseColor(c, rgb);
            const int label = cf_label_map_find(map, rgb);
            if (label < 0) throw std::runtime_error("unknown color to ignore: " + c);   // random_tree_image.cpp:952-960
            cfg.ignored_labels[cfg.num_ignored++] = label;
        }
        const int trees = static_cast<int>(o.integer("trees"));
        cli::info("trees: " + std::to_string(trees));

        // train (train.cpp:84-107)
        const auto t1 = std::chrono::steady_clock::now();
        cf_forest* forest = nullptr;
        cli::check(ctx, cf_forest_train(ctx, images, &cfg, trees, numLabels, CF_MODE_BATCHED, &forest));
        const double trainSeconds = cli::seconds(t1);
        double evaluations = 0, deviceSeconds = 0;
        cf_forest_train_stats(forest, &evaluations, &deviceSeconds);
        std::ostringstream msg;
        msg << "training took " << trainSeconds << " s (" << trainSeconds / 60.0 << " min), " << evaluations / trainSeconds
            << " feature-threshold evaluations/s";
        cli::info(msg.str());
        // the forest summary and feature counts (RandomForestImage::countFeatures)
        std::map<std::string, long> featureCounts;
        for (int t = 0; t < cf_forest_num_trees(forest); t++) {
            const int n = cf_forest_tree_size(forest, t);
            std::vector<cf_tree_node> nodes(n);
            std::vector<uint32_t> hist(static_cast<size_t>(n) * cf_forest_num_classes(forest));
            cli::check(ctx, cf_forest_get_tree(forest, t, nodes.data(), hist.data()));
            int leaves = 0, depth = 0;
            for (const auto& nd : nodes) {
                leaves += nd.is_leaf;
                depth = std::max(depth, nd.level);
                if (!nd.is_leaf) featureCounts[nd.feature[0] == 1 ? "color" : "depth"]++;
            }
            cli::info("tree " + std::to_string(t) + ": " + std::to_string(n) + " nodes, " + std::to_string(leaves) +
                    " leaves, depth " + std::to_string(depth));
        }
        for (const auto& fc : featureCounts) cli::info("feature " + fc.first + ": " + std::to_string(fc.second));
        if (o.flag("profile")) {
            // per-phase device time, as the reference's PROFILING lines (utils.h:159-203)
            double p[12] = {0};
            const int n = cf_forest_profile(forest, p, 12);
            const char* names[12] = {"evaluations", "seconds", "response kernel ms", "response kernel launches",
                    "responses", "feature generation ms", "scores ms", "partition ms", "levels", "replays", "nodes",
                    "node histograms ms"};
            for (int i = 0; i < n && i < 12; i++) {
                std::ostringstream m;
                m << "PROFILING('" << names[i] << "'): " << p[i];
                cli::info(m.str());
            }
        }

        // RandomTreeExport::writeJSON: <outputFolder>/tree<N>.json.gz
        if (!outputFolder.empty()) {
            std::filesystem::create_directories(outputFolder);
            std::vector<uint8_t> palette(3 * static_cast<size_t>(std::max(numLabels, 1)));
            cf_label_map_colors(map, palette.data());
            const std::string absTraining = std::filesystem::absolute(folderTraining).string();
            for (int t = 0; t < cf_forest_num_trees(forest); t++) {
                const std::string f = outputFolder + "/tree" + std::to_string(t) + ".json.gz";
                cli::check(ctx, cf_forest_export_tree(forest, t, f.c_str(), &cfg, palette.data(), absTraining.c_str()));
                cli::info("wrote " + f);
            }
        } else {
            cli::warning("no output folder given. skipping JSON export");
        }
        cf_forest_destroy(forest);
        cf_images_destroy(images);
        cf_dataset_destroy(ds);
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
