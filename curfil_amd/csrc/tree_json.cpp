// Per-tree JSON files, gzip-compressed when the name ends in ".gz": the
// format RandomTreeExport writes and RandomTreeImport reads
// (export.cpp:180-313, import.cpp:42-265).
//
// Layout (every value written as a JSON string, as Boost.PropertyTree does):
//   header keys (hostname, buildType, version, date, randomSeed,
//   folderTraining, numCores, numPhysicalCPUs, processorModels[],
//   accelerationMode, deviceId[], samplesPerImage, featureCount, thresholds,
//   boxRadius, regionSize, maxDepth, numThreads, minSampleCount,
//   maxSamplesPerBatch, maxImages, imageCacheSize, subsamplingType, useCIELab,
//   useDepthFilling, useDepthImages, ignoredColors[]),
//   classLabelPriorDistribution {"<label>": count}, and
//   tree: node {id, level, samples, leaf, histogram {"R,G,B (label)": count},
//               [split {threshold, score, featureId, feature {type "color" |
//                "depth", [channel1, channel2], offset1 [x, y], region1 [x, y],
//                offset2 [x, y], region2 [x, y]}}, left {node}, right {node}]}.
// Floats are written with enough digits to read back to the same value.
#include <unistd.h>
#include <zlib.h>

#include <cctype>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <deque>
#include <memory>
#include <sstream>
#include <string>
#include <utility>
#include <vector>

#include "cf_host.h"

namespace cf {
namespace {

// ---- a small JSON document model -------------------------------------------
struct JVal {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    std::string s;      // Str; also the text of Num
    bool b = false;
    std::vector<JVal> arr;
    std::vector<std::pair<std::string, JVal>> obj;

    const JVal* find(const std::string& key) const {
        for (const auto& kv : obj) {
            if (kv.first == key) return &kv.second;
        }
        return nullptr;
    }
    const JVal& at(const std::string& key) const {
        const JVal* v = find(key);
        if (!v) throw Error(CF_ERR_ARG, "tree file: missing key '" + key + "'");
        return *v;
    }
    // scalar text (strings and numbers alike, as Boost.PropertyTree reads them)
    const std::string& text() const {
        if (kind != Str && kind != Num) throw Error(CF_ERR_ARG, "tree file: expected a value");
        return s;
    }
    long long asInt() const {
        const std::string& t = text();
        char* end = nullptr;
        const long long v = std::strtoll(t.c_str(), &end, 10);
        if (end == t.c_str() || *end != '\0') throw Error(CF_ERR_ARG, "tree file: not an integer: '" + t + "'");
        return v;
    }
    double asDouble() const {
        const std::string& t = text();
        char* end = nullptr;
        const double v = std::strtod(t.c_str(), &end);
        if (end == t.c_str() || *end != '\0') throw Error(CF_ERR_ARG, "tree file: not a number: '" + t + "'");
        return v;
    }
    bool asBool() const {
        if (kind == Bool) return b;
        const std::string& t = text();
        if (t == "true" || t == "1") return true;
        if (t == "false" || t == "0") return false;
        throw Error(CF_ERR_ARG, "tree file: not a boolean: '" + t + "'");
    }
};

class Parser {
public:
    explicit Parser(const std::string& t) : t_(t) {}
    JVal parse() {
        JVal v = value();
        ws();
        if (i_ != t_.size()) fail("trailing characters");
        return v;
    }

private:
    const std::string& t_;
    size_t i_ = 0;
    int depth_ = 0;   // nesting; a tree of depth d nests about 2 d levels
    static constexpr int kMaxDepth = 4096;

    [[noreturn]] void fail(const char* what) {
        throw Error(CF_ERR_ARG, std::string("tree file: JSON ") + what + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < t_.size() && std::isspace(static_cast<unsigned char>(t_[i_]))) i_++;
    }
    char peek() {
        ws();
        if (i_ >= t_.size()) fail("ends early");
        return t_[i_];
    }
    void expect(char c) {
        if (peek() != c) fail("syntax error");
        i_++;
    }
    std::string str() {
        expect('"');
        std::string out;
        while (true) {
            if (i_ >= t_.size()) fail("unterminated string");
            const char c = t_[i_++];
            if (c == '"') break;
            if (c != '\\') {
                out.push_back(c);
                continue;
            }
            if (i_ >= t_.size()) fail("bad escape");
            const char e = t_[i_++];
            switch (e) {
                case '"': out.push_back('"'); break;
                case '\\': out.push_back('\\'); break;
                case '/': out.push_back('/'); break;
                case 'b': out.push_back('\b'); break;
                case 'f': out.push_back('\f'); break;
                case 'n': out.push_back('\n'); break;
                case 'r': out.push_back('\r'); break;
                case 't': out.push_back('\t'); break;
                case 'u': {
                    if (i_ + 4 > t_.size()) fail("bad escape");
                    const unsigned cp = static_cast<unsigned>(std::strtoul(t_.substr(i_, 4).c_str(), nullptr, 16));
                    i_ += 4;
                    if (cp < 0x80) {
                        out.push_back(static_cast<char>(cp));
                    } else if (cp < 0x800) {
                        out.push_back(static_cast<char>(0xC0 | (cp >> 6)));
                        out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
                    } else {
                        out.push_back(static_cast<char>(0xE0 | (cp >> 12)));
                        out.push_back(static_cast<char>(0x80 | ((cp >> 6) & 0x3F)));
                        out.push_back(static_cast<char>(0x80 | (cp & 0x3F)));
                    }
                    break;
                }
                default: fail("bad escape");
            }
        }
        return out;
    }
    JVal value() {
        if (++depth_ > kMaxDepth) fail("nested too deeply");
        JVal v = valueAt();
        depth_--;
        return v;
    }
    JVal valueAt() {
        JVal v;
        const char c = peek();
        if (c == '{') {
            i_++;
            v.kind = JVal::Obj;
            if (peek() == '}') {
                i_++;
                return v;
            }
            while (true) {
                std::string k = str();
                expect(':');
                v.obj.emplace_back(std::move(k), value());
                const char d = peek();
                i_++;
                if (d == '}') break;
                if (d != ',') fail("expected , or }");
            }
        } else if (c == '[') {
            i_++;
            v.kind = JVal::Arr;
            if (peek() == ']') {
                i_++;
                return v;
            }
            while (true) {
                v.arr.push_back(value());
                const char d = peek();
                i_++;
                if (d == ']') break;
                if (d != ',') fail("expected , or ]");
            }
        } else if (c == '"') {
            v.kind = JVal::Str;
            v.s = str();
        } else if (t_.compare(i_, 4, "true") == 0) {
            v.kind = JVal::Bool;
            v.b = true;
            i_ += 4;
        } else if (t_.compare(i_, 5, "false") == 0) {
            v.kind = JVal::Bool;
            i_ += 5;
        } else if (t_.compare(i_, 4, "null") == 0) {
            i_ += 4;
        } else {
            const size_t b = i_;
            while (i_ < t_.size() && (std::isdigit(static_cast<unsigned char>(t_[i_])) || std::strchr("+-.eE", t_[i_]))) i_++;
            if (i_ == b) fail("unexpected character");
            v.kind = JVal::Num;
            v.s = t_.substr(b, i_ - b);
        }
        return v;
    }
};

// ---- files: zlib reads plain and gzip files alike ---------------------------
std::string readFile(const char* path) {
    gzFile f = gzopen(path, "rb");
    if (!f) throw Error(CF_ERR_ARG, std::string("cannot open tree file '") + path + "'");
    std::string data;
    char buf[1 << 16];
    int n;
    while ((n = gzread(f, buf, sizeof(buf))) > 0) data.append(buf, static_cast<size_t>(n));
    const bool bad = n < 0;
    gzclose(f);
    if (bad) throw Error(CF_ERR_ARG, std::string("cannot read tree file '") + path + "'");
    return data;
}

bool endsWith(const std::string& s, const char* suffix) {
    const size_t n = std::strlen(suffix);
    return s.size() >= n && s.compare(s.size() - n, n, suffix) == 0;
}

void writeFile(const char* path, const std::string& data) {
    if (endsWith(path, ".gz")) {
        gzFile f = gzopen(path, "wb");
        if (!f) throw Error(CF_ERR_ARG, std::string("cannot create tree file '") + path + "'");
        const int n = gzwrite(f, data.data(), static_cast<unsigned>(data.size()));
        const int rc = gzclose(f);
        if (n != static_cast<int>(data.size()) || rc != Z_OK) {
            throw Error(CF_ERR_RUNTIME, std::string("cannot write tree file '") + path + "'");
        }
    } else {
        FILE* f = std::fopen(path, "wb");
        if (!f) throw Error(CF_ERR_ARG, std::string("cannot create tree file '") + path + "'");
        const size_t n = std::fwrite(data.data(), 1, data.size(), f);
        const int rc = std::fclose(f);
        if (n != data.size() || rc != 0) throw Error(CF_ERR_RUNTIME, std::string("cannot write tree file '") + path + "'");
    }
}

// ---- writing ------------------------------------------------------------------
class Writer {
public:
    std::string out;

    void open(const char* key) {
        sep();
        if (key) keyed(key);
        out += "{";
        first_ = true;
        depth_++;
    }
    void close() {
        depth_--;
        nl();
        out += "}";
        first_ = false;
    }
    void openArr(const char* key) {
        sep();
        keyed(key);
        out += "[";
        first_ = true;
        depth_++;
    }
    void closeArr() {
        depth_--;
        nl();
        out += "]";
        first_ = false;
    }
    void put(const std::string& key, const std::string& value) {
        sep();
        keyed(key);
        quoted(value);
    }
    void item(const std::string& value) {
        sep();
        quoted(value);
    }

private:
    bool first_ = true;
    int depth_ = 0;

    void nl() {
        out += "\n";
        out.append(static_cast<size_t>(depth_) * 4, ' ');
    }
    void sep() {
        if (!first_) out += ",";
        if (depth_ > 0) nl();
        first_ = false;
    }
    void keyed(const std::string& key) {
        quoted(key);
        out += ": ";
    }
    void quoted(const std::string& v) {
        out += '"';
        for (char c : v) {
            if (c == '"' || c == '\\') {
                out += '\\';
                out += c;
            } else if (static_cast<unsigned char>(c) < 0x20) {
                char b[8];
                std::snprintf(b, sizeof(b), "\\u%04x", c);
                out += b;
            } else {
                out += c;
            }
        }
        out += '"';
    }
};

std::string fmtFloat(float v) {
    char b[32];
    std::snprintf(b, sizeof(b), "%.9g", static_cast<double>(v));
    return b;
}
std::string fmtDouble(double v) {
    char b[40];
    std::snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

std::string colorKey(const uint8_t* palette, int label) {
    std::ostringstream o;
    if (palette) {
        o << static_cast<int>(palette[3 * label]) << "," << static_cast<int>(palette[3 * label + 1]) << ","
          << static_cast<int>(palette[3 * label + 2]);
    } else {
        o << label << "," << label << "," << label;
    }
    o << " (" << label << ")";
    return o.str();
}

void writeXY(Writer& w, const char* key, int x, int y) {
    w.openArr(key);
    w.item(std::to_string(x));
    w.item(std::to_string(y));
    w.closeArr();
}

void writeNode(Writer& w, const TreeBuild& tb, int i, int C, const uint8_t* palette) {
    const cf_tree_node& n = tb.nodes[i];
    w.put("id", std::to_string(n.id));
    w.put("level", std::to_string(n.level));
    w.put("samples", std::to_string(n.num_samples));
    w.put("leaf", n.is_leaf ? "true" : "false");
    w.open("histogram");
    for (int c = 0; c < C; c++) w.put(colorKey(palette, c), std::to_string(tb.hist[static_cast<size_t>(i) * C + c]));
    w.close();
    if (!n.is_leaf) {
        w.open("split");
        w.put("threshold", fmtFloat(n.threshold));
        w.put("score", fmtDouble(n.score));
        w.put("featureId", std::to_string(n.feature_id));
        w.open("feature");
        const bool color = n.feature[P_TYPE] == T_COLOR;
        w.put("type", color ? "color" : "depth");
        if (color) {
            w.put("channel1", std::to_string(static_cast<uint8_t>(n.feature[P_C1])));
            w.put("channel2", std::to_string(static_cast<uint8_t>(n.feature[P_C2])));
        }
        writeXY(w, "offset1", n.feature[P_O1X], n.feature[P_O1Y]);
        writeXY(w, "region1", n.feature[P_R1X], n.feature[P_R1Y]);
        writeXY(w, "offset2", n.feature[P_O2X], n.feature[P_O2Y]);
        writeXY(w, "region2", n.feature[P_R2X], n.feature[P_R2Y]);
        w.close();
        w.close();
        w.open("left");
        writeNode(w, tb, n.left, C, palette);
        w.close();
        w.open("right");
        writeNode(w, tb, n.right, C, palette);
        w.close();
    }
}

// ---- reading ------------------------------------------------------------------
int8_t toParam(long long v, const char* what) {
    if (v < -128 || v > 127) throw Error(CF_ERR_ARG, std::string("tree file: ") + what + " out of range");
    return static_cast<int8_t>(v);
}

void readXY(const JVal& v, int8_t& x, int8_t& y, const char* what) {
    if (v.kind != JVal::Arr || v.arr.size() != 2) throw Error(CF_ERR_ARG, std::string("tree file: bad ") + what);
    x = toParam(v.arr[0].asInt(), what);
    y = toParam(v.arr[1].asInt(), what);
}

// histogram keys "R,G,B (label)": the label and its colour
int parseHistKey(const std::string& key, uint8_t* rgb) {
    const size_t sp = key.find(" (");
    const size_t cl = key.find(')', sp == std::string::npos ? 0 : sp);
    if (sp == std::string::npos || cl == std::string::npos) {
        throw Error(CF_ERR_ARG, "tree file: bad histogram key '" + key + "'");
    }
    int r = 0, g = 0, b = 0;
    if (std::sscanf(key.substr(0, sp).c_str(), "%d,%d,%d", &r, &g, &b) != 3) {
        throw Error(CF_ERR_ARG, "tree file: bad histogram colour '" + key + "'");
    }
    rgb[0] = static_cast<uint8_t>(r);
    rgb[1] = static_cast<uint8_t>(g);
    rgb[2] = static_cast<uint8_t>(b);
    return std::atoi(key.substr(sp + 2, cl - sp - 2).c_str());
}

struct ParsedTree {
    TreeBuild tb;
    std::vector<uint32_t> prior;
    int C = 0;
};

// breadth-first: a node's children are stored as an adjacent pair after it
ParsedTree readTree(const JVal& root, uint8_t* palette, std::vector<bool>& paletteSet) {
    ParsedTree out;
    std::deque<std::pair<const JVal*, int>> q;   // node, its index
    out.tb.nodes.emplace_back();
    q.emplace_back(&root, 0);
    std::vector<int> parentOf(1, -1);
    while (!q.empty()) {
        const JVal& v = *q.front().first;
        const int idx = q.front().second;
        q.pop_front();
        cf_tree_node n{};
        n.id = static_cast<int32_t>(v.at("id").asInt());
        n.level = static_cast<int32_t>(v.at("level").asInt());
        n.num_samples = static_cast<int32_t>(v.at("samples").asInt());
        n.parent = parentOf[idx];
        n.left = n.right = -1;
        n.is_leaf = 1;
        const JVal& hist = v.at("histogram");
        if (hist.kind != JVal::Obj) throw Error(CF_ERR_ARG, "tree file: bad histogram");
        const int C = static_cast<int>(hist.obj.size());
        if (out.C == 0) {
            out.C = C;
            if (C < 1 || C > kMaxClasses) throw Error(CF_ERR_ARG, "tree file: number of classes must be in [1, 32]");
            out.tb.hist.assign(C, 0u);
        } else if (C != out.C) {
            throw Error(CF_ERR_ARG, "tree file: nodes with different numbers of classes");
        }
        out.tb.hist.resize(out.tb.nodes.size() * static_cast<size_t>(C), 0u);
        std::vector<bool> seen(C, false);
        double sum = 0;
        for (const auto& kv : hist.obj) {
            uint8_t rgb[3];
            const int label = parseHistKey(kv.first, rgb);
            if (label < 0 || label >= C || seen[label]) {
                throw Error(CF_ERR_ARG, "tree file: illegal histogram label in '" + kv.first + "'");
            }
            seen[label] = true;
            const long long cnt = kv.second.asInt();
            if (cnt < 0 || cnt > 0xffffffffll) throw Error(CF_ERR_ARG, "tree file: bad histogram count");
            out.tb.hist[static_cast<size_t>(idx) * C + label] = static_cast<uint32_t>(cnt);
            sum += static_cast<double>(cnt);
            {
                if (!paletteSet[label]) {
                    std::memcpy(palette + 3 * label, rgb, 3);
                    paletteSet[label] = true;
                } else if (std::memcmp(palette + 3 * label, rgb, 3) != 0) {
                    throw Error(CF_ERR_ARG, "tree file: label " + std::to_string(label) + " has two colours");
                }
            }
        }
        if (sum < n.num_samples) throw Error(CF_ERR_ARG, "tree file: incorrect histogram sum");   // import.cpp:145
        const JVal* left = v.find("left");
        if (left) {
            const JVal& split = v.at("split");
            const JVal& right = v.at("right");
            n.is_leaf = 0;
            n.threshold = static_cast<float>(split.at("threshold").asDouble());
            n.score = split.at("score").asDouble();
            n.feature_id = static_cast<int32_t>(split.at("featureId").asInt());
            const JVal& f = split.at("feature");
            const std::string& type = f.at("type").text();
            if (type == "color") {
                n.feature[P_TYPE] = T_COLOR;
                const long long c1 = f.at("channel1").asInt(), c2 = f.at("channel2").asInt();
                if (c1 < 0 || c1 > 2 || c2 < 0 || c2 > 2) throw Error(CF_ERR_ARG, "tree file: invalid colour channel");
                n.feature[P_C1] = static_cast<int8_t>(c1);
                n.feature[P_C2] = static_cast<int8_t>(c2);
            } else if (type == "depth") {
                n.feature[P_TYPE] = T_DEPTH;
            } else {
                throw Error(CF_ERR_ARG, "tree file: illegal feature type: " + type);
            }
            readXY(f.at("offset1"), n.feature[P_O1X], n.feature[P_O1Y], "offset1");
            readXY(f.at("region1"), n.feature[P_R1X], n.feature[P_R1Y], "region1");
            readXY(f.at("offset2"), n.feature[P_O2X], n.feature[P_O2Y], "offset2");
            readXY(f.at("region2"), n.feature[P_R2X], n.feature[P_R2Y], "region2");
            const int li = static_cast<int>(out.tb.nodes.size());
            n.left = li;
            n.right = li + 1;
            out.tb.nodes.emplace_back();
            out.tb.nodes.emplace_back();
            parentOf.push_back(idx);
            parentOf.push_back(idx);
            q.emplace_back(left, li);
            q.emplace_back(&right, li + 1);
        }
        out.tb.nodes[idx] = n;
    }
    out.tb.hist.resize(out.tb.nodes.size() * static_cast<size_t>(out.C), 0u);
    return out;
}

std::string nowString() {
    const auto t = std::chrono::system_clock::now();
    const std::time_t tt = std::chrono::system_clock::to_time_t(t);
    const long us = static_cast<long>(
            std::chrono::duration_cast<std::chrono::microseconds>(t.time_since_epoch()).count() % 1000000);
    std::tm tmv{};
    localtime_r(&tt, &tmv);
    char b[64];
    std::strftime(b, sizeof(b), "%Y-%b-%d %H:%M:%S", &tmv);
    char c[80];
    std::snprintf(c, sizeof(c), "%s.%06ld", b, us);
    return c;
}

}  // namespace

void writeTreeFile(const cf_forest* fo, int tree, const char* path, const cf_train_config& cfg,
        const uint8_t* palette, const char* folderTraining) {
    if (tree < 0 || tree >= static_cast<int>(fo->trees.size())) throw Error(CF_ERR_ARG, "tree index out of range");
    const TreeBuild& tb = fo->trees[tree];
    const int C = fo->C;
    Writer w;
    w.open(nullptr);
    char host[256] = {0};
    gethostname(host, sizeof(host) - 1);
    w.put("hostname", host);
    w.put("buildType", "RELEASE");
    w.put("version", "curfil_amd");
    w.put("date", nowString());
    w.put("randomSeed", std::to_string(cfg.random_seed));
    w.put("folderTraining", folderTraining ? folderTraining : "");
    w.put("numCores", std::to_string(sysconf(_SC_NPROCESSORS_ONLN)));
    w.put("numPhysicalCPUs", "1");
    w.openArr("processorModels");
    w.closeArr();
    w.put("accelerationMode", "gpu");
    w.openArr("deviceId");
    w.item(std::to_string(fo->ctx ? fo->ctx->device : 0));
    w.closeArr();
    w.put("samplesPerImage", std::to_string(cfg.samples_per_image));
    w.put("featureCount", std::to_string(cfg.feature_count));
    w.put("thresholds", std::to_string(cfg.thresholds));
    w.put("boxRadius", std::to_string(cfg.box_radius));
    w.put("regionSize", std::to_string(cfg.region_size));
    w.put("maxDepth", std::to_string(cfg.max_depth));
    w.put("numThreads", std::to_string(cfg.num_threads));
    w.put("minSampleCount", std::to_string(cfg.min_sample_count));
    w.put("maxSamplesPerBatch", "0");
    w.put("maxImages", std::to_string(cfg.max_images));
    w.put("imageCacheSize", "0");
    w.put("subsamplingType", cfg.subsampling_pixel_uniform ? "pixelUniform" : "classUniform");
    w.put("useCIELab", cfg.use_cielab ? "true" : "false");
    w.put("useDepthFilling", cfg.use_depth_filling ? "true" : "false");
    w.put("useDepthImages", cfg.use_depth_images ? "true" : "false");
    w.openArr("ignoredColors");
    for (int i = 0; i < cfg.num_ignored && i < 8; i++) {
        const int l = cfg.ignored_labels[i];
        const std::string k = colorKey(palette && l >= 0 && l < C ? palette : nullptr, l);
        w.item(k.substr(0, k.find(' ')));
    }
    w.closeArr();
    w.open("classLabelPriorDistribution");
    const std::vector<uint32_t>& prior = fo->treePriors.empty() ? fo->prior : fo->treePriors[tree];
    for (int c = 0; c < C; c++) w.put(std::to_string(c), std::to_string(prior[c]));
    w.close();
    w.open("tree");
    writeNode(w, tb, 0, C, palette);
    w.close();
    w.close();
    w.out += "\n";
    writeFile(path, w.out);
}

std::unique_ptr<cf_forest> readTreeFiles(cf_ctx* ctx, const char* const* paths, int count, cf_train_config* cfgOut,
        uint8_t* palette) {
    if (!paths || count <= 0) throw Error(CF_ERR_ARG, "no tree files");
    auto fo = std::make_unique<cf_forest>();
    fo->ctx = ctx;
    std::vector<bool> paletteSet(kMaxClasses, false);
    uint8_t pal[3 * kMaxClasses] = {0};
    cf_train_config first{};
    for (int t = 0; t < count; t++) {
        if (!paths[t]) throw Error(CF_ERR_ARG, "null tree file name");
        const JVal doc = Parser(readFile(paths[t])).parse();
        if (doc.kind != JVal::Obj) throw Error(CF_ERR_ARG, "tree file: not a JSON object");
        // keys RandomTreeImport::readJSON requires besides the ones read below (import.cpp:196-236)
        for (const char* k : {"date", "folderTraining", "hostname", "maxSamplesPerBatch", "accelerationMode",
                     "imageCacheSize"}) {
            (void)doc.at(k);
        }
        cf_train_config c{};
        c.random_seed = static_cast<int32_t>(doc.at("randomSeed").asInt());
        c.samples_per_image = static_cast<uint32_t>(doc.at("samplesPerImage").asInt());
        c.feature_count = static_cast<uint32_t>(doc.at("featureCount").asInt());
        c.min_sample_count = static_cast<uint32_t>(doc.at("minSampleCount").asInt());
        c.max_depth = static_cast<int32_t>(doc.at("maxDepth").asInt());
        c.box_radius = static_cast<uint16_t>(doc.at("boxRadius").asInt());
        c.region_size = static_cast<uint16_t>(doc.at("regionSize").asInt());
        c.thresholds = static_cast<uint16_t>(doc.at("thresholds").asInt());
        c.num_threads = static_cast<int32_t>(doc.at("numThreads").asInt());
        const JVal* mi = doc.find("maxImages");
        c.max_images = mi ? static_cast<int32_t>(mi->asInt()) : 0;
        const JVal* udi = doc.find("useDepthImages");
        c.use_depth_images = udi ? (udi->asBool() ? 1 : 0) : 1;
        const JVal* st = doc.find("subsamplingType");
        c.subsampling_pixel_uniform = st ? (st->text() == "pixelUniform" ? 1 : 0) : 1;
        // optional keys and their defaults as RandomTreeImport::readJSON reads them (import.cpp:210-248)
        const JVal* lab = doc.find("useCIELab");
        c.use_cielab = lab ? (lab->asBool() ? 1 : 0) : 1;
        const JVal* fill = doc.find("useDepthFilling");
        c.use_depth_filling = fill ? (fill->asBool() ? 1 : 0) : 0;
        ParsedTree pt = readTree(doc.at("tree"), pal, paletteSet);
        // ignored colours as labels, through the colours of the histogram keys
        const JVal* ig = doc.find("ignoredColors");
        if (ig && ig->kind == JVal::Arr) {
            for (const JVal& e : ig->arr) {
                int r = 0, g = 0, b = 0;
                if (std::sscanf(e.text().c_str(), "%d,%d,%d", &r, &g, &b) != 3) {
                    throw Error(CF_ERR_ARG, "tree file: bad ignored colour '" + e.text() + "'");
                }
                for (int l = 0; l < pt.C; l++) {
                    if (paletteSet[l] && pal[3 * l] == r && pal[3 * l + 1] == g && pal[3 * l + 2] == b &&
                            c.num_ignored < 8) {
                        c.ignored_labels[c.num_ignored++] = l;
                    }
                }
            }
        }
        // TrainingConfiguration::equals without `strict` (random_tree.cpp:168-214): trees
        // trained with different seeds make one forest (random_forest_image.cpp:85-98)
        if (t == 0) {
            first = c;
        } else if (c.samples_per_image != first.samples_per_image || c.feature_count != first.feature_count ||
                c.min_sample_count != first.min_sample_count || c.max_depth != first.max_depth ||
                c.box_radius != first.box_radius || c.region_size != first.region_size ||
                c.thresholds != first.thresholds || c.num_threads != first.num_threads ||
                c.max_images != first.max_images || c.subsampling_pixel_uniform != first.subsampling_pixel_uniform ||
                c.use_depth_images != first.use_depth_images || c.num_ignored != first.num_ignored ||
                c.use_cielab != first.use_cielab || c.use_depth_filling != first.use_depth_filling ||
                std::memcmp(c.ignored_labels, first.ignored_labels, sizeof(c.ignored_labels)) != 0) {
            throw Error(CF_ERR_ARG, "different configurations");
        }
        const JVal& pr = doc.at("classLabelPriorDistribution");
        if (pr.kind != JVal::Obj) throw Error(CF_ERR_ARG, "tree file: bad classLabelPriorDistribution");
        std::vector<uint32_t> prior(pt.C, 0u);
        for (const auto& kv : pr.obj) {
            const int label = std::atoi(kv.first.c_str());
            if (label < 0 || label >= pt.C) throw Error(CF_ERR_ARG, "tree file: prior label out of range");
            const long long v = kv.second.asInt();
            if (v < 0 || v > 0xffffffffll) throw Error(CF_ERR_ARG, "tree file: bad prior count");
            prior[label] = static_cast<uint32_t>(v);
        }
        if (t == 0) {
            fo->C = pt.C;
        } else if (pt.C != fo->C) {
            throw Error(CF_ERR_ARG, "different number of classes in trees");
        }
        fo->trees.push_back(std::move(pt.tb));
        fo->treePriors.push_back(std::move(prior));
    }
    fo->prior = fo->treePriors[0];
    if (cfgOut) *cfgOut = first;
    if (palette) std::memcpy(palette, pal, 3 * static_cast<size_t>(fo->C));
    return fo;
}

}  // namespace cf
