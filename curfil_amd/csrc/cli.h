// Shared pieces of the curfil_train / curfil_predict command-line tools
// (cf_tool_train, cf_tool_predict):
// option parsing in the style of the reference's boost::program_options
// setup (named options `--name value` or `--name=value`, options with an
// implicit value, repeatable options, positional arguments in a fixed order),
// listing of `*_colors.png` files (listImageFilenames, image.cpp:844-866), log
// lines, and the C-ABI status check.
#pragma once

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <ctime>
#include <filesystem>
#include <map>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/curfil_hip.h"

namespace cli {

struct Option {
    std::string name, help, defaultValue;
    bool required = false;
    bool repeatable = false;
    bool hasImplicit = false;
    std::string implicitValue;
};

class Options {
public:
    void add(const std::string& name, const std::string& help, const std::string& def = "", bool required = false) {
        Option o;
        o.name = name;
        o.help = help;
        o.defaultValue = def;
        o.required = required;
        opts_.push_back(o);
    }
    // a boolean option; given without a value it takes `implicit`
    void addFlag(const std::string& name, const std::string& help, bool def, bool implicit = true) {
        add(name, help, def ? "true" : "false");
        opts_.back().hasImplicit = true;
        opts_.back().implicitValue = implicit ? "true" : "false";
    }
    void addRepeatable(const std::string& name, const std::string& help, bool required = false) {
        add(name, help, "", required);
        opts_.back().repeatable = true;
    }
    // positional arguments fill these options in order; the last one may take the rest
    void positional(const std::vector<std::string>& names) { positional_ = names; }

    // returns false for --help / --version / no arguments (usage printed)
    bool parse(int argc, char** argv) {
        if (argc <= 1) {
            usage(argv[0]);
            return false;
        }
        size_t pos = 0;
        for (int i = 1; i < argc; i++) {
            std::string a = argv[i];
            if (a == "--help") {
                usage(argv[0]);
                return false;
            }
            if (a == "--version") {
                std::printf("%s version curfil_amd\n", argv[0]);
                std::fflush(stdout);
                return false;
            }
            if (a.rfind("--", 0) == 0) {
                std::string name = a.substr(2), value;
                bool hasValue = false;
                const size_t eq = name.find('=');
                if (eq != std::string::npos) {
                    value = name.substr(eq + 1);
                    name = name.substr(0, eq);
                    hasValue = true;
                }
                const Option* o = find(name);
                if (!o) throw std::runtime_error("unrecognised option '--" + name + "'");
                if (!hasValue) {
                    const bool nextIsValue = i + 1 < argc && std::string(argv[i + 1]).rfind("--", 0) != 0;
                    if (o->hasImplicit && (!nextIsValue || !isBool(argv[i + 1]))) {
                        value = o->implicitValue;
                    } else if (nextIsValue) {
                        value = argv[++i];
                    } else {
                        throw std::runtime_error("the option '--" + name + "' requires an argument");
                    }
                }
                set(*o, value);
            } else {
                if (pos >= positional_.size()) throw std::runtime_error("too many positional options");
                const Option* o = find(positional_[pos]);
                set(*o, a);
                if (!o->repeatable) pos++;
            }
        }
        for (const Option& o : opts_) {
            if (o.required && !values_.count(o.name)) {
                throw std::runtime_error("the option '--" + o.name + "' is required but missing");
            }
        }
        return true;
    }

    std::string str(const std::string& name) const {
        auto it = values_.find(name);
        if (it != values_.end()) return it->second.back();
        return find(name)->defaultValue;
    }
    std::vector<std::string> all(const std::string& name) const {
        auto it = values_.find(name);
        return it == values_.end() ? std::vector<std::string>() : it->second;
    }
    bool given(const std::string& name) const { return values_.count(name) > 0; }
    long integer(const std::string& name) const {
        const std::string v = str(name);
        char* end = nullptr;
        const long x = std::strtol(v.c_str(), &end, 10);
        if (v.empty() || *end) throw std::runtime_error("the argument ('" + v + "') for option '--" + name + "' is invalid");
        return x;
    }
    double real(const std::string& name) const {
        const std::string v = str(name);
        char* end = nullptr;
        const double x = std::strtod(v.c_str(), &end);
        if (v.empty() || *end) throw std::runtime_error("the argument ('" + v + "') for option '--" + name + "' is invalid");
        return x;
    }
    bool flag(const std::string& name) const {
        const std::string v = str(name);
        if (v == "1" || v == "true" || v == "yes" || v == "on") return true;
        if (v == "0" || v == "false" || v == "no" || v == "off") return false;
        throw std::runtime_error("the argument ('" + v + "') for option '--" + name + "' is invalid");
    }

private:
    std::vector<Option> opts_;
    std::vector<std::string> positional_;
    std::map<std::string, std::vector<std::string>> values_;

    static bool isBool(const std::string& v) {
        return v == "1" || v == "0" || v == "true" || v == "false" || v == "yes" || v == "no" || v == "on" ||
                v == "off";
    }
    const Option* find(const std::string& name) const {
        for (const Option& o : opts_) {
            if (o.name == name) return &o;
        }
        return nullptr;
    }
    void set(const Option& o, const std::string& v) {
        auto& vs = values_[o.name];
        if (!vs.empty() && !o.repeatable) throw std::runtime_error("option '--" + o.name + "' cannot be specified more than once");
        vs.push_back(v);
    }
    void usage(const char* prog) const {
        std::printf("usage: %s [options]\noptions:\n", prog);
        for (const Option& o : opts_) {
            std::printf("  --%-24s %s%s%s\n", o.name.c_str(), o.help.c_str(),
                    o.defaultValue.empty() ? "" : (" (=" + o.defaultValue + ")").c_str(), o.required ? " [required]" : "");
        }
        std::fflush(stdout);
    }
};

inline void log(const char* level, const std::string& msg) {
    const std::time_t t = std::time(nullptr);
    char buf[32];
    std::strftime(buf, sizeof(buf), "%Y-%m-%d %H:%M:%S", std::localtime(&t));
    std::fprintf(stderr, "%s %s %s\n", buf, level, msg.c_str());
}
inline void info(const std::string& msg) { log("INFO", msg); }
inline void warning(const std::string& msg) { log("WARNING", msg); }

inline void check(const cf_ctx* ctx, int status) {
    if (status != CF_OK) throw std::runtime_error(cf_last_error(ctx));
}

// listImageFilenames (image.cpp:844-866): files whose name contains "_colors.png",
// recursively, sorted by path
inline std::vector<std::string> listImageFilenames(const std::string& folder) {
    std::vector<std::string> out;
    namespace fs = std::filesystem;
    if (!fs::is_directory(folder)) throw std::runtime_error("not a folder: " + folder);
    for (const auto& e : fs::recursive_directory_iterator(folder)) {
        if (e.is_regular_file() && e.path().filename().string().find("_colors.png") != std::string::npos) {
            out.push_back(e.path().string());
        }
    }
    std::sort(out.begin(), out.end());
    return out;
}

// RGBColor(colorString) (image.cpp:66-91)
inline void parseColor(const std::string& s, uint8_t* rgb) {
    int v[3];
    char tail;
    if (std::sscanf(s.c_str(), "%d,%d,%d%c", &v[0], &v[1], &v[2], &tail) != 3) {
        throw std::runtime_error("illegal color: '" + s + "'");
    }
    for (int c = 0; c < 3; c++) {
        if (v[c] < 0 || v[c] > 255) throw std::runtime_error("color value out of range [0-255]: " + s);
        rgb[c] = static_cast<uint8_t>(v[c]);
    }
}

inline double seconds(std::chrono::steady_clock::time_point t0) {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace cli
