// Launch-shape tuner of pass 1 (k_eval_bins).
//
// Every shape writes the same bin rows (a shape only changes how a level's
// (sample, feature) responses are spread over workgroups and lanes), so the
// tuner measures on real work: while a (threshold width, level size class)
// key is open, successive pass-1 launches of that key alternate between the
// default shape and each other shape in turn (0, 1, 0, 2, ..., 0, n-1), and
// each shape's time per response is taken relative to the default launch just
// before it.  The launches of a key span several levels, whose time per
// response differs by a few percent; neighbouring launches are alike, so the
// ratio cancels that.  Once every shape has kRounds ratios (launches of at
// least kMinResponses responses), the key keeps the shape with the lowest mean
// ratio (the default's is 1).  Levels below the smallest size class keep the
// default shape: their launches are too short to time.
//
// CURFIL_EVAL_SHAPE=i pins shape i (parity tests run every shape this way;
// read at every launch), CURFIL_TUNE=0 keeps the default shape,
// CURFIL_TUNE_ROUNDS=r sets the ratios per shape, and
// CURFIL_TUNE_FILE=path loads finished keys from a file and writes them back
// when a key finishes.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <map>
#include <mutex>
#include <sstream>
#include <vector>

#include "cf_host.h"

namespace cf {

namespace {

// Ratios per shape.  Timing the shapes in turn, without the default between them, chose shape 1 or 8 on
// the 800-frame workload (17.57 and 17.50 s per step) where shape 7 pinned took 17.41 s: the launch order
// and the levels' differences decided more than the shapes did.  With ratios, 3 rounds chose shape 3 in
// three runs out of three (17.48 s per step); 6 rounds, which reach further down the tree, chose shape 7
// in two runs out of two (17.44 s).
constexpr int kRounds = 6;
// CURFIL_TUNE_ROUNDS=r overrides the ratios per shape (tests on small workloads, which have few launches)
int tuneRounds() {
    const char* v = std::getenv("CURFIL_TUNE_ROUNDS");
    const int r = v ? std::atoi(v) : 0;
    return r > 0 ? r : kRounds;
}
constexpr double kMinResponses = 5e7;

// size classes of a level by permutation entries: < 2^20 (not tuned), < 2^23, < 2^25, larger
int sizeClass(size_t entries) {
    if (entries < (static_cast<size_t>(1) << 20)) return 0;
    if (entries < (static_cast<size_t>(1) << 23)) return 1;
    if (entries < (static_cast<size_t>(1) << 25)) return 2;
    return 3;
}

struct KeyState {
    std::vector<double> ms, responses;   // per shape, for the report
    std::vector<double> ratioSum;        // per shape: time per response over the preceding default launch's
    std::vector<int> ratios, timed;
    int cand = 1;              // the shape the next candidate launch runs
    bool defaultNext = true;   // the next launch runs the default shape
    double lastDefault = -1;   // time per response of the last default launch
    int best = -1;             // >= 0 once finished
};

class Tuner {
public:
    static Tuner& get() {
        static Tuner t;
        return t;
    }

    int choose(int TM, size_t entries, bool* measure) {
        *measure = false;
        const int pin = pinned();
        if (pin >= 0) return pin;
        if (!enabled()) return 0;
        const int cls = sizeClass(entries);
        if (cls == 0 && !tuneSmall()) return 0;
        std::lock_guard<std::mutex> g(mu_);
        load();
        KeyState& k = state(TM, cls);
        if (k.best >= 0) return k.best;
        *measure = true;
        return k.defaultNext ? 0 : k.cand;
    }

    void record(int TM, size_t entries, int shape, double ms, double responses) {
        if (pinned() >= 0 || !enabled()) return;
        const int cls = sizeClass(entries);
        if ((cls == 0 || responses < kMinResponses) && !tuneSmall()) return;
        std::lock_guard<std::mutex> g(mu_);
        KeyState& k = state(TM, cls);
        if (k.best >= 0) return;
        const int n = numEvalShapes();
        k.ms[shape] += ms;
        k.responses[shape] += responses;
        k.timed[shape]++;
        const double per = ms / responses;
        if (shape == 0) {
            k.lastDefault = per;
            if (n > 1) k.defaultNext = false;
        } else if (shape == k.cand && !k.defaultNext) {
            if (k.lastDefault > 0) {
                k.ratioSum[shape] += per / k.lastDefault;
                k.ratios[shape]++;
            }
            k.cand = k.cand + 1 < n ? k.cand + 1 : 1;
            k.defaultNext = true;
        }
        const int rounds = tuneRounds();
        for (int i = 1; i < n; i++) {
            if (k.ratios[i] < rounds) return;
        }
        int best = 0;
        double bestRatio = 1.0;
        for (int i = 1; i < n; i++) {
            const double r = k.ratioSum[i] / k.ratios[i];
            if (r < bestRatio) {
                best = i;
                bestRatio = r;
            }
        }
        k.best = best;
        save();
    }

    // per key: TM, size class, chosen shape (-1: open), ns per response of the chosen shape (estimated from
    // its mean ratio) and of the default shape, launches timed
    std::vector<double> table() {
        std::lock_guard<std::mutex> g(mu_);
        std::vector<double> out;
        for (auto& kv : keys_) {
            const KeyState& k = kv.second;
            int timed = 0;
            for (int v : k.timed) timed += v;
            const double ns0 = k.responses[0] > 0 ? k.ms[0] * 1e6 / k.responses[0] : 0.0;
            // the chosen shape on the default's launches: the default's time times the chosen shape's mean ratio
            double nsBest = 0.0;
            if (k.best == 0) nsBest = ns0;
            if (k.best > 0 && k.ratios[k.best] > 0) nsBest = ns0 * k.ratioSum[k.best] / k.ratios[k.best];
            out.insert(out.end(), {static_cast<double>(kv.first.first), static_cast<double>(kv.first.second),
                    static_cast<double>(k.best), nsBest, ns0, static_cast<double>(timed)});
        }
        return out;
    }

    void reset() {
        std::lock_guard<std::mutex> g(mu_);
        keys_.clear();
        loaded_ = false;
    }

private:
    Tuner() {
        const char* file = std::getenv("CURFIL_TUNE_FILE");
        if (file) file_ = file;
    }

    // read per launch, so a caller can pin a shape for one training and release it afterwards
    static int pinned() {
        const char* pin = std::getenv("CURFIL_EVAL_SHAPE");
        if (!pin || !*pin) return -1;
        const int v = std::atoi(pin);
        if (v < 0 || v >= numEvalShapes()) throw Error(CF_ERR_ARG, "CURFIL_EVAL_SHAPE out of range");
        return v;
    }

    // CURFIL_TUNE_SMALL=1: levels and launches of any size are tuned and timed (for tests on small workloads)
    static bool tuneSmall() {
        const char* v = std::getenv("CURFIL_TUNE_SMALL");
        return v && std::string(v) == "1";
    }

    static bool enabled() {
        const char* tune = std::getenv("CURFIL_TUNE");
        return !(tune && std::string(tune) == "0");
    }

    KeyState& state(int TM, int cls) {
        KeyState& k = keys_[{TM, cls}];
        if (k.ms.empty()) {
            const int n = numEvalShapes();
            k.ms.assign(n, 0.0);
            k.responses.assign(n, 0.0);
            k.ratioSum.assign(n, 0.0);
            k.ratios.assign(n, 0);
            k.timed.assign(n, 0);
        }
        return k;
    }

    // file lines: "TM class shape"
    void load() {
        if (loaded_ || file_.empty()) return;
        loaded_ = true;
        std::ifstream in(file_);
        std::string line;
        while (std::getline(in, line)) {
            std::istringstream ls(line);
            int tm = 0, cls = 0, shape = -1;
            if (!(ls >> tm >> cls >> shape) || shape < 0 || shape >= numEvalShapes()) continue;
            state(tm, cls).best = shape;
        }
    }

    void save() {
        if (file_.empty()) return;
        std::ofstream o(file_, std::ios::trunc);
        for (auto& kv : keys_) {
            if (kv.second.best >= 0) o << kv.first.first << ' ' << kv.first.second << ' ' << kv.second.best << '\n';
        }
    }

    std::mutex mu_;
    std::map<std::pair<int, int>, KeyState> keys_;
    bool loaded_ = false;
    std::string file_;
};

}  // namespace

int evalTunerChoose(int TM, size_t entries, bool* measure) { return Tuner::get().choose(TM, entries, measure); }

void evalTunerRecord(int TM, size_t entries, int shape, double ms, double responses) {
    Tuner::get().record(TM, entries, shape, ms, responses);
}

}  // namespace cf

extern "C" {

int cf_eval_shape_info(int shape, int* lanes, int* per_lane, int* features_per_step, int* workgroups_per_cu) {
    if (shape < 0 || shape >= cf::numEvalShapes()) return CF_ERR_ARG;
    const cf::EvalShape& s = cf::evalShape(shape);
    if (lanes) *lanes = s.lanes;
    if (per_lane) *per_lane = s.perLane;
    if (features_per_step) *features_per_step = s.featPerStep;
    if (workgroups_per_cu) *workgroups_per_cu = s.wgPerCu;
    return CF_OK;
}

int cf_eval_tuning(double* out, int n) {
    try {
        const std::vector<double> t = cf::Tuner::get().table();
        if (out) {
            for (int i = 0; i < n && i < static_cast<int>(t.size()); i++) out[i] = t[i];
        }
        return static_cast<int>(t.size() / 6);
    } catch (const cf::Error& e) {
        return e.code;
    }
}

int cf_eval_tuning_reset(void) {
    try {
        cf::Tuner::get().reset();
        return CF_OK;
    } catch (const cf::Error& e) {
        return e.code;
    }
}

}  // extern "C"
