// Launch-shape tuner of pass 1 (k_eval_bins).
//
// Every shape writes the same bin rows (a shape only changes how a level's
// (sample, feature) responses are spread over workgroups and lanes), so the
// tuner measures on real work: while a (threshold width, level size class)
// key is open, successive pass-1 launches of that key run the candidate
// shapes in turn and are timed; once every shape has been timed kRounds
// times on launches of at least kMinResponses responses, the key keeps the
// shape with the lowest time per response.  Levels below the smallest size
// class keep the default shape: their launches are too short to time.
//
// CURFIL_EVAL_SHAPE=i pins shape i (parity tests run every shape this way;
// read at every launch), CURFIL_TUNE=0 keeps the default shape, and
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

constexpr int kRounds = 2;
constexpr double kMinResponses = 5e7;

// size classes of a level by permutation entries: < 2^20 (not tuned), < 2^23, < 2^25, larger
int sizeClass(size_t entries) {
    if (entries < (static_cast<size_t>(1) << 20)) return 0;
    if (entries < (static_cast<size_t>(1) << 23)) return 1;
    if (entries < (static_cast<size_t>(1) << 25)) return 2;
    return 3;
}

struct KeyState {
    std::vector<double> ms, responses;
    std::vector<int> timed;
    int next = 0;        // next shape to time
    int best = -1;       // >= 0 once finished
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
        return k.next;
    }

    void record(int TM, size_t entries, int shape, double ms, double responses) {
        if (pinned() >= 0 || !enabled()) return;
        const int cls = sizeClass(entries);
        if ((cls == 0 || responses < kMinResponses) && !tuneSmall()) return;
        std::lock_guard<std::mutex> g(mu_);
        KeyState& k = state(TM, cls);
        if (k.best >= 0) return;
        k.ms[shape] += ms;
        k.responses[shape] += responses;
        k.timed[shape]++;
        const int n = numEvalShapes();
        k.next = (shape + 1) % n;
        for (int i = 0; i < n; i++) {
            if (k.timed[i] < kRounds) return;
        }
        int best = 0;
        for (int i = 1; i < n; i++) {
            if (k.ms[i] / k.responses[i] < k.ms[best] / k.responses[best]) best = i;
        }
        k.best = best;
        save();
    }

    // per key: TM, size class, chosen shape (-1: open), ns per response of the chosen and of the default
    // shape, launches timed
    std::vector<double> table() {
        std::lock_guard<std::mutex> g(mu_);
        std::vector<double> out;
        for (auto& kv : keys_) {
            const KeyState& k = kv.second;
            int timed = 0;
            for (int v : k.timed) timed += v;
            const auto nsPer = [&](int i) { return k.responses[i] > 0 ? k.ms[i] * 1e6 / k.responses[i] : 0.0; };
            out.insert(out.end(), {static_cast<double>(kv.first.first), static_cast<double>(kv.first.second),
                    static_cast<double>(k.best), k.best >= 0 ? nsPer(k.best) : 0.0, nsPer(0),
                    static_cast<double>(timed)});
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
