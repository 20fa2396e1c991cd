// Synthetic labelled RGB-D frames (the benchmark and test workloads; not part
// of the reference).  Deterministic for a given seed:
//   * labels: Voronoi partition of `sites` uniform sites, site k -> label
//     k mod (classes + 1), label 0 = void;
//   * colour: per-class mean in U{0..255}^3 drawn from `class_seed`, each
//     pixel mean + 20 * N(0,1) per channel, rounded and clipped to [0, 255];
//   * depth: per Voronoi cell a tilted plane d0 + gx*x + gy*y with d0 in
//     U[500, 5000] mm and gx, gy in U[-2, 2] mm/px, clipped to [500, 5000],
//     then a fraction `zero_frac` of pixels set to 0 (missing).
#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

#include "../../include/curfil_hip.h"

namespace {

double normal01(std::mt19937& g) {
    // Box-Muller on two uniforms in (0, 1]
    const double u1 = (static_cast<double>(g()) + 1.0) / 4294967296.0;
    const double u2 = (static_cast<double>(g()) + 1.0) / 4294967296.0;
    return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
}

double uniform(std::mt19937& g, double lo, double hi) {
    return lo + (hi - lo) * (static_cast<double>(g()) / 4294967296.0);
}

}  // namespace

extern "C" int cf_synth_rgbd(uint32_t seed, uint32_t class_seed, int width, int height, int sites, int classes,
        double zero_frac, uint8_t* rgb, uint16_t* depth, uint8_t* labels) {
    if (width <= 0 || height <= 0 || sites <= 0 || classes <= 0 || classes > 254) return CF_ERR_ARG;
    std::mt19937 cg(class_seed);
    std::vector<double> mu(static_cast<size_t>(classes + 1) * 3);
    for (auto& m : mu) m = static_cast<double>(cg() % 256u);
    std::mt19937 g(seed);
    std::vector<double> sx(sites), sy(sites), d0(sites), gx(sites), gy(sites);
    for (int k = 0; k < sites; k++) {
        sx[k] = uniform(g, 0, width);
        sy[k] = uniform(g, 0, height);
        d0[k] = uniform(g, 500, 5000);
        gx[k] = uniform(g, -2, 2);
        gy[k] = uniform(g, -2, 2);
    }
    for (int y = 0; y < height; y++) {
        for (int x = 0; x < width; x++) {
            int best = 0;
            double bd = 1e300;
            for (int k = 0; k < sites; k++) {
                const double dx = x - sx[k], dy = y - sy[k];
                const double d = dx * dx + dy * dy;
                if (d < bd) {
                    bd = d;
                    best = k;
                }
            }
            const size_t p = static_cast<size_t>(y) * width + x;
            const int label = best % (classes + 1);
            labels[p] = static_cast<uint8_t>(label);
            for (int c = 0; c < 3; c++) {
                double v = std::nearbyint(mu[label * 3 + c] + 20.0 * normal01(g));
                v = v < 0 ? 0 : (v > 255 ? 255 : v);
                rgb[p * 3 + c] = static_cast<uint8_t>(v);
            }
            double dd = d0[best] + gx[best] * (x - sx[best]) + gy[best] * (y - sy[best]);
            dd = dd < 500 ? 500 : (dd > 5000 ? 5000 : dd);
            uint16_t dm = static_cast<uint16_t>(std::lround(dd));
            if (uniform(g, 0, 1) < zero_frac) dm = 0;
            depth[p] = dm;
        }
    }
    return CF_OK;
}
