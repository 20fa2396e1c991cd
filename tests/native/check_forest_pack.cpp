// Checks the block packing of forests (curfil_amd/csrc/forest_pack.cpp) on the host: for complete trees
// of 1..14 split levels, random ragged trees to depth 24, one-leaf trees and a forest of all of them,
// every walk through the packed blocks -- stepped the way the prediction kernel steps (heap positions,
// leaf and child-block masks, popcount numbering) -- stays inside the blocks and leaves, visits the
// records of the nodes the plain node walk visits (threshold and feature bytes), and reaches the leaf
// whose posterior row the plain walk reaches.  Decisions come from a hash of (walk, node), so each walk
// takes its own path.  Prints "ok" and the number of walks, or the first failure.
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "../../curfil_amd/csrc/forest_pack.h"

using namespace cf;

namespace {

struct Tree {
    std::vector<cf_tree_node> nodes;
    std::vector<double> post;
};

const int C = 5;

uint64_t mix(uint64_t x) {
    x ^= x >> 33;
    x *= 0xff51afd7ed558ccdULL;
    x ^= x >> 33;
    x *= 0xc4ceb9fe1a85ec53ULL;
    return x ^ (x >> 33);
}

// node i's threshold names it (exact in float below 2^24); feature bytes from a hash
void fillNode(cf_tree_node& nd, int i, uint64_t salt) {
    nd.threshold = static_cast<float>(i);
    const uint64_t h = mix(salt * 1000003u + static_cast<uint64_t>(i));
    for (int k = 0; k < 11; k++) nd.feature[k] = static_cast<int8_t>((h >> (5 * k)) & 0xff);
}

// creation order as the trainer makes it: per level, children pairs in frontier order
Tree makeTree(std::mt19937_64& rng, int maxDepth, double leafProb, uint64_t salt) {
    Tree t;
    std::vector<int> frontier{0};
    t.nodes.push_back(cf_tree_node{});
    for (int d = 0; !frontier.empty(); d++) {
        std::vector<int> next;
        for (int i : frontier) {
            const bool leaf = d >= maxDepth || (d > 0 && std::uniform_real_distribution<double>(0, 1)(rng) < leafProb);
            cf_tree_node& nd = t.nodes[i];
            nd.is_leaf = leaf ? 1 : 0;
            nd.left = nd.right = -1;
            fillNode(nd, i, salt);
            if (leaf) continue;
            const int l = static_cast<int>(t.nodes.size());
            t.nodes.push_back(cf_tree_node{});
            t.nodes.push_back(cf_tree_node{});
            t.nodes[i].left = l;
            t.nodes[i].right = l + 1;
            next.push_back(l);
            next.push_back(l + 1);
        }
        frontier.swap(next);
    }
    t.post.resize(t.nodes.size() * C);
    for (size_t i = 0; i < t.nodes.size(); i++) {
        for (int c = 0; c < C; c++) t.post[i * C + c] = static_cast<double>(salt) * 1e7 + static_cast<double>(i) + 0.125 * c;
    }
    return t;
}

int popc(uint32_t v) { return __builtin_popcount(v); }

bool fail(const std::string& m) {
    std::printf("FAIL %s\n", m.c_str());
    return false;
}

bool checkForest(const std::vector<Tree>& forest, const std::vector<uint64_t>& salts, int walksPerTree, long& walks) {
    std::vector<const cf_tree_node*> trees;
    std::vector<int> sizes;
    std::vector<const double*> post;
    for (const Tree& t : forest) {
        trees.push_back(t.nodes.data());
        sizes.push_back(static_cast<int>(t.nodes.size()));
        post.push_back(t.post.data());
    }
    PackedForest pf;
    std::string err;
    if (packForest(trees, sizes, post, C, pf, err) != CF_OK) return fail("pack: " + err);
    const size_t S = static_cast<size_t>(rowStride(C));
    if (pf.post.size() % S != 0) return fail("posterior rows not whole");
    const size_t nb = pf.blocks.size() / kBlockWords, nl = pf.post.size() / S;
    if (pf.roots.size() != forest.size()) return fail("roots");
    // every leaf row once, every split record once
    size_t splits = 0, leafNodes = 0;
    for (const Tree& t : forest) {
        for (const auto& nd : t.nodes) (nd.is_leaf ? leafNodes : splits)++;
    }
    if (nl != leafNodes) return fail("leaf rows " + std::to_string(nl) + " vs " + std::to_string(leafNodes));
    for (size_t ti = 0; ti < forest.size(); ti++) {
        const Tree& t = forest[ti];
        for (int w = 0; w < walksPerTree; w++) {
            // plain walk
            std::vector<int> path;
            int i = 0;
            while (!t.nodes[i].is_leaf) {
                path.push_back(i);
                const bool right = mix((static_cast<uint64_t>(w) << 32) ^ static_cast<uint64_t>(i) ^ (salts[ti] << 48)) & 1;
                i = right ? t.nodes[i].right : t.nodes[i].left;
            }
            // packed walk, stepped as the kernel steps
            const int root = pf.roots[ti];
            size_t leaf;
            size_t steps = 0;
            if (root < 0) {
                leaf = static_cast<size_t>(-root - 1);
            } else {
                size_t blk = static_cast<size_t>(root);
                if (blk >= nb) return fail("root block out of range");
                const uint32_t* b = &pf.blocks[blk * kBlockWords];
                uint32_t tx = b[28], ty = b[29], tz = b[30], tw = b[31];
                int p = 0;
                for (;;) {
                    if (p >= static_cast<int>(tw >> 16) || p > 6) return fail("position past the block's nodes");
                    const uint32_t* rec = b + 4 * p;
                    float thr;
                    std::memcpy(&thr, &rec[0], 4);
                    const int id = static_cast<int>(thr);
                    if (steps >= path.size() || id != path[steps]) return fail("visited a node off the path");
                    const cf_tree_node& nd = t.nodes[id];
                    const uint32_t y = rec[1], z = rec[2], ww = rec[3];
                    auto sb = [](uint32_t v, int k) { return static_cast<int8_t>((v >> (8 * k)) & 0xff); };
                    const int8_t* f = nd.feature;
                    if (sb(y, 0) != f[FC_O1X] || sb(y, 1) != f[FC_O1Y] || sb(y, 2) != f[FC_O2X] || sb(y, 3) != f[FC_O2Y] ||
                            sb(z, 0) != f[FC_R1X] || sb(z, 1) != f[FC_R1Y] || sb(z, 2) != f[FC_R2X] ||
                            sb(z, 3) != f[FC_R2Y] || sb(ww, 0) != f[FC_TYPE] || sb(ww, 1) != f[FC_C1] ||
                            sb(ww, 2) != f[FC_C2]) {
                        return fail("feature bytes");
                    }
                    steps++;
                    const bool right = mix((static_cast<uint64_t>(w) << 32) ^ static_cast<uint64_t>(id) ^ (salts[ti] << 48)) & 1;
                    const int q = 2 * p + (right ? 2 : 1);
                    const uint32_t below = (1u << q) - 1u;
                    if ((tz >> q) & 1u) {
                        leaf = ty + popc(tz & below);
                        break;
                    }
                    if (q < static_cast<int>(tw >> 16)) {
                        p = q;
                        continue;
                    }
                    if (!((tw >> q) & 1u)) return fail("exit that is neither a leaf nor a block");
                    blk = tx + popc(tw & 0xffffu & below);
                    if (blk >= nb) return fail("child block out of range");
                    b = &pf.blocks[blk * kBlockWords];
                    tx = b[28], ty = b[29], tz = b[30], tw = b[31];
                    p = 0;
                }
            }
            if (steps != path.size()) return fail("walk length");
            if (leaf >= nl) return fail("leaf out of range");
            for (int c = 0; c < C; c++) {
                if (pf.post[leaf * S + c] != t.post[static_cast<size_t>(i) * C + c]) return fail("leaf row");
            }
            for (size_t c = C; c < S; c++) {
                if (pf.post[leaf * S + c] != 0.0) return fail("row padding");
            }
            walks++;
        }
    }
    return true;
}

}  // namespace

int main() {
    std::mt19937_64 rng(20240601);
    long walks = 0;
    std::vector<Tree> all;
    std::vector<uint64_t> allSalts;
    uint64_t salt = 1;
    // complete trees: every top-block height
    for (int d = 1; d <= 14; d++) {
        Tree t = makeTree(rng, d, 0.0, salt);
        if (!checkForest({t}, {salt}, 400, walks)) return 1;
        all.push_back(t);
        allSalts.push_back(salt++);
    }
    // ragged trees
    for (int k = 0; k < 60; k++) {
        const double leafProb = 0.02 + 0.3 * (k % 7) / 6.0;
        Tree t = makeTree(rng, 4 + k % 21, leafProb, salt);
        if (!checkForest({t}, {salt}, 300, walks)) return 1;
        all.push_back(t);
        allSalts.push_back(salt++);
    }
    // one-leaf tree
    {
        Tree t = makeTree(rng, 0, 0.0, salt);
        if (!checkForest({t}, {salt}, 3, walks)) return 1;
        all.push_back(t);
        allSalts.push_back(salt++);
    }
    // all of them as one forest: block and leaf numbering across trees
    if (!checkForest(all, allSalts, 200, walks)) return 1;
    // an error case: a child before its parent
    {
        Tree t = makeTree(rng, 3, 0.0, salt);
        t.nodes[2].left = 1;
        PackedForest pf;
        std::string err;
        if (packForest({t.nodes.data()}, {static_cast<int>(t.nodes.size())}, {t.post.data()}, C, pf, err) != CF_ERR_ARG) {
            return fail("malformed tree accepted"), 1;
        }
    }
    std::printf("ok %ld\n", walks);
    return 0;
}
