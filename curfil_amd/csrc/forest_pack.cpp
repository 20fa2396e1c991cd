// Block packing of a forest's split nodes (layout in forest_pack.h).  A walk reads one 128-byte line per
// three levels: for a large forest the deep levels are in no cache, and one record per node would cost
// a line per level there.  A tree's top block takes 1 + (D - 1) mod 3 levels for D split levels, so that
// a complete tree ends with full blocks.
#include "forest_pack.h"

#include <algorithm>
#include <cstring>

namespace cf {

namespace {

uint32_t byteOf(const cf_tree_node& nd, int k) { return static_cast<uint32_t>(static_cast<uint8_t>(nd.feature[k])); }

}  // namespace

int packForest(const std::vector<const cf_tree_node*>& trees, const std::vector<int>& sizes,
        const std::vector<const double*>& post, int C, PackedForest& out, std::string& error) {
    out = PackedForest{};
    const size_t kMaxIndex = static_cast<size_t>(INT32_MAX);
    if (C <= 0 || trees.size() != sizes.size() || trees.size() != post.size()) {
        error = "inconsistent forest";
        return CF_ERR_ARG;
    }
    for (size_t t = 0; t < trees.size(); t++) {
        const cf_tree_node* nodes = trees[t];
        const int n = sizes[t];
        const double* tpost = post[t];
        const size_t stride = static_cast<size_t>(rowStride(C));
        auto leaves = [&] { return out.post.size() / stride; };
        auto addLeaf = [&](int i) {
            const size_t leaf = leaves();
            out.post.insert(out.post.end(), tpost + static_cast<size_t>(i) * C, tpost + static_cast<size_t>(i + 1) * C);
            out.post.resize((leaf + 1) * stride, 0.0);
            return static_cast<uint32_t>(leaf);
        };
        if (n <= 0) {
            out.roots.push_back(-1);
            continue;
        }
        // split depth of every node reachable from the root (children follow their parents)
        std::vector<int> depth(n, -1);
        depth[0] = 0;
        int maxSplitDepth = 0;
        for (int i = 0; i < n; i++) {
            const cf_tree_node& nd = nodes[i];
            if (nd.is_leaf) continue;
            if (nd.left <= i || nd.right <= i || nd.left >= n || nd.right >= n) {
                error = "child indices must follow their parent";
                return CF_ERR_ARG;
            }
            if (depth[i] < 0) continue;
            maxSplitDepth = std::max(maxSplitDepth, depth[i]);
            depth[nd.left] = depth[i] + 1;
            depth[nd.right] = depth[i] + 1;
        }
        if (nodes[0].is_leaf) {
            if (leaves() >= kMaxIndex) {
                error = "forest too large";
                return CF_ERR_UNSUPPORTED;
            }
            out.roots.push_back(-static_cast<int32_t>(addLeaf(0)) - 1);
            continue;
        }
        // breadth-first over blocks; a queued block's index is reserved when it is queued
        struct Pending {
            size_t block;
            int node, levels;
        };
        std::vector<Pending> queue;
        size_t numBlocks = out.blocks.size() / kBlockWords;
        if (numBlocks >= kMaxIndex) {
            error = "forest too large";
            return CF_ERR_UNSUPPORTED;
        }
        out.roots.push_back(static_cast<int32_t>(numBlocks));
        queue.push_back({numBlocks, 0, 1 + maxSplitDepth % 3});
        numBlocks++;
        out.blocks.resize(numBlocks * kBlockWords, 0u);
        for (size_t qi = 0; qi < queue.size(); qi++) {
            const Pending pb = queue[qi];
            const int firstExit = (1 << pb.levels) - 1;
            int at[15];   // node at each heap position, -1 where none
            std::fill(at, at + 15, -1);
            at[0] = pb.node;
            for (int p = 0; p < firstExit; p++) {
                if (at[p] < 0 || nodes[at[p]].is_leaf) continue;
                at[2 * p + 1] = nodes[at[p]].left;
                at[2 * p + 2] = nodes[at[p]].right;
            }
            uint32_t leafMask = 0, blockMask = 0;
            for (int q = 1; q <= 2 * firstExit; q++) {
                if (at[q] < 0) continue;
                if (nodes[at[q]].is_leaf) {
                    leafMask |= 1u << q;
                } else if (q >= firstExit) {
                    blockMask |= 1u << q;
                }
            }
            uint32_t* b = &out.blocks[pb.block * kBlockWords];
            for (int p = 0; p < firstExit; p++) {
                if (at[p] < 0 || nodes[at[p]].is_leaf) continue;
                const cf_tree_node& nd = nodes[at[p]];
                uint32_t thr;
                std::memcpy(&thr, &nd.threshold, sizeof(thr));
                b[4 * p + 0] = thr;
                b[4 * p + 1] = byteOf(nd, FC_O1X) | byteOf(nd, FC_O1Y) << 8 | byteOf(nd, FC_O2X) << 16 |
                        byteOf(nd, FC_O2Y) << 24;
                b[4 * p + 2] = byteOf(nd, FC_R1X) | byteOf(nd, FC_R1Y) << 8 | byteOf(nd, FC_R2X) << 16 |
                        byteOf(nd, FC_R2Y) << 24;
                b[4 * p + 3] = byteOf(nd, FC_TYPE) | byteOf(nd, FC_C1) << 8 | byteOf(nd, FC_C2) << 16;
            }
            const size_t firstChild = numBlocks;
            uint32_t firstLeaf = 0;
            bool anyLeaf = false;
            for (int q = 1; q <= 2 * firstExit; q++) {
                if ((leafMask >> q) & 1u) {
                    if (leaves() >= kMaxIndex) {
                        error = "forest too large";
                        return CF_ERR_UNSUPPORTED;
                    }
                    const uint32_t leaf = addLeaf(at[q]);
                    if (!anyLeaf) firstLeaf = leaf;
                    anyLeaf = true;
                } else if ((blockMask >> q) & 1u) {
                    if (numBlocks >= kMaxIndex) {
                        error = "forest too large";
                        return CF_ERR_UNSUPPORTED;
                    }
                    queue.push_back({numBlocks++, at[q], 3});
                }
            }
            b[28] = static_cast<uint32_t>(firstChild);
            b[29] = firstLeaf;
            b[30] = leafMask;
            b[31] = blockMask | static_cast<uint32_t>(firstExit) << 16;
            out.blocks.resize(numBlocks * kBlockWords, 0u);   // b is not used past this point
        }
    }
    return CF_OK;
}

}  // namespace cf
